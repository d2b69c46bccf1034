"""The library's host-side setup (qbh_species_*, qbh_ewald_esr in libqbh.so) against the
oracle (oracle/species.py) on the CPU: the species read from the Si HSCV LDA file, its
spline values Vloc(G) and the KB projector transforms with their derivatives, the KB
weights, eself, and the Ewald esr of diamond cells (distinct pairs, folded into the
Wigner-Seitz cell, image count from the reference's rule).  No GPU is used.

Bar: 1e-12 relative (the radial transforms are direct sums in the library and DST/DCT in
the oracle, equal up to rounding).
"""
import os

import numpy as np
import pytest

from oracle import species as S

HERE = os.path.dirname(os.path.abspath(__file__))
SI_XML = os.path.join(HERE, "golden", "Si_HSCV_LDA-1.0.xml")


@pytest.fixture(scope="module")
def pair():
    from qball_amd import Species
    lib_sp = Species(SI_XML, rcps=1.5)
    yield lib_sp, S.Species(S.read_species_xml(SI_XML), 1.5)
    lib_sp.close()


def test_species_scalars_and_projector_list(pair):
    lib_sp, ora = pair
    assert lib_sp.zval == 4.0 and lib_sp.rcps == 1.5
    assert abs(lib_sp.eself - ora.eself()) <= 1e-15 * ora.eself()
    assert lib_sp.lproj.tolist() == ora.projector_ls() == [0, 1, 1, 1]
    assert np.allclose(lib_sp.wt, [ora.wsg[l] for l in ora.projector_ls()], rtol=1e-12, atol=0)


@pytest.mark.parametrize("which", [-1, 0, 1])
def test_spline_values_and_derivatives(pair, which):
    lib_sp, ora = pair
    g = np.concatenate([np.linspace(0.0, 12.0, 1501), [ora.gspl[-1] * 1.01]])
    v, dv = lib_sp.eval(which, g)
    if which < 0:
        rv = ora.vlocg(g)
        _, rdv = ora.local_spline.derivative(g)
        rdv = np.where(g > ora.gspl[-1], 0.0, rdv)
    else:
        rv, rdv = ora.dvnlg(which, g)
    assert np.max(np.abs(v - rv)) <= 1e-12 * np.max(np.abs(rv))
    assert np.max(np.abs(dv - rdv)) <= 1e-11 * np.max(np.abs(rdv))
    assert v[-1] == 0.0  # past the last spline knot


def test_local_channel_has_no_projector(pair):
    lib_sp, _ = pair
    v, dv = lib_sp.eval(2, np.array([0.5, 1.0]))
    assert np.all(v == 0.0) and np.all(dv == 0.0)


@pytest.mark.parametrize("ncell,rcps", [(1, 1.5), (2, 1.0), (1, 3.0)])
def test_ewald_esr_matches_oracle(ncell, rcps):
    from qball_amd import Species, ewald_esr
    from qball_amd.workload import make_system
    s = make_system(ncell=ncell, ecut_ry=10.0, nbands=4, seed=0, generate_psi=False)
    sp = Species(SI_XML, rcps=rcps)
    e = ewald_esr(s.a, [sp], [s.na], s.tau)
    sp.close()
    ref = S.esr(s.a, [4.0], [rcps], [s.tau])
    assert ref > 0.0
    assert abs(e - ref) <= 1e-12 * abs(ref)


def test_species_from_arrays_equals_species_from_file(pair):
    from qball_amd import Species
    lib_sp, _ = pair
    d = S.read_species_xml(SI_XML)
    phi = np.stack([f if f is not None else np.zeros_like(d["vps"][0]) for f in d["phi"]])
    sp = Species(rcps=1.5, zval=d["zval"], lmax=d["lmax"], llocal=d["llocal"],
                 deltar=d["deltar"], vps=np.stack(d["vps"]), phi=phi)
    g = np.linspace(0.0, 10.0, 101)
    for which in (-1, 0, 1):
        assert np.array_equal(sp.eval(which, g)[0], lib_sp.eval(which, g)[0])
    sp.close()


def test_auto_rcps_matches_oracle():
    from oracle import species as S
    from qball_amd import auto_rcps
    for a, zv, ecut in [(np.diag([41.04] * 3), 4.0, 12.5), (np.diag([38.25] * 3), 3.0, 15.0),
                        (np.diag([5.0] * 3), 4.0, 40.0),
                        (np.array([[0.0, 2.0, 2.0], [2.0, 0.0, 2.0], [2.0, 2.0, 0.0]]), 4.0, 40.0)]:
        assert auto_rcps(zv, a, ecut) == S.auto_rcps(zv, a, ecut)
