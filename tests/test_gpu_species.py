"""The setup row on the GPU through the C ABI against the oracle (oracle/species.py):

* qbh_species_twnl (NonLocalPotential::update_twnl) for the Si HSCV LDA species (s, p) and
  a synthetic p+d+f species, on Γ and complex bases, with and without a vector potential;
* qbh_local_tables (cell_moved + atoms_moved: v_ion(G), rho_ps(G)) on the density basis of
  Si8 and of two species;
* Hψ with the projectors built from the pseudopotential on the GPU equals the oracle's Hψ
  with the oracle's tables.

Bar: 1e-12 relative (max-norm for the tables, 2-norm for Hψ).
"""
import os

import numpy as np
import pytest

from helpers import (from_dev, make_gpu_basis, oracle_for, real_to_dev, rel_err,
                     synthetic_spdf, to_dev)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SI_XML = os.path.join(HERE, "golden", "Si_HSCV_LDA-1.0.xml")


@pytest.fixture(scope="module")
def si_pair():
    from oracle import species as S
    from qball_amd import Species
    sp = Species(SI_XML, rcps=1.5)
    yield sp, S.Species(S.read_species_xml(SI_XML), 1.5)
    sp.close()


def _twnl_gpu(ctx, basis, sp, A=None):
    import torch
    out = torch.full((sp.nproj * basis.ngw,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sp.twnl(basis, out.data_ptr(), A)
    ctx.sync()
    return out.cpu().numpy().reshape(sp.nproj, basis.ngw)


@pytest.mark.parametrize("complex_basis,A", [(False, None), (True, None),
                                             (True, (0.03, -0.02, 0.05))])
def test_twnl_si_matches_oracle(gpu_ctx, si_pair, complex_basis, A):
    from oracle import species as S
    from qball_amd.workload import make_system
    sp, ora = si_pair
    s = make_system(ncell=1, ecut_ry=25.0, nbands=4, seed=0, complex_basis=complex_basis)
    basis = make_gpu_basis(gpu_ctx, s)
    got = _twnl_gpu(gpu_ctx, basis, sp, A)
    ob, _, _ = oracle_for(s)
    kpgx = ob.kpgx - (0.0 if A is None else np.asarray(A)[None, :])
    ref, ls, wt = S.twnl_table(ora, kpgx)
    basis.close()
    assert got.shape == ref.shape and np.isfinite(got).all()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert ls.tolist() == sp.lproj.tolist()


def test_twnl_spdf_matches_oracle(gpu_ctx):
    """p, d and f projectors (rows in the reference's order) at a vector potential."""
    from oracle import species as S
    from qball_amd import Species
    from qball_amd.workload import make_system
    d = synthetic_spdf()
    ora = S.Species(d, 1.0)
    sp = Species(rcps=1.0, zval=d["zval"], lmax=d["lmax"], llocal=d["llocal"],
                 deltar=d["deltar"], vps=np.stack(d["vps"]), phi=np.stack(d["phi"]))
    assert sp.lproj.tolist() == ora.projector_ls() and sp.nproj == 15
    s = make_system(ncell=1, ecut_ry=25.0, nbands=4, seed=0, complex_basis=True)
    basis = make_gpu_basis(gpu_ctx, s)
    A = (0.02, 0.01, -0.03)
    got = _twnl_gpu(gpu_ctx, basis, sp, A)
    ob, _, _ = oracle_for(s)
    ref, _, _ = S.twnl_table(ora, ob.kpgx - np.asarray(A)[None, :])
    basis.close()
    sp.close()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_local_tables_match_oracle(gpu_ctx, si_pair):
    """v_ion(G), rho_ps(G) on the Si8 density basis: one species, then two (the synthetic
    one on half of the atoms)."""
    import torch
    from oracle import species as S
    from qball_amd import Basis, Species, capi, local_tables
    from qball_amd.workload import make_system, plane_waves
    sp, ora = si_pair
    s = make_system(ncell=1, ecut_ry=25.0, nbands=4, seed=0, complex_basis=True)
    vr = capi.rods_build(s.a, 4.0 * s.ecut, False)
    vb = Basis(gpu_ctx, False, vr["rod_h"], vr["rod_k"], vr["rod_lmin"], vr["rod_size"], s.np, s.b)
    _, gx = plane_waves(vr, s.b)
    d = synthetic_spdf()
    sp2 = Species(rcps=1.0, zval=d["zval"], lmax=d["lmax"], llocal=d["llocal"],
                  deltar=d["deltar"], vps=np.stack(d["vps"]), phi=np.stack(d["phi"]))
    ora2 = S.Species(d, 1.0)
    cases = [([sp], [ora], [s.tau]), ([sp, sp2], [ora, ora2], [s.tau[:5], s.tau[5:]])]
    for lib_sps, ora_sps, taus in cases:
        na = [len(t) for t in taus]
        tau_d = real_to_dev(np.concatenate(taus).ravel())
        vion = torch.full((2 * vb.ngw,), float("nan"), dtype=torch.float64, device="cuda")
        rhops = torch.full_like(vion, float("nan"))
        torch.cuda.synchronize()
        local_tables(vb, lib_sps, na, tau_d.data_ptr(), s.omega, vion.data_ptr(), rhops.data_ptr())
        gpu_ctx.sync()
        gv = vion.cpu().numpy().view(np.complex128)
        gr = rhops.cpu().numpy().view(np.complex128)
        rv, rr = S.local_tables(ora_sps, taus, gx, s.omega)
        assert np.max(np.abs(gv - rv)) <= 1e-12 * np.max(np.abs(rv))
        assert np.max(np.abs(gr - rr)) <= 1e-12 * np.max(np.abs(rr))
    sp2.close()
    vb.close()


def test_hpsi_with_pseudopotential_projectors(gpu_ctx, si_pair):
    """Hψ (all terms) with twnl, lproj and wsg from the Si species on the GPU equals the
    oracle's Hψ with the oracle's tables."""
    import torch
    from oracle import hpsi as oh
    from oracle import species as S
    from qball_amd import Hamiltonian
    from qball_amd.workload import make_system
    sp, ora = si_pair
    s = make_system(ncell=1, ecut_ry=25.0, nbands=12, seed=7)
    basis = make_gpu_basis(gpu_ctx, s)
    tw = _twnl_gpu(gpu_ctx, basis, sp)
    ham = Hamiltonian(gpu_ctx, basis)
    vd = real_to_dev(s.v)
    ham.set_vloc(vd.data_ptr())
    ham.set_projectors([s.na], [sp.nproj], sp.lproj, sp.wt, tw.ravel(), s.tau.ravel(), s.omega)
    cd = to_dev(s.psi)
    hd = torch.zeros_like(cd)
    enl = ham.hpsi(cd.data_ptr(), basis.ngw, s.nst, s.occ, hd.data_ptr(), basis.ngw,
                   want_enl=True)
    gpu_ctx.sync()
    got = from_dev(hd, basis.ngw, s.nst)
    ob, maps, _ = oracle_for(s)
    rows, ls, wt = S.twnl_table(ora, ob.kpgx)
    ref, enl_ref = oh.hpsi(maps, s.psi, s.v, [oh.Projectors(ls, wt, rows, s.tau)], s.occ, s.omega)
    ham.close()
    basis.close()
    assert rel_err(got, ref) <= 1e-12
    assert abs(enl - enl_ref) <= 1e-10 * abs(enl_ref)
