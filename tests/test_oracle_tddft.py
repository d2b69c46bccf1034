"""Oracle TDDFT adjoint pinned by central finite differences on a tiny cell."""
import math

import numpy as np
import pytest

from oracle import hpsi as oh
from oracle.basis import Basis, FFTMaps, UnitCell
from oracle.tddft import TDSystem, fxc_pz81, hartree_apply, td_forward, td_gradient
from oracle.vhxc import pz81


def test_fxc_is_the_derivative_of_vxc():
    rho = np.concatenate([np.logspace(-5, 1.5, 300), [3.0 / (4 * math.pi) * 1.001,
                                                      3.0 / (4 * math.pi) * 0.999]])
    h = 1e-6
    d = (pz81(rho * (1 + h))[1] - pz81(rho * (1 - h))[1]) / (2 * h * rho)
    assert np.max(np.abs(d - fxc_pz81(rho)) / np.abs(d)) < 1e-7
    assert fxc_pz81(np.array([0.0, -1.0])).tolist() == [0.0, 0.0]


def _tiny(seed=3, nst=2):
    cell = UnitCell(np.array([[5.0, 0.2, 0.0], [0.0, 4.6, 0.3], [0.2, 0.0, 5.3]]))
    wb = Basis(cell, 2.5, force_complex=True)
    vb = Basis(cell, 10.0, force_complex=True)
    n = [max(a, b) for a, b in zip(wb.np, vb.np)]
    maps, vmaps = FFTMaps(wb, *n), FFTMaps(vb, *n)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((wb.size, nst)) + 1j * rng.standard_normal((wb.size, nst))
    c = oh.gram(c * (1.0 / (1.0 + wb.g2))[:, None], False)
    g = wb.kpg
    gi = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    s14, s34 = math.sqrt(1 / (4 * math.pi)), math.sqrt(3 / (4 * math.pi))
    vs, vp = 1.2 * np.exp(-0.4 * g * g), 0.8 * g * np.exp(-0.4 * g * g)
    twnl = np.stack([s14 * vs] + [s34 * wb.kpgx[:, d] * gi * vp for d in range(3)])
    sp = oh.Projectors([0, 1, 1, 1], [1.1, -0.6, -0.6, -0.6], twnl, rng.uniform(0, 4, (2, 3)))
    vion = (rng.standard_normal(vb.size) + 1j * rng.standard_normal(vb.size)) * 0.3 * np.exp(-0.5 * vb.g2)
    rhops = (0.03 * np.exp(-0.5 * vb.g2)).astype(np.complex128)
    sysm = TDSystem(maps, vmaps, [sp], np.array([2.0, 1.0][:nst]), cell.volume, vion, rhops)
    vext = 0.05 * np.cos(np.arange(maps.N) * 0.37)
    return sysm, c, vext, rng


def test_adjoint_matches_central_differences():
    sysm, c, vext, rng = _tiny()
    nsteps, dt = 3, 0.2
    A = rng.uniform(-0.1, 0.1, (nsteps + 1, 3))
    w = np.array([0.3, 1.0, -0.5, 0.7])
    e = np.array([1.0, 0.0, 0.0])
    J, gv, gA = td_gradient(sysm, c, vext, A, w, e, dt, nsteps)
    J2 = td_forward(sysm, c, vext, A, w, e, dt, nsteps)[0]
    assert abs(J - J2) <= 1e-14 * max(1.0, abs(J))
    h = 1e-5
    dv = np.sin(np.arange(sysm.maps.N) * 0.11) * 0.5
    fd = (td_forward(sysm, c, vext + h * dv, A, w, e, dt, nsteps)[0]
          - td_forward(sysm, c, vext - h * dv, A, w, e, dt, nsteps)[0]) / (2 * h)
    assert abs(fd - gv @ dv) <= 1e-6 * abs(fd), (fd, gv @ dv)
    for n, d in [(0, 0), (1, 1), (2, 0), (3, 2)]:
        Ap, Am = A.copy(), A.copy()
        Ap[n, d] += h
        Am[n, d] -= h
        fd = (td_forward(sysm, c, vext, Ap, w, e, dt, nsteps)[0]
              - td_forward(sysm, c, vext, Am, w, e, dt, nsteps)[0]) / (2 * h)
        assert abs(fd - gA[n, d]) <= 1e-6 * max(abs(fd), 1e-8), (n, d, fd, gA[n, d])


def test_hartree_operator_is_symmetric_and_linear():
    sysm, _, _, rng = _tiny()
    a = rng.standard_normal(sysm.maps.N)
    b = rng.standard_normal(sysm.maps.N)
    ka, kb = hartree_apply(sysm.vmaps, a), hartree_apply(sysm.vmaps, b)
    assert abs(a @ kb - b @ ka) <= 1e-12 * abs(a @ kb)
    assert np.max(np.abs(hartree_apply(sysm.vmaps, 2 * a - b) - (2 * ka - kb))) <= 1e-12 * np.max(np.abs(ka))


def test_adjoint_with_retabled_projectors_matches_central_differences():
    """Projector tables re-tabled at |k+G-A| from two species (the Si HSCV s+p species and a
    synthetic p+d+f one): the adjoint's A gradient, which then carries the derivative of the
    tables, against central differences of the forward map."""
    import os
    from helpers import synthetic_spdf
    from oracle import species as S
    sysm, c, vext, rng = _tiny(seed=5)
    here = os.path.dirname(os.path.abspath(__file__))
    si = S.Species(S.read_species_xml(os.path.join(here, "golden", "Si_HSCV_LDA-1.0.xml")), 1.5)
    spdf = S.Species(synthetic_spdf(), 1.0)
    sysm.retable = [(si, rng.uniform(0, 4, (2, 3))), (spdf, rng.uniform(0, 4, (1, 3)))]
    nsteps, dt = 2, 0.2
    A = rng.uniform(-0.3, 0.3, (nsteps + 1, 3))
    w = np.array([0.3, 1.0, -0.5])
    e = np.array([0.6, 0.8, 0.0])
    J, gv, gA = td_gradient(sysm, c, vext, A, w, e, dt, nsteps)
    # the tables do move with A: the same run with them frozen at A = 0 differs
    frozen = TDSystem(sysm.maps, sysm.vmaps, sysm.projectors(np.zeros(3)), sysm.occ,
                      sysm.omega, sysm.vion_g, sysm.rhops_g)
    assert abs(td_forward(frozen, c, vext, A, w, e, dt, nsteps)[0] - J) > 1e-6 * abs(J)
    h = 1e-5
    for n, d in [(0, 0), (1, 1), (2, 2), (1, 0)]:
        Ap, Am = A.copy(), A.copy()
        Ap[n, d] += h
        Am[n, d] -= h
        fd = (td_forward(sysm, c, vext, Ap, w, e, dt, nsteps)[0]
              - td_forward(sysm, c, vext, Am, w, e, dt, nsteps)[0]) / (2 * h)
        assert abs(fd - gA[n, d]) <= 1e-6 * max(abs(fd), 1e-8), (n, d, fd, gA[n, d])
