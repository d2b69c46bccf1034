"""Oracle SCF pieces pinned by their defining optimality conditions and by the SCF fixed
point."""
import math

import numpy as np
import pytest

from oracle import hpsi as oh
from oracle.basis import Basis, FFTMaps, UnitCell
from oracle.scf import AndersonMixer, PSDA, kinetic_energy, mixing_weights, prec_diag, scf


def _metric(a, b, real):
    s = (a.conj() * b).real.sum(axis=0)
    return 2.0 * s - (a[0].conj() * b[0]).real if real else s


@pytest.mark.parametrize("real", [True, False])
def test_psda_residual_is_metric_orthogonal_and_theta_minimises(real):
    cell = UnitCell(np.array([[6.0, 0.2, 0.0], [0.0, 5.5, 0.4], [0.3, 0.0, 6.5]]))
    b = Basis(cell, 3.0, force_complex=not real)
    rng = np.random.default_rng(5)
    nst = 4
    c = rng.standard_normal((b.size, nst)) + 1j * rng.standard_normal((b.size, nst))
    if real:
        c[0] = c[0].real
    c = oh.gram(c, real)
    H = rng.standard_normal((b.size, b.size))
    H = H + H.T
    hc = H @ c.real + 1j * (H @ c.imag)
    if real:
        hc[0] = hc[0].real
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    p = PSDA(b, 1.5)
    new, theta = p.update(c, hc, occ)
    # residual HC - C A is orthogonal to C in the basis metric (pins A, incl. the G=0 term)
    K = prec_diag(b, 1.5)
    r = -p.dc_last / K[:, None]
    for i in range(nst):
        for j in range(nst):
            prod = _metric(c[:, i:i + 1], r[:, j:j + 1], real)[0]
            assert abs(prod) <= 1e-10 * np.max(np.abs(r))
    assert theta == 0.0 and np.allclose(new, c + p.dc_last)
    # second step: theta minimises the occupation-weighted metric norm of f + t (f - f_last)
    c2 = oh.gram(new, real)
    hc2 = H @ c2.real + 1j * (H @ c2.imag)
    if real:
        hc2[0] = hc2[0].real
    f_last = p.dc_last.copy()
    _, theta2 = p.update(c2, hc2, occ)
    f = p.dc_last
    ts = np.linspace(-1.0, 2.0, 30001)
    norms = [float((occ * _metric(f + t * (f - f_last), f + t * (f - f_last), real)).sum()) for t in ts]
    assert abs(ts[int(np.argmin(norms))] - theta2) <= 2e-4


def test_preconditioner_floor():
    cell = UnitCell(np.diag([7.0, 7.0, 7.0]))
    b = Basis(cell, 4.0)
    K = prec_diag(b, 2.0)
    e = 0.5 * b.kpg2
    assert np.all(K[e < 2.0] == 0.25) and np.allclose(K[e >= 2.0], 0.5 / e[e >= 2.0])


def test_anderson_mixer_least_squares_and_ring():
    rng = np.random.default_rng(3)
    m = AndersonMixer(50, 3)
    xs, fs = [], []
    for it in range(6):
        x, f = rng.standard_normal(50), 0.1 * rng.standard_normal(50)
        xs.append(x)
        fs.append(f)
        xbar, fbar, theta = m.update(x, f)
        n = min(it, 3)
        assert len(theta) == n
        if n == 0:
            assert np.allclose(xbar, x) and np.allclose(fbar, f)
            continue
        D = np.stack([fs[-1] - fs[-2 - i] for i in range(n)], axis=1)
        t_ls = np.linalg.lstsq(D, fs[-1], rcond=None)[0]
        if np.sum(t_ls ** 2) <= 1.0:
            assert np.allclose(theta, t_ls, rtol=1e-6, atol=1e-9)
        assert np.allclose(fbar, fs[-1] - D @ theta)
        Dx = np.stack([xs[-1] - xs[-2 - i] for i in range(n)], axis=1)
        assert np.allclose(xbar, xs[-1] - Dx @ theta)
        assert np.sum(theta ** 2) <= 1.0 + 1e-12


def test_scf_reaches_a_self_consistent_fixed_point():
    """A tiny Γ system: energies settle and the final orbitals are eigenvectors of the
    Hamiltonian of their own density."""
    from oracle.vhxc import update_vhxc
    cell = UnitCell(np.diag([6.0, 6.0, 6.0]))
    wb = Basis(cell, 4.0)
    vb = Basis(cell, 16.0)
    n = [max(a, b) for a, b in zip(wb.np, vb.np)]
    maps, vmaps = FFTMaps(wb, *n), FFTMaps(vb, *n)
    rng = np.random.default_rng(2)
    c = rng.standard_normal((wb.size, 2)) + 1j * rng.standard_normal((wb.size, 2))
    c[0] = c[0].real
    c = oh.gram(c * (1.0 / (1.0 + wb.g2))[:, None], True)
    vion = -0.8 * np.exp(-0.2 * vb.g2).astype(np.complex128)
    rhops = (0.0 * vb.g2).astype(np.complex128)
    occ = np.array([2.0, 2.0])
    c, hist = scf(maps, vmaps, [], c, occ, cell.volume, vion, rhops, nitscf=25, nite=3,
                  ecutprec=4.0)
    e = [s[-1]["etotal"] for s in hist]
    assert e[-1] < e[0] - 40.0 and abs(e[-1] - e[-2]) < 1e-4
    rho, _, rhog = oh.update_density(maps, vmaps, c, occ, 1.0, cell.volume)
    v = update_vhxc(vmaps, rho, rhog, vion, rhops, cell.volume)[0]
    hc = oh.hpsi(maps, c, v, [], occ, cell.volume)[0]
    lam = 2.0 * (c.real.T @ hc.real + c.imag.T @ hc.imag) - np.outer(c[0].real, hc[0].real)
    r = hc - c @ lam
    assert np.linalg.norm(r) <= 1e-2 * np.linalg.norm(hc)
    wk, wls = mixing_weights(vb)
    assert wk[0] == 0.0 and wls[0] == 1.0
    assert kinetic_energy(wb, c, occ) > 0.0


def test_mixing_weights_half_points():
    """Kerker weight g2/(g2 + q0^2) is 1/2 where |G| = q0 = 2 pi / rcut and the row weight
    sqrt(g2/(g2 + q1^2)) is 1/sqrt(2) where |G| = q1 = 2 pi / rc1 (BOSampleStepper.cc:425-444);
    a q0 = 1/rcut slip moves both half points."""
    from types import SimpleNamespace
    q0, q1 = 2.0 * math.pi / 10.0, 2.0 * math.pi / 3.0
    vb = SimpleNamespace(g2=np.array([0.0, q0 * q0, q1 * q1, 1e6]))
    wk, wls = mixing_weights(vb, rc_kerker=10.0, rc1=3.0)
    assert abs(wk[1] - 0.5) < 1e-15 and abs(wls[2] - math.sqrt(0.5)) < 1e-15
    assert wk[0] == 0.0 and wls[0] == 1.0 and abs(wk[3] - 1.0) < 1e-5
    wk0, _ = mixing_weights(vb, rc_kerker=0.0)
    assert np.all(wk0 == 1.0)


def test_fermi_occupations_minimise_the_free_energy():
    """Wavefunction::update_occ / SlaterDet::fermi / SlaterDet::entropy: the occupations hold
    nel electrons to the reference's 1e-8, mu agrees with a root of N(mu) = nel, and with
    kT = width / 2 (width in Rydberg) they minimise F = sum occ e - kT S over occupations of
    the same total, S the entropy the reference's ets uses (ets = -S width / 2)."""
    from scipy.optimize import brentq
    from oracle import scf as O
    rng = np.random.default_rng(4)
    eig = np.sort(rng.uniform(-0.6, 0.4, 30))
    nel, width = 33.7, 0.02
    occ, mu = O.update_occ(eig, nel, width)
    assert abs(occ.sum() - nel) <= 1e-8
    mu_root = brentq(lambda m: np.sum(2.0 / (1.0 + np.exp((eig - m) / (0.5 * width)))) - nel,
                     -2.0, 2.0, xtol=1e-15)
    dn_dmu = np.sum(occ * (2.0 - occ) / 2.0) / (0.5 * width)
    assert abs(mu - mu_root) <= 2e-8 / dn_dmu
    kt = 0.5 * width
    F = lambda o: float(np.sum(o * eig)) - kt * O.entropy(o)
    f0 = F(occ)
    wgt = occ * (2.0 - occ)     # perturbations that keep every occupation inside (0, 2)
    for _ in range(20):
        d = 0.05 * rng.standard_normal(eig.size) * wgt
        d -= wgt * d.sum() / wgt.sum()
        o = occ + d
        assert np.all((o >= 0.0) & (o <= 2.0)) and abs(o.sum() - occ.sum()) < 1e-12
        assert F(o) > f0
    # limits: a full or empty state has no entropy, f = 1/2 gives 2 ln 2
    assert O.entropy(np.array([2.0, 0.0])) == 0.0
    assert abs(O.entropy(np.array([1.0])) - 2.0 * np.log(2.0)) <= 1e-15
    # a gapped spectrum at vanishing width fills the lowest nel / 2 states
    occ0, _ = O.update_occ(np.array([-1.0, -0.5, 0.5, 1.0]), 4.0, 1e-6)
    assert np.allclose(occ0, [2.0, 2.0, 0.0, 0.0])
