"""Oracle Hψ, projectors, density, Gram and exponential pinned by identities and closed forms."""
import math

import numpy as np
import pytest

from oracle import fft as offt
from oracle import hpsi as oh
from oracle.basis import Basis, FFTMaps, UnitCell


def setup(real=True, nst=5, seed=0):
    cell = UnitCell(np.array([[6.0, 0.2, 0.0], [0.0, 5.5, 0.4], [0.3, 0.0, 6.5]]))
    b = Basis(cell, 5.0, force_complex=not real)
    maps = FFTMaps(b, *b.np)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((b.size, nst)) + 1j * rng.standard_normal((b.size, nst))
    if real:
        c[0] = c[0].real
    return cell, b, maps, c, rng


def gdot(a, b, real):
    """<a|b> per column: Γ half-sphere storage counts G and -G (2 Re - G=0 term)."""
    if real:
        return 2.0 * np.sum((a.conj() * b).real, axis=0) - (a[0].conj() * b[0]).real
    return np.sum(a.conj() * b, axis=0)


def species(b, rng, na=3):
    g = b.kpg
    gi = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    vs = 1.7 * np.exp(-0.4 * g * g)
    vp = 0.9 * g * np.exp(-0.4 * g * g)
    s14 = math.sqrt(1 / (4 * math.pi))
    s34 = math.sqrt(3 / (4 * math.pi))
    twnl = np.stack([s14 * vs, s34 * b.kpgx[:, 0] * gi * vp, s34 * b.kpgx[:, 1] * gi * vp,
                     s34 * b.kpgx[:, 2] * gi * vp])
    tau = rng.uniform(0, 5, size=(na, 3))
    return oh.Projectors([0, 1, 1, 1], [1.3, -0.7, -0.7, -0.7], twnl, tau)


@pytest.mark.parametrize("real,nst", [(True, 4), (True, 5), (False, 3)])
def test_constant_potential(real, nst):
    _, b, maps, c, _ = setup(real, nst)
    v = np.full(maps.N, 0.37)
    cp = oh.rs_mul_add(maps, c, v, np.zeros_like(c))
    assert np.max(np.abs(cp - 0.37 * c)) <= 1e-14


def test_cosine_potential_closed_form():
    """complex basis, single plane wave G1, v = cos(G0.r): V psi = (psi(G1+G0)+psi(G1-G0))/2."""
    _, b, maps, _, _ = setup(False, 1)
    idx = {tuple(g): i for i, g in enumerate(b.idx)}
    g1 = (1, 0, -1)
    g0 = (0, 1, 1)
    n0, n1, n2 = maps.np0, maps.np1, maps.np2
    z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    v = np.cos(2 * np.pi * (g0[0] * x / n0 + g0[1] * y / n1 + g0[2] * z / n2)).ravel()
    c = np.zeros((b.size, 1), dtype=np.complex128)
    c[idx[g1], 0] = 1.0
    cp = oh.rs_mul_add(maps, c, v, np.zeros_like(c))
    ref = np.zeros(b.size, dtype=np.complex128)
    for s in (+1, -1):
        g = tuple(a + s * d for a, d in zip(g1, g0))
        if g in idx:
            ref[idx[g]] += 0.5
    assert np.max(np.abs(cp[:, 0] - ref)) <= 1e-14


@pytest.mark.parametrize("real", [True, False])
def test_hamiltonian_is_hermitian(real):
    cell, b, maps, c, rng = setup(real, 4)
    sp = species(b, rng)
    v = rng.standard_normal(maps.N)
    occ = np.zeros(4)
    h, _ = oh.hpsi(maps, c, v, [sp], occ, cell.volume)
    m1 = np.array([[gdot(c[:, [i]], h[:, [j]], real)[0] for j in range(4)] for i in range(4)])
    assert np.max(np.abs(m1 - m1.conj().T)) <= 1e-12 * np.max(np.abs(m1))


def test_projector_contraction_real_space():
    """Γ: F = 2 Re(anl^H c) - anl0 c0 equals (1/N) sum_r beta(r) psi(r) (Parseval)."""
    cell, b, maps, c, rng = setup(True, 3)
    sp = species(b, rng)
    anl = oh.anl_block(sp, b.kpgx, 0, sp.na)
    occ = np.array([2.0, 1.0, 0.5])
    F = (2.0 * (anl.real.T @ c.real + anl.imag.T @ c.imag)
         - np.outer(anl[0].real, c[0].real))
    Fr = np.zeros_like(F)
    for p in range(anl.shape[1]):
        beta = offt.backward(maps, anl[:, p]).ravel()
        assert np.max(np.abs(beta.imag)) <= 1e-12 * np.max(np.abs(beta))  # real projector
        for n in range(c.shape[1]):
            psi = offt.backward(maps, c[:, n]).ravel()
            Fr[p, n] = np.sum(beta.real * psi.real) / maps.N
    assert np.max(np.abs(F - Fr)) <= 1e-12 * np.max(np.abs(F))
    # Enl = sum occ wt/Omega F^2, and <psi_n|V_NL psi_n> summed with occ gives the same
    cp = np.zeros_like(c)
    enl = oh.nonlocal_energy(b, [sp], c, occ, cell.volume, cp=cp)
    wt = np.repeat(sp.wt, sp.na)
    assert abs(enl - np.sum(occ[None, :] * (wt / cell.volume)[:, None] * Fr ** 2)) <= 1e-12 * abs(enl)
    assert abs(np.sum(occ * gdot(c, cp, True)) - enl) <= 1e-12 * abs(enl)


def test_projector_blocks_do_not_change_result():
    """na_block (128 atoms in the reference) only reorders sums."""
    cell, b, maps, c, rng = setup(False, 3)
    sp = species(b, rng, na=5)
    occ = np.ones(3)
    cp1, cp2 = np.zeros_like(c), np.zeros_like(c)
    e1 = oh.nonlocal_energy(b, [sp], c, occ, cell.volume, cp=cp1, na_block=2)
    e2 = oh.nonlocal_energy(b, [sp], c, occ, cell.volume, cp=cp2)
    assert abs(e1 - e2) <= 1e-13 * abs(e1) and np.max(np.abs(cp1 - cp2)) <= 1e-13


def species_all_l(b, rng, na=3):
    """One projector per l = 0..3 with smooth radial tables (the force identities below
    need no particular table, only one per l pattern of the phase)."""
    g = b.kpg
    twnl = np.stack([(1.0 + 0.3 * l) * g ** l * np.exp(-0.35 * g * g) for l in range(4)])
    tau = rng.uniform(0, 5, size=(na, 3))
    return oh.Projectors([0, 1, 2, 3], [1.3, -0.7, 0.4, 0.25], twnl, tau)


def _fd_forces(b, sp, c, occ, omega, kpgx, h=1e-5):
    """-dEnl/dtau by central differences of the oracle's Enl."""
    f = np.zeros((sp.na, 3))
    for ia in range(sp.na):
        for j in range(3):
            e = []
            for s in (1, -1):
                tau = sp.tau.copy()
                tau[ia, j] += s * h
                p = oh.Projectors(sp.lproj, sp.wt, sp.twnl, tau)
                e.append(oh.nonlocal_energy(b, [p], c, occ, omega, kpgx=kpgx))
            f[ia, j] = -(e[0] - e[1]) / (2 * h)
    return f


@pytest.mark.parametrize("real,vp", [(True, False), (False, False), (False, True)])
def test_nonlocal_forces_are_minus_enl_gradient(real, vp):
    """NonLocalPotential.cc:2175-2323: fion = -dEnl/dtau for every l pattern, Γ and complex,
    and with a vector potential (phases and derivative factors with k+G-A)."""
    cell, b, maps, c, rng = setup(real, 4, seed=3)
    sp = species_all_l(b, rng)
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    kpgx = b.kpgx - np.array([0.11, -0.07, 0.05]) if vp else b.kpgx
    enl, fion = oh.nonlocal_forces(b, [sp], c, occ, cell.volume, kpgx=kpgx)
    assert enl == oh.nonlocal_energy(b, [sp], c, occ, cell.volume, kpgx=kpgx)
    fd = _fd_forces(b, sp, c, occ, cell.volume, kpgx)
    assert np.max(np.abs(fion[0] - fd)) <= 1e-7 * np.max(np.abs(fd))


def test_nonlocal_forces_blocks_and_species():
    """Atom blocks (na_block) only reorder sums; two species give the forces of each alone."""
    cell, b, maps, c, rng = setup(True, 3, seed=5)
    s1 = species_all_l(b, rng, na=5)
    s2 = species(b, rng, na=2)
    occ = np.array([2.0, 1.5, 0.0])
    e, f = oh.nonlocal_forces(b, [s1, s2], c, occ, cell.volume)
    e1, f1 = oh.nonlocal_forces(b, [s1], c, occ, cell.volume, na_block=2)
    e2, f2 = oh.nonlocal_forces(b, [s2], c, occ, cell.volume)
    assert abs(e - (e1 + e2)) <= 1e-13 * abs(e)
    assert np.max(np.abs(f[0] - f1[0])) <= 1e-13 * np.max(np.abs(f1[0]))
    assert np.max(np.abs(f[1] - f2[0])) <= 1e-13 * np.max(np.abs(f2[0]))


@pytest.mark.parametrize("real,l", [(True, 2), (False, 3)])
def test_nonlocal_force_check_catches_a_wrong_phase(monkeypatch, real, l):
    """A derivative table whose l pattern has its phase off by i (a swapped real/imaginary
    part) fails the gradient check: the identity pins each l pattern, not only their sum."""
    cell, b, maps, c, rng = setup(real, 3, seed=7)
    sp = species_all_l(b, rng, na=2)
    occ = np.array([2.0, 1.0, 1.0])
    fd = _fd_forces(b, sp, c, occ, cell.volume, b.kpgx)
    good = oh.danl_block

    def planted(sp_, kpgx, j, ia0, ia1):
        d = good(sp_, kpgx, j, ia0, ia1)
        nblk = ia1 - ia0
        cols = slice(l * nblk, (l + 1) * nblk)     # projector ipr = l has angular momentum l
        d[:, cols] = 1j * d[:, cols]
        return d

    monkeypatch.setattr(oh, "danl_block", planted)
    _, f = oh.nonlocal_forces(b, [sp], c, occ, cell.volume)
    assert np.max(np.abs(f[0] - fd)) > 1e-3 * np.max(np.abs(fd))


@pytest.mark.parametrize("real", [True, False])
def test_density_integral_and_direct_sum(real):
    cell, b, maps, c, rng = setup(real, 4)
    c = oh.gram(c, real)
    occ = np.array([2.0, 2.0, 1.0, 0.0])
    rho = oh.compute_density(maps, c, occ, 1.0, cell.volume)
    assert np.min(rho) >= 0.0
    nel = np.sum(rho) * cell.volume / maps.N
    assert abs(nel - occ.sum()) <= 1e-12
    # direct evaluation at a few grid points
    n0, n1, n2 = maps.np0, maps.np1, maps.np2
    for (x, y, z) in [(0, 0, 0), (3, 1, 2), (n0 - 1, n1 - 2, n2 - 3)]:
        tot = 0.0
        for n in range(4):
            s = 0.0
            terms = [(g, c[i, n]) for i, g in enumerate(b.idx)]
            if real:
                terms += [(-g, np.conj(c[i, n])) for i, g in enumerate(b.idx) if np.any(g != 0)]
            for g, cv in terms:
                s += cv * np.exp(2j * np.pi * (g[0] * x / n0 + g[1] * y / n1 + g[2] * z / n2))
            tot += occ[n] / cell.volume * abs(s) ** 2
        assert abs(rho[x + n0 * (y + n1 * z)] - tot) <= 1e-12 * max(1.0, tot)


@pytest.mark.parametrize("real", [True, False])
def test_gram_orthonormal(real):
    _, b, maps, c, _ = setup(real, 6)
    q = oh.gram(c, real)
    s = np.array([[gdot(q[:, [i]], q[:, [j]], real)[0] for j in range(6)] for i in range(6)])
    assert np.max(np.abs(s - np.eye(6))) <= 1e-12
    # Cholesky factor is lower triangular: q spans c, the first column is c0 normalised
    c0 = c[:, 0] / math.sqrt(gdot(c[:, [0]], c[:, [0]], real)[0].real)
    assert np.max(np.abs(q[:, 0] - c0)) <= 1e-12


def test_exponential_is_fourth_order_taylor():
    rng = np.random.default_rng(9)
    a = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    H = a + a.conj().T
    psi = rng.standard_normal((6, 2)) + 1j * rng.standard_normal((6, 2))
    dt = 0.05
    got = oh.exponential(lambda w: H @ w, psi, dt)
    A = -1j * dt * H
    taylor = np.eye(6) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    assert np.max(np.abs(got - taylor @ psi)) <= 1e-14
    # and it is exp(-i dt H) to O(dt^5)
    w, U = np.linalg.eigh(H)
    exact = U @ np.diag(np.exp(-1j * dt * w)) @ U.conj().T
    assert np.max(np.abs(got - exact @ psi)) <= (dt * np.max(np.abs(w))) ** 5


def test_kpgpa2():
    _, b, _, _, _ = setup(False, 1)
    A = np.array([0.1, -0.2, 0.05])
    d = b.kpgx - A[None, :]
    assert np.max(np.abs(oh.kpgpa2(b, A) - np.sum(d * d, axis=1))) <= 1e-12


@pytest.mark.parametrize("real", [True, False])
def test_subspace_diag_identities(real):
    """H Z = Z w in the Γ metric: <Cz_i|H|Cz_j> = w_i delta_ij for orthonormal C, and the
    rotated block stays orthonormal; H built independently over the full sphere."""
    cell, b, maps, c, rng = setup(real, 5)
    c = oh.gram(c, real)
    sp = species(b, rng)
    v = rng.standard_normal(maps.N)
    hc, _ = oh.hpsi(maps, c, v, [sp], np.zeros(5), cell.volume)
    w, cz = oh.subspace_diag(c, hc, real)
    assert np.all(np.diff(w) >= 0)
    hcz, _ = oh.hpsi(maps, cz, v, [sp], np.zeros(5), cell.volume)
    m = np.array([[gdot(cz[:, [i]], hcz[:, [j]], real)[0] for j in range(5)] for i in range(5)])
    assert np.max(np.abs(m - np.diag(w))) <= 1e-10 * np.max(np.abs(w))
    s = np.array([[gdot(cz[:, [i]], cz[:, [j]], real)[0] for j in range(5)] for i in range(5)])
    assert np.max(np.abs(s - np.eye(5))) <= 1e-12


def test_etrs_step_splits_and_order():
    """ETRS = exp(-i dt/2 H(t+dt)) exp(-i dt/2 H(t)) to O(dt^5): pins the half steps and
    which Hamiltonian acts first (the two orders differ at O(dt^2) [H1, H2])."""
    rng = np.random.default_rng(10)
    n = 8
    mk = lambda: (lambda a: 0.25 * (a + a.conj().T))(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H1, H2 = mk(), mk()
    psi = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    dt = 0.02

    def expm(H, t):
        w, U = np.linalg.eigh(H)
        return U @ np.diag(np.exp(-1j * t * w)) @ U.conj().T

    seen = []

    def h2_of(tilde):
        seen.append(tilde)
        return lambda w: H2 @ w
    got, tilde = oh.etrs_step(lambda w: H1 @ w, h2_of, psi, dt)
    lam = max(np.max(np.abs(np.linalg.eigvalsh(H1))), np.max(np.abs(np.linalg.eigvalsh(H2))))
    bound = 3 * (dt * lam) ** 5
    # H(t+dt) is built from psi~(t+dt) = exp(-i dt H(t)) psi
    assert np.max(np.abs(seen[0] - expm(H1, dt) @ psi)) <= bound
    ref = expm(H2, 0.5 * dt) @ expm(H1, 0.5 * dt) @ psi
    assert np.max(np.abs(got - ref)) <= bound
    wrong = expm(H1, 0.5 * dt) @ expm(H2, 0.5 * dt) @ psi
    assert np.max(np.abs(got - wrong)) > 100 * bound


@pytest.mark.parametrize("real", [False, True])
def test_s_projector_sits_at_its_atom(real):
    """beta_s(r) = sum_G anl(G) e^{iG.r} (direct sum, no transform) with a positive radial
    table is real and largest at r = tau: the structure-factor phase e^{-iG.tau} of comp_eigr
    (NonLocalPotential.cc:1969, kpgr = -(k+G).tau) puts the projector on the atom.  A phase
    e^{+iG.tau} would put it at -tau instead."""
    cell = UnitCell(np.array([[6.0, 0.2, 0.0], [0.0, 5.5, 0.4], [0.3, 0.0, 6.5]]))
    b = Basis(cell, 5.0, force_complex=not real)
    tau = np.array([[1.3, 2.1, 0.7]])
    tw = np.exp(-0.3 * b.kpg2)[None, :]
    sp = oh.Projectors([0], [1.0], tw, tau)
    a = oh.anl_block(sp, b.kpgx, 0, 1)[:, 0]
    g = b.kpgx
    if real:  # the half sphere stands for G and -G: beta(-G) = conj(beta(G)), G=0 once
        g = np.concatenate([g, -g[1:]])
        a = np.concatenate([a, np.conj(a[1:])])

    def beta(r):
        return np.sum(a * np.exp(1j * (g @ r)))

    rng = np.random.default_rng(1)
    at = beta(tau[0])
    assert abs(at.imag) < 1e-9 * abs(at) and abs(at.real - np.sum(np.exp(-0.3 * (g * g).sum(1)))) < 1e-9 * abs(at)
    for _ in range(20):
        d = rng.normal(size=3)
        d *= rng.uniform(0.2, 2.0) / np.linalg.norm(d)
        assert abs(beta(tau[0] + d)) < at.real
    assert abs(beta(-tau[0])) < 0.5 * at.real
