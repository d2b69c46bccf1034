"""Synthetic inputs for the Hψ path (host side, numpy).

Systems are the benchmark configurations: diamond Si (8-atom conventional cell,
a = 10.26 bohr, replicated n x n x n), Γ point, ecut 25 Ry.  Inputs follow the
measurement plan: seeded complex N(0,1) coefficients on the stored G vectors damped by
1/(1+|G|^2) with real c(G=0); a seeded smooth random v_loc (white noise Gaussian-filtered
with sigma = 1 bohr, scaled to max |v| = 1 Ha); s + p Kleinman-Bylander projectors.

The projector radial form factors are a synthetic analytic pair (s: exp(-g^2 rc^2/2),
p: g*exp(-g^2 rc^2/2)) with the angular tables of NonLocalPotential::update_twnl
(s14pi*v for l=0, s34pi*(k+G)_m/|k+G|*v for l=1, NonLocalPotential.cc:334, :475-481);
the Hψ kernels take twnl as input, so the form factor only fixes the data, not the path.
The rod / plane-wave layout comes from the library's restatement of Basis::resize.
"""
import math

import numpy as np

from . import capi

RY = 0.5  # Hartree per Rydberg


class System:
    pass


def diamond_positions(ncell, alat0=10.26):
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
    basis = np.concatenate([base, base + 0.25])
    pos = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                pos.append((basis + np.array([i, j, k], dtype=np.float64)) * alat0)
    return np.concatenate(pos)


def cell_vectors(a):
    """UnitCell::set (UnitCell.cc:36-90): volume and b = 2 pi A^-T (same operation order)."""
    a0, a1, a2 = [np.asarray(x, dtype=np.float64) for x in a]

    def cross(x, y):
        return np.array([x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2],
                         x[0] * y[1] - x[1] * y[0]])
    vol = float(a0[0] * (a1[1] * a2[2] - a1[2] * a2[1]) + a0[1] * (a1[2] * a2[0] - a1[0] * a2[2])
                + a0[2] * (a1[0] * a2[1] - a1[1] * a2[0]))
    f = 1.0 / vol
    b = np.array([2.0 * math.pi * cross(f * a1, a2), 2.0 * math.pi * cross(f * a2, a0),
                  2.0 * math.pi * cross(f * a0, a1)])
    return vol, b


def plane_waves(rods, b, kpoint=(0.0, 0.0, 0.0)):
    """Miller indices and k+G of the basis in rod order (Basis::update_g order)."""
    h = np.repeat(rods["rod_h"], rods["rod_size"]).astype(np.float64)
    k = np.repeat(rods["rod_k"], rods["rod_size"]).astype(np.float64)
    first = np.concatenate([[0], np.cumsum(rods["rod_size"])[:-1]])
    l = (np.repeat(rods["rod_lmin"], rods["rod_size"])
         + (np.arange(int(rods["rod_size"].sum())) - np.repeat(first, rods["rod_size"]))
         ).astype(np.float64)
    kp = np.asarray(kpoint, dtype=np.float64)
    kpg = ((kp[0] + h)[:, None] * b[0] + (kp[1] + k)[:, None] * b[1]) + (kp[2] + l)[:, None] * b[2]
    return np.stack([h, k, l], axis=1).astype(np.int64), kpg


def density_grid_np(a, ecut):
    """ChargeDensity.cc:76-99: grid of the 4*ecut basis, np(i)+2 bumped to factorizable."""
    r = capi.rods_build(a, 4.0 * ecut, real=True)
    out = []
    for n in r["np"]:
        n += 2
        while not capi.factorizable(n):
            n += 2
        out.append(n)
    return tuple(out)


def make_system(ncell=2, ecut_ry=25.0, nbands=None, seed=0, complex_basis=False,
                generate_psi=True):
    """Diamond Si with 8*ncell^3 atoms.  Returns a System with numpy host arrays."""
    s = System()
    alat = 10.26 * ncell
    s.a = np.diag([alat, alat, alat])
    s.ecut = ecut_ry * RY
    s.omega, s.b = cell_vectors(s.a)
    s.real = not complex_basis
    s.rods = capi.rods_build(s.a, s.ecut, real=s.real)
    s.np = density_grid_np(s.a, s.ecut)
    s.idx, s.kpg = plane_waves(s.rods, s.b)
    s.ngw = s.idx.shape[0]
    s.g2 = (s.kpg ** 2).sum(axis=1)
    s.tau = diamond_positions(ncell)
    s.na = s.tau.shape[0]
    s.nst = nbands if nbands is not None else 2 * s.na  # 4 valence e / atom, doubly occupied
    s.occ = np.full(s.nst, 2.0)
    s.seed = seed
    rng = np.random.default_rng(seed)
    _fill_tables(s, rng, alat)
    if generate_psi:
        s.psi = random_psi(s.ngw, s.nst, s.g2, rng, real=s.real)
    return s


def make_fcc_al(ncell=5, alat0=7.65, ecut_ry=30.0, nbands=900, noccupied=750, seed=0,
                generate_psi=True):
    """fcc Al with 4*ncell^3 atoms (conventional cells of alat0 bohr), Γ, ecut_ry, s + p + d
    projectors (9 per atom: the d form factor g^2 exp(-0.45 g^2) with the l = 2 angular
    tables of NonLocalPotential::update_twnl, NonLocalPotential.cc:712-747), noccupied
    doubly occupied bands and the rest empty.  The large-cell form: 500 atoms, 900 bands."""
    s = System()
    alat = alat0 * ncell
    s.a = np.diag([alat, alat, alat])
    s.ecut = ecut_ry * RY
    s.omega, s.b = cell_vectors(s.a)
    s.real = True
    s.rods = capi.rods_build(s.a, s.ecut, real=True)
    s.np = density_grid_np(s.a, s.ecut)
    s.idx, s.kpg = plane_waves(s.rods, s.b)
    s.ngw = s.idx.shape[0]
    s.g2 = (s.kpg ** 2).sum(axis=1)
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
    s.tau = np.concatenate([(base + np.array([i, j, k], dtype=np.float64)) * alat0
                            for i in range(ncell) for j in range(ncell) for k in range(ncell)])
    s.na = s.tau.shape[0]
    s.nst = nbands
    s.occ = np.where(np.arange(nbands) < noccupied, 2.0, 0.0)
    s.seed = seed
    rng = np.random.default_rng(seed)
    _fill_tables(s, rng, alat)
    g = np.sqrt(s.g2)
    gi = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    x, y, z = (s.kpg[:, d] * gi for d in range(3))
    vd = 0.7 * g * g * np.exp(-0.45 * g * g)
    s15 = math.sqrt(15.0 / (4 * math.pi))
    s5 = math.sqrt(5.0 / (16 * math.pi))
    dtab = [s15 * x * y * vd, s15 * y * z * vd, s15 * z * x * vd,
            0.5 * s15 * (x * x - y * y) * vd, s5 * (3 * z * z - 1.0) * vd]
    s.twnl = np.concatenate([s.twnl, np.stack(dtab)])
    s.lproj = np.array(list(s.lproj) + [2] * 5, dtype=np.int32)
    s.wt = np.array(list(s.wt) + [0.8] * 5)
    if generate_psi:
        s.psi = random_psi(s.ngw, s.nst, s.g2, rng, real=True)
    return s


def _fill_tables(s, rng, box):
    """s + p projector tables (4 per atom) and the smooth random v_loc of a cubic box."""
    g = np.sqrt(s.g2)
    rc = 1.0
    vs = 2.0 * np.exp(-0.5 * (g * rc) ** 2)
    vp = 1.5 * g * np.exp(-0.5 * (g * rc) ** 2)
    fpi = 4.0 * math.pi
    s14pi = math.sqrt(1.0 / fpi)
    s34pi = math.sqrt(3.0 / fpi)
    gi = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    s.lproj = np.array([0, 1, 1, 1], dtype=np.int32)
    s.wt = np.array([1.1, -0.6, -0.6, -0.6])
    s.twnl = np.stack([s14pi * vs, s34pi * s.kpg[:, 0] * gi * vp, s34pi * s.kpg[:, 1] * gi * vp,
                       s34pi * s.kpg[:, 2] * gi * vp])
    # v_loc: white noise filtered with exp(-sigma^2 G^2 / 2), sigma = 1 bohr, max |v| = 1 Ha
    n0, n1, n2 = s.np
    noise = rng.standard_normal((n2, n1, n0))
    fz = np.fft.fftfreq(n2) * n2
    fy = np.fft.fftfreq(n1) * n1
    fx = np.fft.fftfreq(n0) * n0
    bb = 2 * math.pi / box
    G2 = (fz[:, None, None] * bb) ** 2 + (fy[None, :, None] * bb) ** 2 + (fx[None, None, :] * bb) ** 2
    v = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-0.5 * G2)))
    s.v = (v / np.abs(v).max()).ravel().astype(np.float64)


def smooth5(n):
    """Smallest even 5-smooth integer >= n."""
    m = max(2, n + (n % 2))
    while True:
        q = m
        for p in (2, 3, 5):
            while q % p == 0:
                q //= p
        if q == 1:
            return m
        m += 2


def make_batch_cell(box, natoms, ecut_ry, nbands, seed, generate_psi=True):
    """One cell of the multi-cell workload: a cubic box of edge `box` bohr holding `natoms`
    atoms at jittered diamond sites (first natoms sites of an m^3 diamond supercell scaled
    to the box, jitter N(0, 0.1 bohr)), Γ, ecut_ry, nbands bands, s + p projectors; the
    Hψ grid is the density-grid rule rounded up to 5-smooth lengths."""
    s = System()
    rng = np.random.default_rng(seed)
    s.a = np.diag([box, box, box])
    s.ecut = ecut_ry * RY
    s.omega, s.b = cell_vectors(s.a)
    s.real = True
    s.rods = capi.rods_build(s.a, s.ecut, real=True)
    s.np = tuple(smooth5(n) for n in density_grid_np(s.a, s.ecut))
    s.idx, s.kpg = plane_waves(s.rods, s.b)
    s.ngw = s.idx.shape[0]
    s.g2 = (s.kpg ** 2).sum(axis=1)
    m = 1
    while 8 * m ** 3 < natoms:
        m += 1
    sites = diamond_positions(m, alat0=box / m)
    s.tau = sites[:natoms] + rng.normal(0.0, 0.1, (natoms, 3))
    s.na = natoms
    s.nst = nbands
    s.occ = np.full(nbands, 2.0)
    s.seed = seed
    _fill_tables(s, rng, box)
    if generate_psi:
        s.psi = random_psi(s.ngw, s.nst, s.g2, rng, real=True)
    return s


def batch_workload(ncells=12, seed=0, ecut_ry=30.0):
    """Parameters of the uneven multi-cell workload: per cell (box in [12, 28] bohr,
    natoms in [32, 48], nbands in [64, 320]), uniform draws from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ncells):
        box = float(rng.uniform(12.0, 28.0))
        natoms = int(rng.integers(32, 49))
        nbands = int(rng.integers(64, 321))
        out.append(dict(box=box, natoms=natoms, nbands=nbands, ecut_ry=ecut_ry, seed=seed * 100 + i))
    return out


def random_psi(ngw, nst, g2, rng, real=True):
    """Complex N(0,1) coefficients / (1+|G|^2); Γ: c(G=0) real (SlaterDet::cleanup).
    Column-major complex128 array (ngw, nst), Fortran order."""
    c = np.empty((ngw, nst), dtype=np.complex128, order="F")
    damp = 1.0 / (1.0 + g2)
    for n0 in range(0, nst, 64):
        n1 = min(nst, n0 + 64)
        blk = rng.standard_normal((n1 - n0, ngw, 2))
        c[:, n0:n1] = ((blk[..., 0] + 1j * blk[..., 1]) * damp[None, :]).T
    if real:
        c[0, :] = c[0, :].real
    return c


# --------------------------------------------------------------------------------------
# real-time TDDFT inputs (the differentiation form)
def field_shape(nsteps, dt, omega=0.3, t0=None, tau=None, direction=(1.0, 0.0, 0.0)):
    """Time profile of the uniform velocity-gauge field A(t) = A0 sin(omega t)
    exp(-(t - t0)^2 / (2 tau^2)) along `direction` at t_n = n dt, n = 0..nsteps; returns
    (nsteps+1, 3) so that A_n = A0 * shape_n."""
    T = nsteps * dt
    t0 = 0.5 * T if t0 is None else t0
    tau = T / 6.0 if tau is None else tau
    t = dt * np.arange(nsteps + 1)
    s = np.sin(omega * t) * np.exp(-((t - t0) ** 2) / (2.0 * tau * tau))
    d = np.asarray(direction, dtype=np.float64)
    return np.outer(s, d / math.sqrt(float(d @ d)))


def dipole_weights(nsteps, dt):
    """Weights w_n for J = integral_0^T d(t) dt with d(t_n) = sum_{m<n} dt j(t_m) (the
    dipole accumulated from the total current, rectangle rule): w_m = dt^2 (nsteps - m)."""
    m = np.arange(nsteps + 1)
    return dt * dt * (nsteps - m).astype(np.float64)


def td_ionic_inputs(sysm, vrods, valence=4.0, vloc=-10.0):
    """Ionic inputs of physical size on the density basis with rods `vrods` (its G order):
    Gaussian ion charges rho_ps(G) = -(valence/Omega) S(G) exp(-G^2/4), neutral with
    valence * natoms electrons, and a short-range local potential v_ion(G) = (vloc/Omega)
    S(G) exp(-G^2/2), S(G) = sum_a exp(-i G.tau_a).  Both stay a few Ha in real space for
    any cell size, unlike per-G inputs of fixed size, whose Hartree potential grows with the
    cell.  Returns (vion, rhops), complex per G."""
    _, kpg = plane_waves(vrods, sysm.b)
    g2 = (kpg ** 2).sum(axis=1)
    sg = np.exp(-1j * (kpg @ sysm.tau.T)).sum(axis=1)
    rhops = -(valence / sysm.omega) * sg * np.exp(-0.25 * g2)
    vion = (vloc / sysm.omega) * sg * np.exp(-0.5 * g2)
    return vion, rhops
