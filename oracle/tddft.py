"""Real-time TDDFT propagation and its discrete adjoint -- oracle restatement.
Test infrastructure only.

Forward: the EhrenSampleStepper iteration (``src/qball/EhrenSampleStepper.cc:280-1152``)
for fixed ions, one spin, one k-point, complex basis:

* iteration start (``:323-330``): rho(psi_n) (ChargeDensity::update_density), v_a =
  update_vhxc (EnergyFunctional.cc:353-977) + v_ext; total current j(psi_n, A_n)
  (CurrentDensity::update_current, ``:359``);
* ETRS preupdate under H[v_a, A_n] (ExponentialWavefunctionStepper.cc:151-240);
* A_n -> A_{n+1} (``:480``);
* update (ExponentialWavefunctionStepper.cc:243-276): v_b = update_vhxc(rho(psi~)) + v_ext
  and psi_{n+1} = T(dt/2) psi(t+dt/2) under H[v_b, A_{n+1}].

T is the order-4 Taylor exponential (ExponentialWavefunctionStepper.cc:51-148).  A enters H
through 0.5|k+G-A|^2 and, when the system holds the species (``retable``), through the
projector tables re-tabled at |k+G-A| on every change of A (NonLocalPotential::update_twnl
under a vector potential, EnergyFunctional::vector_potential_changed,
EnergyFunctional.cc:2047-2057); given fixed tables (``species``) they stay at |k+G|.  The
projector phases at k+G-A cancel exactly in beta D beta^H either way.

Objective: J = sum_n w_n e.j(psi_n, A_n) with j_d = -sum_b f_b sum_G (k+G)_d |c_bG|^2 +
A_d N_e (the G-space form of the total current, equal to the grid sum of update_current
by Parseval).

Adjoint: the exact derivative of this discrete map (Wirtinger gradients, dJ = 2 Re <g,
dpsi>), with f_xc of the PZ81 functional and the Hartree operator; pinned by central
finite differences in the tests.
"""
import math

import numpy as np

from . import fft as offt
from . import hpsi as oh
from . import species as osp
from .vhxc import update_vhxc


def fxc_pz81(rho):
    """d v_xc / d rho of LDAFunctional::xc_unpolarized (PZ81, same constants), zero where
    rho <= 0."""
    rho = np.asarray(rho, dtype=np.float64)
    c1 = 0.6203504908994001
    c3 = -0.610887057711
    A, B, b1, b2, G = 0.0311, -0.048, 1.0529, 0.3334, -0.1423
    D = G / (1.0 + b1 + b2) - B
    C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)))
    out = np.zeros_like(rho)
    m = rho > 0.0
    r = rho[m]
    rs = c1 / np.cbrt(r)
    vx = c3 / rs
    dvx = vx / (3.0 * r)
    dvc_drs = np.empty_like(rs)
    lo = rs < 1.0
    lr = np.log(rs[lo])
    dvc_drs[lo] = A / rs[lo] + (2.0 / 3.0) * C * (lr + 1.0) + (2.0 * D - C) / 3.0
    hi = ~lo
    s = np.sqrt(rs[hi])
    den = 1.0 + b1 * s + b2 * rs[hi]
    num = 1.0 + (7.0 / 6.0) * b1 * s + (4.0 / 3.0) * b2 * rs[hi]
    dnum = (7.0 / 6.0) * b1 + (8.0 / 3.0) * b2 * s
    dden = b1 + 2.0 * b2 * s
    dvc_ds = G * (dnum * den - 2.0 * num * dden) / den ** 3
    dvc_drs[hi] = dvc_ds / (2.0 * s)
    out[m] = dvx + dvc_drs * (-rs / (3.0 * r))
    return out


def hartree_apply(vmaps, mu):
    """K_H mu = Re FFT^-1[4 pi / G^2 FFT(mu)] on the density sphere (G != 0): the linear
    map rho -> v_H of update_vhxc, symmetric under sum_r."""
    vb = vmaps.basis
    f = np.asarray(mu, dtype=np.complex128).reshape(vmaps.np2, vmaps.np1, vmaps.np0)
    mg = offt.forward(vmaps, f, pair=False)
    g = np.sqrt(vb.g2)
    gi = np.where(g > 0.0, 1.0 / np.where(g > 0.0, g, 1.0), 0.0)
    return offt.backward(vmaps, 4.0 * math.pi * gi * gi * mg).ravel().real


class TDSystem:
    """Everything the propagation needs besides psi, v_ext and A(t)."""

    def __init__(self, maps, vmaps, species, occ, omega, vion_g, rhops_g, weight=1.0,
                 retable=None):
        """species: oracle.hpsi.Projectors with fixed tables; or retable: [(oracle.species
        .Species, tau)] whose tables follow A."""
        self.maps, self.vmaps = maps, vmaps
        self.species = species
        self.retable = retable
        self.occ = np.asarray(occ, dtype=np.float64)
        self.omega = omega
        self.vion_g, self.rhops_g = vion_g, rhops_g
        self.weight = weight
        self.basis = maps.basis

    def potential(self, c, vext):
        rho, _, rhog = oh.update_density(self.maps, self.vmaps, c, self.occ, self.weight, self.omega)
        v = update_vhxc(self.vmaps, rho, rhog, self.vion_g, self.rhops_g, self.omega,
                        hamil_rhog=rhog)[0]
        if vext is not None:
            v = v + vext
        return v, rho

    def projectors(self, A):
        """Projectors at A: the fixed ones, or the species' tables at |k+G-A|."""
        if self.retable is None:
            return self.species
        kpgx = self.basis.kpgx - np.asarray(A, dtype=np.float64)[None, :]
        out = []
        for sp, tau in self.retable:
            rows, ls, wt = osp.twnl_table(sp, kpgx)
            out.append(oh.Projectors(ls, wt, rows, tau))
        return out

    def apply_h(self, v, A):
        b = self.basis
        kpg2 = oh.kpgpa2(b, A)
        kpgx = b.kpgx - np.asarray(A, dtype=np.float64)[None, :]
        occ0 = np.zeros(len(self.occ))
        proj = self.projectors(A)
        return lambda w: oh.hpsi(self.maps, w, v, proj, occ0, self.omega, kpg2=kpg2,
                                 kpgx=kpgx)[0]

    def nl_dA(self, A, z, U):
        """2 Re sum_b <z_b| dH_NL/dA_d |U_b> (3,) for tables re-tabled at |k+G-A|:
        dH_NL/dA_d = sum_p |d_d beta_p> D_p <beta_p| + h.c. with d_d beta_p = -(-i)^l
        (d twnl_p / dq_d) exp(-i (k+G-A).tau) (the A-derivative of the phases cancels)."""
        out = np.zeros(3)
        if self.retable is None:
            return out
        kpgx = self.basis.kpgx - np.asarray(A, dtype=np.float64)[None, :]
        for sp, tau in self.retable:
            rows, ls, wt = osp.twnl_table(sp, kpgx)
            drows = osp.twnl_grad_table(sp, kpgx)
            na = len(tau)
            fac = np.repeat(np.asarray(wt) / self.omega, na)   # column ia + ipr * na
            anl = oh.anl_block(oh.Projectors(ls, wt, rows, tau), kpgx, 0, na)
            Fz, FU = anl.conj().T @ z, anl.conj().T @ U
            for d in range(3):
                dbeta = -oh.anl_block(oh.Projectors(ls, wt, drows[d], tau), kpgx, 0, na)
                dFz, dFU = dbeta.conj().T @ z, dbeta.conj().T @ U
                out[d] += 2.0 * float(np.sum(fac[:, None] * (np.conj(dFz) * FU +
                                                             np.conj(Fz) * dFU)).real)
        return out

    def current(self, c, A):
        """(total current (3,), N_e): G-space form of CurrentDensity::update_current."""
        f = self.weight * self.occ
        w2 = np.abs(c) ** 2 * f[None, :]
        ne = float(w2.sum())
        return -(w2.sum(axis=1) @ self.basis.kpgx) + np.asarray(A) * ne, ne


def taylor_coeffs(dt, order=4):
    """alpha_N = (-i dt)^N / N!, built as the reference builds factor (factor *= -i dt/N)."""
    a = [1.0 + 0.0j]
    f = 1.0 + 0.0j
    for N in range(1, order + 1):
        f *= -1j * dt / float(N)
        a.append(f)
    return a


def expo_krylov(apply_h, x, dt, order=4):
    """exponential() with the Krylov vectors u_k = H^k x, k < order."""
    a = taylor_coeffs(dt, order)
    u = [x]
    y = x.copy()
    w = x
    for N in range(1, order + 1):
        w = apply_h(w)
        y = y + a[N] * w
        if N < order:
            u.append(w)
    return y, u


def td_forward(sysm, psi0, vext, A, w, e, dt, nsteps, order=4, keep=False):
    """J = sum_{n<=nsteps} w_n e.j(psi_n, A_n).  A: (nsteps+1, 3).  Returns (J, psi_final,
    trajectory or None)."""
    e = np.asarray(e, dtype=np.float64)
    c = psi0
    J = 0.0
    traj = []
    for n in range(nsteps):
        j, _ = sysm.current(c, A[n])
        J += w[n] * float(e @ j)
        va, rho_a = sysm.potential(c, vext)
        Ha = sysm.apply_h(va, A[n])
        half, u1 = expo_krylov(Ha, c, 0.5 * dt, order)
        tilde, u2 = expo_krylov(Ha, half, 0.5 * dt, order)
        vb, rho_b = sysm.potential(tilde, vext)
        Hb = sysm.apply_h(vb, A[n + 1])
        c_new, u3 = expo_krylov(Hb, half, 0.5 * dt, order)
        if keep:
            traj.append(dict(psi=c, rho_a=rho_a, Ha=Ha, u1=u1, u2=u2, half=half, tilde=tilde,
                             rho_b=rho_b, Hb=Hb, u3=u3))
        c = c_new
    j, _ = sysm.current(c, A[nsteps])
    J += w[nsteps] * float(e @ j)
    return J, c, (traj if keep else None)


def _expo_adjoint(apply_h, lam, dt, order):
    """T^H lam (= Taylor with dt -> -dt) and z_k = H^k lam, k < order."""
    return expo_krylov(apply_h, lam, -dt, order)


def _cross(sysm, z, u, dt, order, A):
    """(mu(r), gA) of one exponential under A: mu = (2/N) Re sum_b sum_k conj(z_k(r)) U_k(r)
    with U_k = sum_{m <= order-1-k} alpha_{k+m+1} u_m; gA_d = -2 Re sum conj(z_k) kvec_d U_k
    (kvec = k+G-A) plus, with re-tabled projectors, 2 Re sum <z_k| dH_NL/dA_d |U_k>."""
    kvec = sysm.basis.kpgx - np.asarray(A, dtype=np.float64)[None, :]
    a = taylor_coeffs(dt, order)
    maps = sysm.maps
    N = maps.N
    mu = np.zeros(N)
    gA = np.zeros(3)
    for k in range(order):
        U = sum(a[k + m + 1] * u[m] for m in range(order - k))
        for b in range(U.shape[1]):
            zr = offt.backward(maps, z[k][:, b]).ravel()
            Ur = offt.backward(maps, U[:, b]).ravel()
            mu += (2.0 / N) * (np.conj(zr) * Ur).real
        gA += -2.0 * np.einsum("gb,gd,gb->d", np.conj(z[k]), kvec, U).real
        gA += sysm.nl_dA(A, z[k], U)
    return mu, gA


def _density_adjoint(sysm, c, nu):
    """g_b = (weight occ_b N / Omega) FFT[nu psi_b] on the sphere."""
    out = np.zeros_like(c)
    oh.rs_mul_add(sysm.maps, c, nu, out)
    fac = sysm.weight * sysm.occ * sysm.maps.N / sysm.omega
    fac = np.where(sysm.weight * sysm.occ > 0.0, fac, 0.0)
    return out * fac[None, :]


def td_gradient(sysm, psi0, vext, A, w, e, dt, nsteps, order=4):
    """(J, dJ/dv_ext (grid), dJ/dA (nsteps+1, 3)) by the discrete adjoint."""
    e = np.asarray(e, dtype=np.float64)
    J, cM, traj = td_forward(sysm, psi0, vext, A, w, e, dt, nsteps, order, keep=True)
    b = sysm.basis
    f = sysm.weight * sysm.occ
    N = sysm.maps.N
    gv = np.zeros(N)
    gA = np.zeros((nsteps + 1, 3))

    def dj(c, An):
        return -(f[None, :] * ((b.kpgx - An[None, :]) @ e)[:, None]) * c

    _, ne = sysm.current(cM, A[nsteps])
    gA[nsteps] += w[nsteps] * e * ne
    lam = w[nsteps] * dj(cM, A[nsteps])
    h = 0.5 * dt
    for n in range(nsteps - 1, -1, -1):
        t = traj[n]
        lam_h, z3 = _expo_adjoint(t["Hb"], lam, h, order)
        mu_b, g = _cross(sysm, z3, t["u3"], h, order, A[n + 1])
        gA[n + 1] += g
        gv += mu_b
        nu_b = fxc_pz81(t["rho_b"]) * mu_b + hartree_apply(sysm.vmaps, mu_b)
        lam_t = _density_adjoint(sysm, t["tilde"], nu_b)
        l2, z2 = _expo_adjoint(t["Ha"], lam_t, h, order)
        lam_h = lam_h + l2
        mu_a, g = _cross(sysm, z2, t["u2"], h, order, A[n])
        gA[n] += g
        lam_psi, z1 = _expo_adjoint(t["Ha"], lam_h, h, order)
        m1, g = _cross(sysm, z1, t["u1"], h, order, A[n])
        mu_a += m1
        gA[n] += g
        gv += mu_a
        nu_a = fxc_pz81(t["rho_a"]) * mu_a + hartree_apply(sysm.vmaps, mu_a)
        lam_psi = lam_psi + _density_adjoint(sysm, t["psi"], nu_a)
        _, ne = sysm.current(t["psi"], A[n])
        gA[n] += w[n] * e * ne
        lam = lam_psi + w[n] * dj(t["psi"], A[n])
    return J, gv, gA
