"""SCF driver pieces -- oracle restatement.  Test infrastructure only.

* ``prec_diag``: Preconditioner::update (``src/qball/Preconditioner.cc:47-85``, no
  confinement): K_G = 0.5 / max(e_G, ecutprec) with e_G = |k+G|^2 / 2.
* ``PSDA``: PSDAWavefunctionStepper::update (``PSDAWavefunctionStepper.cc:38-375``): residual
  HC - C A with A = C^T(HC) in the basis metric, preconditioned descent -K(HC - CA), and the
  Anderson extrapolation over one previous step (occupation-weighted theta, clamps
  theta < -1 -> 0 and theta <= 2); ``preprocess`` clears the history (``.h:53-58``).
* ``AndersonMixer``: ``AndersonMixer.cc`` (history of nmax previous pairs, Tikhonov-
  regularised symmetric solve until sum theta^2 <= 1).
* ``mixing_weights``: Kerker and ``wls`` weights of ``BOSampleStepper.cc:394-447``.
* ``scf``: the BOSampleStepper SCF loop for fixed ions (``BOSampleStepper.cc:880-1130``):
  update_density, Anderson/Kerker charge mixing on rhog, update_rhor
  (``ChargeDensity.cc:558-580``), update_vhxc, then nite x (energy(compute_hpsi) -> PSDA
  update -> gram).
"""
import math

import numpy as np

from . import fft as offt
from . import hpsi as oh
from .vhxc import update_vhxc


def prec_diag(basis, ecutprec):
    e = 0.5 * basis.kpg2
    return np.where(e < ecutprec, 0.5 / ecutprec, 0.5 / np.where(e < ecutprec, 1.0, e))


def _metric_product(c, hc, real):
    """A = 2 C^T HC - c0 hc0^T over the real view (Γ) or C^H HC."""
    if real:
        a = 2.0 * (np.ascontiguousarray(c.real).T @ np.ascontiguousarray(hc.real)
                   + np.ascontiguousarray(c.imag).T @ np.ascontiguousarray(hc.imag))
        a -= np.outer(c[0].real, hc[0].real)
        return a
    return c.conj().T @ hc


class PSDA:
    def __init__(self, basis, ecutprec):
        self.real = basis.real
        self.K = prec_diag(basis, ecutprec)
        self.c_last = None
        self.dc_last = None
        self.extrapolate = False

    def preprocess(self):
        self.extrapolate = False

    def update(self, c, hc, occ):
        """Returns (new c before orthonormalisation, theta)."""
        a = _metric_product(c, hc, self.real)
        r = hc - c @ a
        dc = -(self.K[:, None] * r)
        theta = 0.0
        x = c
        if self.extrapolate:
            df = dc - self.dc_last
            wa = (occ[None, :] * (dc.real * df.real + dc.imag * df.imag)).sum()
            wb = (occ[None, :] * (df.real ** 2 + df.imag ** 2)).sum()
            if self.real:
                wa = 2.0 * wa - (occ * (dc[0].real * df[0].real + dc[0].imag * df[0].imag)).sum()
                wb = 2.0 * wb - (occ * (df[0].real ** 2 + df[0].imag ** 2)).sum()
            if wb != 0.0:
                theta = -wa / wb
            if theta < -1.0:
                theta = 0.0
            theta = min(2.0, theta)
            xbar = x + theta * (x - self.c_last)
            fbar = dc + theta * (dc - self.dc_last)
            new = xbar + fbar
        else:
            new = x + dc
        self.c_last = x.copy()
        self.dc_last = dc.copy()
        self.extrapolate = True
        return new, theta


class AndersonMixer:
    def __init__(self, m, nmax):
        self.m, self.nmax = m, nmax
        self.x = [None] * (nmax + 1)
        self.f = [None] * (nmax + 1)
        self.restart()

    def restart(self):
        self.n = -1
        self.k = -1

    def update(self, x, f):
        nm = self.nmax
        self.k = (self.k + 1) % (nm + 1)
        if self.n < nm:
            self.n += 1
        k, n = self.k, self.n
        self.x[k] = np.array(x, dtype=np.float64)
        self.f[k] = np.array(f, dtype=np.float64)
        theta = np.zeros(max(n, 0))
        if n > 0:
            a = np.zeros((n, n))
            b = np.zeros(n)
            for i in range(n):
                kmi = (k - i + nm) % (nm + 1)
                for j in range(i + 1):
                    kmj = (k - j + nm) % (nm + 1)
                    a[i, j] = float(np.sum((self.f[k] - self.f[kmi]) * (self.f[k] - self.f[kmj])))
                b[i] = float(np.sum((self.f[k] - self.f[kmi]) * self.f[k]))
            a = np.tril(a) + np.tril(a, -1).T
            tp = 1e-12
            for _ in range(100):
                t = np.linalg.solve(a + tp * np.eye(n), b)
                theta = t
                if float(np.sum(t * t)) <= 1.0:
                    break
                tp *= 2.0
        fbar = self.f[k].copy()
        xbar = self.x[k].copy()
        for i in range(n):
            kmi = (k - i + nm) % (nm + 1)
            fbar -= theta[i] * (self.f[k] - self.f[kmi])
            xbar -= theta[i] * (self.x[k] - self.x[kmi])
        return xbar, fbar, theta


def mixing_weights(vbasis, rc_kerker=10.0, rc1=3.0):
    g2 = vbasis.g2
    q1 = 2.0 * math.pi / rc1
    wls = np.where(g2 != 0.0, np.sqrt(g2 / np.where(g2 != 0.0, g2 + q1 * q1, 1.0)), 1.0)
    if rc_kerker > 0.0:
        q0 = 2.0 * math.pi / rc_kerker
        wk = g2 / (g2 + q0 * q0)
    else:
        wk = np.ones_like(g2)
    return wk, wls


def kinetic_energy(basis, c, occ, weight=1.0):
    """sum_n occ_n <c_n| |k+G|^2/2 |c_n> in the basis metric (Γ: 2 sum - G=0 term)."""
    t = (0.5 * basis.kpg2)[:, None] * (np.abs(c) ** 2) * occ[None, :]
    s = float(t.sum())
    if basis.real:
        s = 2.0 * s - float(t[0].sum())
    return weight * s


def update_rhor(vmaps, rhog, omega):
    return offt.backward(vmaps, rhog).ravel().real / omega


EVOLT = 0.036749023  # 1 eV in Hartree (Wavefunction.cc:1246)


def fermi(e, mu, width):
    """SlaterDet::fermi (SlaterDet.cc:1897-1915): e, mu in Hartree, width in Rydberg."""
    e = np.asarray(e, dtype=np.float64)
    if width == 0.0:
        return np.where(e < mu, 1.0, np.where(e == mu, 0.5, 0.0))
    kt = 0.5 * width
    arg = (e - mu) / kt
    out = 1.0 / (1.0 + np.exp(np.clip(arg, -30.0, 30.0)))
    return np.where(arg < -30.0, 1.0, np.where(arg > 30.0, 0.0, out))


def update_occ(eig, nel, width, maxiter=10000):
    """Wavefunction::update_occ, one spin, one k-point of weight 1, Fermi smearing
    (Wavefunction.cc:1214-1361, SlaterDet::update_occ :1840-1881): mu by the reference's
    step-halving search from 0 with steps of 10 eV until |sum occ - nel| <= 1e-8;
    occ = 2 f(eig, mu, width).  Returns (occ, mu)."""
    mu, dmu, up = 0.0, 10.0 * EVOLT, True
    occ = 2.0 * fermi(eig, mu, width)
    rhosum = float(np.sum(occ))
    niter = 0
    while niter < maxiter and abs(rhosum - nel) > 1.0e-8:
        niter += 1
        if rhosum < nel:
            if not up:
                dmu /= 2.0
            mu += dmu
            up = True
        else:
            if up:
                dmu /= 2.0
            mu -= dmu
            up = False
        occ = 2.0 * fermi(eig, mu, width)
        rhosum = float(np.sum(occ))
    return occ, mu


def entropy(occ):
    """SlaterDet::entropy, one spin (SlaterDet.cc:1948-1962): -sum 2 (f ln f + (1-f)
    ln(1-f)), f = occ / 2, over 0 < f < 1."""
    f = np.asarray(occ, dtype=np.float64) / 2.0
    m = (f > 0.0) & (f < 1.0)
    fm = f[m]
    return float(-np.sum(2.0 * (fm * np.log(fm) + (1.0 - fm) * np.log(1.0 - fm))))


def scf(maps, vmaps, species, c, occ, omega, vion_g, rhops_g, nitscf, nite, ecutprec,
        alpha=0.5, ndim=3, rc_kerker=10.0, hpsi=None, smearing=0.0, nel=None, diag=False):
    """Returns (c, history) with history[itscf] = list of per-ite dicts (etotal and parts,
    PSDA theta).  ``hpsi(c)`` may replace the oracle Hψ (returns (hc, enl)).

    smearing > 0 (Rydberg, Fermi): fractional occupations as BOSampleStepper runs them
    (BOSampleStepper.cc:1150-1185): after each cycle's iterations Hψ, the subspace
    diagonalisation with eigenvectors (Wavefunction::diag), update_occ for nel electrons;
    etotal carries ets = -entropy * smearing / 2 (EnergyFunctional.cc:1602-1609).  The
    history then also holds each cycle's eigenvalues, occupations and mu (``cycles``).
    diag (wf_diag = T, BOSampleStepper.cc:197, 1150-1163): the same Hψ and subspace
    diagonalisation with eigenvectors after each cycle, occupations unchanged."""
    occ = np.array(occ, dtype=np.float64)
    cycles = []
    basis, vb = maps.basis, vmaps.basis
    if hpsi is None:
        def hpsi(x):
            return oh.hpsi(maps, x, v, species, occ, omega)
    psda = PSDA(basis, ecutprec)
    mixer = AndersonMixer(2 * vb.size, ndim)
    wk, wls = mixing_weights(vb, rc_kerker)
    rhog_in = None
    hist = []
    for itscf in range(nitscf):
        rhor, nel, rhog = oh.update_density(maps, vmaps, c, occ, 1.0, omega)
        if itscf == 0:
            rhog_in = rhog.copy()
        else:
            drhog = (rhog - rhog_in) / wls
            xbar, fbar, _ = mixer.update(rhog_in.view(np.float64), drhog.view(np.float64))
            rhobar = xbar.view(np.complex128)
            drhobar = fbar.view(np.complex128) * wls
            rhog_in = rhobar + alpha * drhobar * wk
            rhog = rhog_in.copy()
            rhor = update_rhor(vmaps, rhog, omega)
        v, eps, ehart, exc = update_vhxc(vmaps, rhor, rhog, vion_g, rhops_g, omega)
        psda.preprocess()
        steps = []
        for ite in range(max(nite, 1)):
            hc, enl = hpsi(c)
            ekin = kinetic_energy(basis, c, occ)
            ets = -entropy(occ) * smearing * 0.5 if smearing > 0.0 else 0.0
            etot = ekin + enl + eps + ehart + exc + ets
            c, theta = psda.update(c, hc, occ)
            c = oh.gram(c, basis.real)
            steps.append(dict(etotal=etot, ekin=ekin, enl=enl, eps=eps, ehart=ehart, exc=exc,
                              theta=theta, ets=ets))
        hist.append(steps)
        if smearing > 0.0 or diag:
            hc, _ = hpsi(c)
            eig, c = oh.subspace_diag(c, hc, basis.real)
            mu = None
            if smearing > 0.0:
                occ, mu = update_occ(eig, nel, smearing)
            cycles.append(dict(eig=eig, occ=occ.copy(), mu=mu))
    if smearing > 0.0 or diag:
        return c, hist, cycles
    return c, hist
