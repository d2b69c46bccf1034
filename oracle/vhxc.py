"""Self-consistent local potential -- oracle restatement.  Test infrastructure only.

Follows ``src/qball/EnergyFunctional.cc:353-977`` (``update_vhxc``, nspin = 1, LDA, no ESM)
and ``src/functionals/LDAFunctional.cc`` (``xc_unpolarized``).
"""
import math

import numpy as np

from . import fft as offt


def pz81(rho):
    """LDAFunctional::xc_unpolarized: Slater exchange + Perdew-Zunger 1981 correlation
    (Phys. Rev. B 23, 5048), C and D from continuity at rs = 1.  Returns (e_xc, v_xc)
    arrays (Hartree); zero where rho <= 0."""
    rho = np.asarray(rho, dtype=np.float64)
    c1 = 0.6203504908994001
    c3 = -0.610887057711
    A, B, b1, b2, G = 0.0311, -0.048, 1.0529, 0.3334, -0.1423
    D = G / (1.0 + b1 + b2) - B
    C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)))
    ee = np.zeros_like(rho)
    vv = np.zeros_like(rho)
    m = rho > 0.0
    rs = c1 / np.cbrt(rho[m])
    vx = c3 / rs
    ex = 0.75 * vx
    lo = rs < 1.0
    ec = np.empty_like(rs)
    vc = np.empty_like(rs)
    lr = np.log(rs[lo])
    ec[lo] = A * lr + B + C * rs[lo] * lr + D * rs[lo]
    vc[lo] = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs[lo] * lr + ((2.0 * D - C) / 3.0) * rs[lo]
    hi = ~lo
    sq = np.sqrt(rs[hi])
    den = 1.0 + b1 * sq + b2 * rs[hi]
    ec[hi] = G / den
    vc[hi] = ec[hi] * (1.0 + (7.0 / 6.0) * b1 * sq + (4.0 / 3.0) * b2 * rs[hi]) / den
    ee[m] = ex + ec
    vv[m] = vx + vc
    return ee, vv


def update_vhxc(vmaps, rhor, rhog, vion_g, rhops_g, omega, hamil_rhog=None):
    """EnergyFunctional::update_vhxc for one spin.  ``rhog`` is ChargeDensity::rhog (the
    forward transform of Omega*rho); returns (v_r, eps, ehart, exc)."""
    vb = vmaps.basis
    N = vmaps.N
    rhoelg = rhog / omega                                       # :364-368
    ee, vxc = pz81(rhor)                                        # XCPotential::update
    exc = float(np.sum(rhor * ee)) * omega / N
    ddot = float(np.sum(rhoelg.real * vion_g.real + rhoelg.imag * vion_g.imag))
    if vb.real:                                                 # :459-467
        eps = 2.0 * ddot - float((np.conj(rhoelg[0]) * vion_g[0]).real)
    else:
        eps = ddot
    eps *= omega
    g = np.sqrt(vb.g2)
    gi = np.where(g > 0.0, 1.0 / np.where(g > 0.0, g, 1.0), 0.0)
    g2i = gi * gi
    rhogt = rhoelg + rhops_g
    if hamil_rhog is not None:                                  # :473-487
        hrhogt = hamil_rhog / omega + rhops_g
        ehsum = float(np.sum((rhogt * np.conj(hrhogt)).real * g2i))
    else:
        hrhogt = rhogt
        ehsum = float(np.sum((rhogt.real ** 2 + rhogt.imag ** 2) * g2i))
    vfact = 1.0 if vb.real else 0.5
    ehart = vfact * omega * 4.0 * math.pi * ehsum
    vlocal_g = vion_g + 4.0 * math.pi * hrhogt * g2i            # :512-514
    tmp = offt.backward(vmaps, vlocal_g).ravel()               # :931
    return vxc + tmp.real, eps, ehart, exc
