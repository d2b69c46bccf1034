"""TDDFT observables -- oracle restatement.  Test infrastructure only.

Follows ``src/qball/CurrentDensity.cc:39-106`` (``update_current``) with the two-state
``SlaterDet::compute_density`` of ``src/qball/SlaterDet.cc:935-966``.
"""
import numpy as np

from . import fft as offt


def current(maps, c, occ, weight, omega, kpgx=None, A=None, nelectrons=0.0):
    """CurrentDensity::update_current for one spin and k-point.

    For each direction d: rwf = i (k+G)_d c (``:70-73``, the basis' kpgx, not k+G-A);
    tmp(r) = sum_n (weight/Omega) occ_n conj(psi_n(r)) rwf_n(r) over bands with positive
    factor (SlaterDet.cc:950-962); current_d(r) = -Im tmp(r) (``:79-81``);
    total_d = (Omega/N) sum_r current_d(r) (+ A_d * nelectrons when a vector potential is
    present, ``:94``).  Returns (current (3, N), total (3,))."""
    b = maps.basis
    kx = b.kpgx if kpgx is None else kpgx
    N = maps.N
    cur = np.zeros((3, N))
    prefac = weight / omega
    for n in range(c.shape[1]):
        fac = prefac * occ[n]
        if fac <= 0.0:
            continue
        f1 = offt.backward(maps, c[:, n]).ravel()
        for d in range(3):
            f2 = offt.backward(maps, 1j * kx[:, d] * c[:, n]).ravel()
            cur[d] += -(fac * np.conj(f1) * f2).imag
    total = (omega / N) * cur.sum(axis=1)
    if A is not None:
        total = total + np.asarray(A, dtype=np.float64) * nelectrons
    return cur, total
