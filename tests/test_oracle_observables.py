"""Oracle current density pinned by Parseval, plane waves and Γ realness."""
import numpy as np
import pytest

from oracle import hpsi as oh
from oracle.basis import Basis, FFTMaps, UnitCell
from oracle.observables import current


def _setup(real, nst=3, seed=2):
    cell = UnitCell(np.array([[6.0, 0.2, 0.0], [0.0, 5.5, 0.4], [0.3, 0.0, 6.5]]))
    b = Basis(cell, 4.0, force_complex=not real)
    maps = FFTMaps(b, *b.np)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((b.size, nst)) + 1j * rng.standard_normal((b.size, nst))
    if real:
        c[0] = c[0].real
    return cell, b, maps, oh.gram(c, real)


def test_total_current_is_the_g_space_sum():
    cell, b, maps, c = _setup(False)
    occ = np.array([2.0, 1.5, 0.0])
    cur, tot = current(maps, c, occ, 1.0, cell.volume)
    # Parseval: (Omega/N) sum_r conj(psi) (i K psi) = i Omega sum_G K |c|^2
    g = -(np.abs(c) ** 2 * occ[None, :]).sum(axis=1) @ b.kpgx
    assert np.max(np.abs(tot - g)) <= 1e-12 * np.max(np.abs(g))
    A = np.array([0.1, -0.2, 0.3])
    _, tot2 = current(maps, c, occ, 1.0, cell.volume, A=A, nelectrons=3.5)
    assert np.max(np.abs(tot2 - tot - 3.5 * A)) <= 1e-14


def test_single_plane_wave_current_is_uniform():
    cell, b, maps, _ = _setup(False)
    ig = 7
    c = np.zeros((b.size, 1), complex)
    c[ig, 0] = 1.0
    cur, tot = current(maps, c, np.array([2.0]), 1.0, cell.volume)
    expect = -2.0 / cell.volume * b.kpgx[ig]
    assert np.max(np.abs(cur - expect[:, None])) <= 1e-13
    assert np.max(np.abs(tot + 2.0 * b.kpgx[ig])) <= 1e-12


def test_gamma_current_vanishes():
    cell, b, maps, c = _setup(True)
    cur, tot = current(maps, c, np.array([2.0, 2.0, 2.0]), 1.0, cell.volume)
    assert np.max(np.abs(cur)) <= 1e-13 and np.max(np.abs(tot)) <= 1e-13
