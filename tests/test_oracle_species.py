"""Pins of the oracle's setup row (oracle/species.py) by what the mathematics fixes:

* the radial sine / cosine transforms against their defining sums (brute force);
* the Bessel transforms of each l against the closed form of r^l exp(-a r^2):
  int_0^inf r^(l+2) j_l(G r) exp(-a r^2) dr = sqrt(pi) G^l / (2^(l+2) a^(l+3/2)) exp(-G^2/4a);
* the cubic spline against scipy's CubicSpline with the same end conditions;
* the angular factors of twnl: sum_m Y_lm^2 = (2l+1) / 4 pi for every direction, and the
  real harmonics of each l orthonormal over the sphere (Gauss-Legendre x trapezoid grid);
* the Ewald split: esr + the G-space energy of the Gaussian pseudo-charges - eself
  - pi rcps^2 (sum z)^2 / Omega equals the Madelung energy with a neutralising background
  (simple cubic -1.4186487 z^2/a, fcc -2.2924 z^2/a per ion), whatever rcps;
* the Si HSCV LDA file (the reference's share/pseudopotentials, held under tests/golden)
  read with its published sizes, and Vloc(G) -> -4 pi zval / G^2 at small G (the long-range
  tail the Gaussian subtraction removes is that of a point charge).
"""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import erfc

from oracle import species as S

HERE = os.path.dirname(os.path.abspath(__file__))
SI_XML = os.path.join(HERE, "golden", "Si_HSCV_LDA-1.0.xml")


@pytest.fixture(scope="module")
def si():
    return S.Species(S.read_species_xml(SI_XML), 1.5)


def test_sine_and_cosine_transforms_are_their_sums():
    rng = np.random.default_rng(1)
    n = 64
    y = rng.standard_normal(n)
    j = np.arange(n)
    F = np.array([np.sum(y[1:] * np.sin(np.pi * j[1:] * k / n)) for k in range(n)])
    assert np.max(np.abs(S.sinft(y) - F)) < 1e-12
    yc = rng.standard_normal(n + 1)
    C = np.array([0.5 * (yc[0] + (-1) ** k * yc[n]) +
                  np.sum(yc[1:n] * np.cos(np.pi * j[1:] * k / n)) for k in range(n + 1)])
    assert np.max(np.abs(S.cosft1(yc) - C)) < 1e-12


def test_simpsn_integrates_a_cubic_exactly():
    x = np.linspace(0.0, 2.0, 201)
    h = x[1] - x[0]
    f = 1.0 + x - 3.0 * x * x + 0.5 * x ** 3
    exact = 2.0 + 2.0 - 8.0 + 0.5 * 16.0 / 4.0
    assert abs(S.simpsn(f) * h - exact) < 1e-12


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_bessel_transform_of_gaussian_closed_form(si, l):
    a = 0.7
    fnr = 4.0 * math.pi * si.deltar * si.rps ** l * np.exp(-a * si.rps ** 2)
    fng = si.bessel_trans(l, fnr)[:si.ndft]
    G = si.gspl
    exact = (4.0 * math.pi * math.sqrt(math.pi) * G ** l / (2 ** (l + 2) * a ** (l + 1.5))
             * np.exp(-G * G / (4.0 * a)))
    m = G < 12.0
    assert np.max(np.abs(fng[m] - exact[m])) < 1e-9 * np.max(np.abs(exact))


@pytest.mark.parametrize("natural0", [False, True])
def test_spline_matches_scipy_with_the_same_end_conditions(natural0):
    x = np.linspace(0.0, 3.0, 41)
    y = np.cos(2.0 * x) * np.exp(-x)
    sp = S.Spline(x, y, None if natural0 else 0.0)
    cs = CubicSpline(x, y, bc_type=((2, 0.0) if natural0 else (1, 0.0), (2, 0.0)))
    q = np.linspace(0.0, 3.0, 997)
    v, dv = sp.derivative(q)
    assert np.max(np.abs(v - cs(q))) < 1e-13
    assert np.max(np.abs(dv - cs(q, 1))) < 1e-12


def test_si_pseudopotential_file_and_tables(si):
    sp = S.read_species_xml(SI_XML)
    assert sp["zval"] == 4.0 and sp["lmax"] == 2 and sp["llocal"] == 2
    assert sp["deltar"] == 0.01 and all(v.size == 956 for v in sp["vps"])
    assert si.ndft == 4096 and si.projector_ls() == [0, 1, 1, 1]
    assert si.wsg[0] > 0 and si.wsg[1] > 0 and si.wsg[2] == 0.0
    # Vloc(G) of the full local potential tends to that of a point charge -4 pi z / G^2 at
    # small G: the smeared-charge part removed is -4 pi z exp(-rcps^2 G^2 / 4) / G^2
    g = np.array([0.05, 0.1, 0.2])
    full = si.vlocg(g) - 4.0 * math.pi * si.zval * np.exp(-0.25 * si.rcps ** 2 * g * g) / (g * g)
    assert np.all(np.abs(full * g * g / (-4.0 * math.pi * si.zval) - 1.0) < 0.03)


@pytest.mark.parametrize("lmax", [1, 2, 3])
def test_twnl_angular_sum_rule_and_orthonormality(lmax):
    """With v(|k+G|) = 1 the rows of each l are the real harmonics of the reference's
    order: sum_m Y_lm^2 = (2l+1) / 4 pi, and they are orthonormal over the sphere."""
    class Flat:
        llocal = -1

        def __init__(self, lmax):
            self.lmax = lmax
            self.wsg = np.ones(lmax + 1)

        def dvnlg(self, l, q):
            return np.ones_like(q), np.zeros_like(q)
    # quadrature grid on the unit sphere: Gauss-Legendre in cos(theta), trapezoid in phi
    xg, wg = np.polynomial.legendre.leggauss(24)
    ph = np.arange(48) * (2.0 * np.pi / 48)
    ct, pp = np.meshgrid(xg, ph, indexing="ij")
    w = (np.outer(wg, np.full(48, 2.0 * np.pi / 48))).ravel()
    st = np.sqrt(1.0 - ct * ct)
    dirs = np.stack([(st * np.cos(pp)).ravel(), (st * np.sin(pp)).ravel(), ct.ravel()], axis=1)
    rows, ls, _ = S.twnl_table(Flat(lmax), 2.5 * dirs)
    for l in range(lmax + 1):
        Y = rows[ls == l]
        assert Y.shape[0] == 2 * l + 1
        assert np.max(np.abs(np.sum(Y * Y, axis=0) - (2 * l + 1) / (4.0 * np.pi))) < 1e-12
        gram = (Y * w[None, :]) @ Y.T
        assert np.max(np.abs(gram - np.eye(2 * l + 1))) < 1e-12


def _ewald_total(a, zv, rcps, tau, nmax=18):
    a = np.asarray(a, dtype=np.float64)
    om = abs(np.linalg.det(a))
    b = 2.0 * np.pi * np.linalg.inv(a).T
    r = np.arange(-nmax, nmax + 1)
    n = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    G = n @ b
    G2 = np.sum(G * G, axis=1)
    keep = G2 > 0
    G, G2 = G[keep], G2[keep]
    rho = np.zeros(G.shape[0], dtype=np.complex128)
    for z, rc, t in zip(zv, rcps, tau):
        rho += -z * np.exp(-0.25 * rc * rc * G2) / om * S.structure_factor(t, G)
    eg = 0.5 * om * np.sum(4.0 * np.pi / G2 * np.abs(rho) ** 2)
    es = S.esr(a, zv, rcps, tau)
    eself = sum(z * z / (math.sqrt(2.0 * math.pi) * rc) * len(t) for z, rc, t in zip(zv, rcps, tau))
    ztot = sum(z * len(t) for z, t in zip(zv, tau))
    # G = 0 term of the Gaussian split for a neutral cell (one rcps for all ions)
    e0 = -math.pi * rcps[0] ** 2 * ztot * ztot / om
    return eg + es - eself + e0, es


@pytest.mark.parametrize("rc", [0.8, 1.2, 1.5])
def test_ewald_split_gives_madelung_energies(rc):
    """The reference's esr sums pairs of distinct atoms only (an atom and its own images
    are left out, EnergyFunctional.cc:1937-1943), which is exact while erfc(a / rckj) is
    negligible: widths up to 1.5 bohr in a 10-bohr cell (erfc(4.7) ~ 1e-11)."""
    a = 10.0
    e_sc, _ = _ewald_total(np.diag([a] * 3), [1.0], [rc], [np.zeros((1, 3))])
    assert abs(e_sc - (-1.4186487 / a)) < 1e-7
    tau = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    e_fcc, es = _ewald_total(np.diag([a] * 3), [1.0], [rc], [tau])
    assert abs(e_fcc / 4 * a - (-2.2924)) < 1e-4
    if rc == 1.5:
        assert es > 1e-6  # the real-space pair sum contributes at this width


def test_fold_in_ws_and_min_wsdist():
    a = np.array([[0.0, 5.0, 5.0], [5.0, 0.0, 5.0], [5.0, 5.0, 0.0]])  # fcc primitive
    assert abs(S.min_wsdist(a) - math.sqrt(50.0)) < 1e-12
    rng = np.random.default_rng(3)
    an, an2h = S.ws_planes(a)
    for _ in range(50):
        v = rng.uniform(-20, 20, 3)
        f = S.fold_in_ws(v, a)
        assert np.all(np.abs(an @ f) <= an2h + 1e-9)
        frac = np.linalg.solve(a.T, f - v)
        assert np.max(np.abs(frac - np.round(frac))) < 1e-9  # a lattice translation


def test_twnl_row_order_follows_the_reference():
    """Along (1,1,1) the rows that vanish fix the order of NonLocalPotential.cc's listing:
    d: y4 (3z^2 - 1) and y5 (x^2 - y^2) vanish, y6..y8 do not; f: y13 (z (x^2 - y^2))
    y13 vanishes, y12 (xyz) does not; along z only y4 and y9 survive."""
    class Flat:
        llocal = -1
        lmax = 3
        wsg = np.ones(4)

        def dvnlg(self, l, q):
            return np.ones_like(q), np.zeros_like(q)
    d = np.array([[1.0, 1.0, 1.0]]) * 1.7
    rows, ls, _ = S.twnl_table(Flat(), d)
    y2 = rows[ls == 2][:, 0]
    y3 = rows[ls == 3][:, 0]
    assert abs(y2[0]) < 1e-15 and abs(y2[1]) < 1e-15 and np.all(np.abs(y2[2:]) > 0.1)
    assert abs(y3[4]) < 1e-15 and abs(y3[3]) > 0.1
    # and along z only y4 (d) and y9 (f) survive
    rows, ls, _ = S.twnl_table(Flat(), np.array([[0.0, 0.0, 2.0]]))
    assert np.flatnonzero(np.abs(rows[ls == 2][:, 0]) > 1e-15).tolist() == [0]
    assert np.flatnonzero(np.abs(rows[ls == 3][:, 0]) > 1e-15).tolist() == [0]


@pytest.mark.parametrize("a,zv,ecut", [(np.diag([41.04] * 3), 4.0, 12.5),
                                        (np.diag([10.26] * 3), 4.0, 12.5),
                                        (np.diag([6.0] * 3), 3.0, 30.0),
                                        (np.array([[0.0, 2.0, 2.0], [2.0, 0.0, 2.0],
                                                   [2.0, 2.0, 0.0]]), 4.0, 40.0)])
def test_auto_rcps_is_the_first_width_meeting_both_thresholds(a, zv, ecut):
    """AtomSet::set_rcps: the reciprocal tail test holds at the result and fails one step
    below it (unless the cell caps the width), and the enclosed-charge test holds."""
    r = S.auto_rcps(zv, a, ecut)
    ell = S.min_wsdist(a)
    tail = lambda x: (zv / math.sqrt(math.pi)) * erfc(x * math.sqrt(ecut)) - 1e-7 * x / zv
    from scipy.special import erfc as erfc_
    assert zv * erfc_(ell / (math.sqrt(2.0) * r)) <= 1e-8
    if r < ell / (2.0 * math.pi) - 1e-12 and abs((r - 0.3) / 0.05 - round((r - 0.3) / 0.05)) < 1e-9:
        assert tail(r) <= 0.0 and tail(r - 0.05) > 0.0
