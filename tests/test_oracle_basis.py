"""Oracle plane-wave basis: pinned by the reference's measured counts and brute force."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle.basis import Basis, FFTMaps, UnitCell, density_grid, factorizable

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "basis_counts.json")))


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_reference_counts(case):
    cell = UnitCell.cubic(case["alat"])
    ecut = 0.5 * case["ecut_ry"]
    b = Basis(cell, ecut)
    vb, npd = density_grid(cell, ecut)
    maps = FFTMaps(b, *npd)
    assert b.size == case["ngw"]
    assert b.nrods == case["nrods"]
    assert maps.nvec == case["nvec"]
    assert maps.ntrans0 == case["ntrans0"]
    assert list(npd) == case["np_density"]
    if "np_wf" in case:
        assert b.np == case["np_wf"]
        assert vb.np == case["np_vbasis"]
    if "ngw_density" in case:
        assert vb.size == case["ngw_density"]


def test_factorizable_brute_force():
    ok = set()
    for h in range(0, 12):
        for i in range(0, 3):
            for j in range(0, 2):
                for k in range(0, 2):
                    for m in range(0, 2):
                        n = 2 ** h * 3 ** i * 5 ** j * 7 ** k * 11 ** m
                        if n <= 4000:
                            ok.add(n)
    for n in range(1, 4001):
        assert factorizable(n) == (n in ok), n


def _tri_cell():
    return UnitCell(np.array([[7.1, 0.3, -0.2], [0.9, 6.4, 0.5], [-0.4, 1.1, 8.3]]))


@pytest.mark.parametrize("real", [True, False])
def test_sphere_matches_enumeration(real):
    """Rods cover exactly {G : |G|^2 < 2 ecut} (Γ: the half space h>0, or h=0,k>0, or
    h=k=0,l>=0), each rod contiguous in l, rod(0,0) first, sizes non-decreasing after."""
    cell = _tri_cell()
    ecut = 9.0
    b = Basis(cell, ecut, force_complex=not real)
    B = cell.b
    lim = 40
    want = set()
    for h, k, l in itertools.product(range(-lim, lim + 1), repeat=3):
        g = h * B[0] + k * B[1] + l * B[2]
        if g @ g < 2 * ecut:
            if real and not (h > 0 or (h == 0 and k > 0) or (h == 0 and k == 0 and l >= 0)):
                continue
            want.add((h, k, l))
    got = [tuple(int(x) for x in row) for row in b.idx]
    assert len(got) == len(set(got))
    assert set(got) == want
    assert b.rod_h[0] == 0 and b.rod_k[0] == 0
    # ascending sizes, except for the rod that rod(0,0) was swapped with
    s = b.rod_size.copy()
    bad = np.nonzero(np.diff(s[1:]) < 0)[0]
    assert len(bad) <= 1
    if len(bad):
        s = np.delete(s, bad[0] + 2)
    assert np.all(np.diff(s[1:]) >= 0)


def test_fft_maps_are_a_bijection():
    cell = UnitCell.cubic(10.26)
    b = Basis(cell, 6.0)
    m = FFTMaps(b, 28, 28, 28)
    gp = m.zvec_index_to_grid(m.ifftp)
    gm = m.zvec_index_to_grid(m.ifftm)
    # +G and -G images land on the grid points of h, k, l and -h, -k, -l
    hkl = b.idx % 28
    want_p = hkl[:, 0] + 28 * (hkl[:, 1] + 28 * hkl[:, 2])
    mhkl = (-b.idx) % 28
    want_m = mhkl[:, 0] + 28 * (mhkl[:, 1] + 28 * mhkl[:, 2])
    assert np.array_equal(gp, want_p)
    assert np.array_equal(gm, want_m)
    allpts = np.concatenate([gp, gm[1:]])
    assert len(np.unique(allpts)) == len(allpts)
