"""Plane-wave basis (z-rods), FFT grid sizes and G<->grid maps -- oracle restatement.

Test infrastructure only (see ``oracle/__init__.py``).

Follows:
* ``src/qball/UnitCell.cc:36-90``   reciprocal vectors b = 2*pi*A^-T, volume = a0.(a1 x a2)
* ``src/qball/Basis.cc:126-154``    ``factorizable`` (2^h 3^i 5^j 7^k 11^m, i<=2, j,k,m<=1)
* ``src/qball/Basis.cc:302-700``    ``resize``: rods of |k+G|^2 < 2 ecut, Γ half sphere,
                                      rod order (multiset by size, stable), rod(0,0) first,
                                      np rule ``n = 2*hmax+2`` bumped to factorizable
* ``src/qball/Basis.cc:703-751``    ``update_g``: g, g2, kpg2, gx per G
* ``src/qball/ChargeDensity.cc:66-151`` density basis at 4*ecut, grid np(i)+2 bumped
* ``src/qball/FourierTransform.cc:186-205`` nvec / ntrans0
* ``src/qball/FourierTransform.cc:262-320`` ifftp / ifftm (Γ) and ``:440-460`` (complex)
"""
import math

import numpy as np


def cross(a, b):
    # math/d3vector.h:137-142, same evaluation order
    return np.array([a[1] * b[2] - a[2] * b[1],
                     a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def norm3(v):
    # math/d3vector.h:163-166
    return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]


class UnitCell:
    """UnitCell.cc:36-90.  ``a`` rows are the lattice vectors a0, a1, a2 (bohr)."""

    def __init__(self, a):
        a = np.asarray(a, dtype=np.float64).reshape(3, 3)
        self.a = a.copy()
        a0, a1, a2 = a
        self.volume = float(a0[0] * (a1[1] * a2[2] - a1[2] * a2[1])
                            + a0[1] * (a1[2] * a2[0] - a1[0] * a2[2])
                            + a0[2] * (a1[0] * a2[1] - a1[1] * a2[0]))
        fac = 1.0 / self.volume
        amt0 = cross(fac * a1, a2)
        amt1 = cross(fac * a2, a0)
        amt2 = cross(fac * a0, a1)
        self.b = np.array([2.0 * math.pi * amt0, 2.0 * math.pi * amt1,
                           2.0 * math.pi * amt2])

    @staticmethod
    def cubic(alat):
        return UnitCell(np.diag([alat, alat, alat]))


def factorizable(n):
    """Basis.cc:126-154 (ESSL lengths criterion used for every FFT backend)."""
    if n % 11 == 0:
        n //= 11
    if n % 7 == 0:
        n //= 7
    if n % 5 == 0:
        n //= 5
    if n % 3 == 0:
        n //= 3
    if n % 3 == 0:
        n //= 3
    while n % 2 == 0:
        n //= 2
    return n == 1


class Basis:
    """Basis::resize for a single process row (nprow = 1), Basis.cc:302-700.

    Attributes mirror the reference accessors: ``rod_h, rod_k, rod_lmin, rod_size``
    (rod order of the single process), ``idx`` (ngw x 3 Miller indices), ``np`` (grid
    sizes), ``idxmin/idxmax`` (used ranges), ``g2/kpg2/gx/kpgx``.
    """

    def __init__(self, cell, ecut, kpoint=(0.0, 0.0, 0.0), force_complex=False):
        self.cell = cell
        self.ecut = float(ecut)
        self.kpoint = np.asarray(kpoint, dtype=np.float64)
        # Basis.cc:284
        self.real = bool(np.all(self.kpoint == 0.0)) and not force_complex
        self._resize()

    def _resize(self):
        cell = self.cell
        two_ecut = 2.0 * self.ecut
        twopi = 2.0 * math.pi
        kpx, kpy, kpz = self.kpoint
        b0, b1, b2 = cell.b
        normb2 = norm3(b2)
        b2inv2 = 1.0 / normb2
        fac = math.sqrt(two_ecut) / twopi
        length = lambda v: math.sqrt(norm3(v))
        hmax = int(0.5 + fac * length(cell.a[0]))
        kmax = int(0.5 + fac * length(cell.a[1]))
        lmax = int(0.5 + fac * length(cell.a[2]))
        hmin, kmin, lmin = -hmax, -kmax, -lmax

        rods = []  # (h, k, lstart, size) in insertion order

        def scan(h, k, shift):
            lstart, lend, found = lmax, lmin, False
            for l in range(lmin - 1, lmax + 2):
                v = ((shift[0] + h) * b0 + (shift[1] + k) * b1) + (shift[2] + l) * b2
                if norm3(v) < two_ecut:
                    lstart = min(l, lstart)
                    lend = max(l, lend)
                    found = True
            return found, lstart, lend

        used = dict(h=[], k=[], l=[])
        if self.real:
            # Basis.cc:417-496
            lend0 = int(math.sqrt(two_ecut * b2inv2))
            rods.append((0, 0, 0, lend0 + 1))
            used['h'] += [0]
            used['k'] += [0]
            used['l'] += [0, lend0]
            zero = (0.0, 0.0, 0.0)
            for k in range(1, kmax + 2):
                found, ls, le = scan(0, k, zero)
                if found:
                    rods.append((0, k, ls, le - ls + 1))
                    used['k'] += [k]
                    used['l'] += [ls, le]
            for h in range(1, hmax + 2):
                for k in range(kmin - 1, kmax + 2):
                    found, ls, le = scan(h, k, zero)
                    if found:
                        rods.append((h, k, ls, le - ls + 1))
                        used['h'] += [h]
                        used['k'] += [k]
                        used['l'] += [ls, le]
        else:
            # Basis.cc:498-536
            for h in range(hmin - 1, hmax + 2):
                for k in range(kmin - 1, kmax + 2):
                    found, ls, le = scan(h, k, (kpx, kpy, kpz))
                    if found:
                        rods.append((h, k, ls, le - ls + 1))
                        used['h'] += [h]
                        used['k'] += [k]
                        used['l'] += [ls, le]

        # Basis.cc:549-557 (hmin_used etc. start at the opposite bound)
        self.idxmax = [max(used['h'] + [hmin]), max(used['k'] + [kmin]),
                       max(used['l'] + [lmin])]
        self.idxmin = [min(used['h'] + [hmax]), min(used['k'] + [kmax]),
                       min(used['l'] + [lmax])]
        self.hkl_max = (hmax, kmax, lmax)

        # Basis.cc:563-574: np rule
        npv = []
        for m in (hmax, kmax, lmax):
            n = 2 * m + 2
            while not factorizable(n):
                n += 2
            npv.append(n)
        self.np = npv

        # Basis.cc:579-663 with nprow = 1: multiset<Rod> ordered by size (equal sizes
        # keep insertion order), all rods land on node 0, then rod(0,0) is swapped to
        # the front.
        order = sorted(range(len(rods)), key=lambda i: rods[i][3])
        rl = [rods[i] for i in order]
        r0 = [i for i, r in enumerate(rl) if r[0] == 0 and r[1] == 0]
        if r0:
            j = r0[0]
            rl[0], rl[j] = rl[j], rl[0]
        self.rods = rl
        self.nrods = len(rl)
        self.rod_h = np.array([r[0] for r in rl], dtype=np.int32)
        self.rod_k = np.array([r[1] for r in rl], dtype=np.int32)
        self.rod_lmin = np.array([r[2] for r in rl], dtype=np.int32)
        self.rod_size = np.array([r[3] for r in rl], dtype=np.int32)
        self.rod_first = np.concatenate([[0], np.cumsum(self.rod_size)[:-1]]).astype(np.int64)
        self.size = int(self.rod_size.sum())

        # Basis.cc:680-692 idx; Basis.cc:703-751 update_g
        idx = np.empty((self.size, 3), dtype=np.int64)
        i = 0
        for (h, k, lm, sz) in rl:
            idx[i:i + sz, 0] = h
            idx[i:i + sz, 1] = k
            idx[i:i + sz, 2] = lm + np.arange(sz)
            i += sz
        self.idx = idx
        b = cell.b
        hh = idx[:, 0:1].astype(np.float64)
        kk = idx[:, 1:2].astype(np.float64)
        ll = idx[:, 2:3].astype(np.float64)
        gt = (hh * b[0] + kk * b[1]) + ll * b[2]
        kpgt = ((self.kpoint[0] + hh) * b[0] + (self.kpoint[1] + kk) * b[1]) \
            + (self.kpoint[2] + ll) * b[2]
        self.gx = gt                                  # ngw x 3
        self.kpgx = kpgt
        self.g2 = gt[:, 0] * gt[:, 0] + gt[:, 1] * gt[:, 1] + gt[:, 2] * gt[:, 2]
        self.kpg2 = kpgt[:, 0] * kpgt[:, 0] + kpgt[:, 1] * kpgt[:, 1] + kpgt[:, 2] * kpgt[:, 2]
        self.kpg = np.sqrt(self.kpg2)
        with np.errstate(divide='ignore'):
            self.kpgi = np.where(self.kpg > 0.0, 1.0 / np.where(self.kpg > 0, self.kpg, 1.0), 0.0)

    @property
    def ngw(self):
        return self.size


def density_grid(wf_basis_cell, ecut, ecutden=0.0):
    """ChargeDensity.cc:76-99: density basis at 4*ecut (or ecutden); grid np(i)+2 bumped
    by 2 until factorizable.  Returns (vbasis, (np0v, np1v, np2v))."""
    vb = Basis(wf_basis_cell, ecutden if ecutden > 0.0 else 4.0 * ecut)
    out = []
    for i in range(3):
        n = vb.np[i] + 2
        while not factorizable(n):
            n += 2
        out.append(n)
    return vb, tuple(out)


class FFTMaps:
    """Index maps of FourierTransform(basis, np0, np1, np2) for one process.

    FourierTransform.cc:186-205 (nvec, ntrans0), :262-320 (Γ ifftp/ifftm),
    :440-460 (complex ifftp), :362-382 / :473-494 (iunpack: column -> grid index
    h + np0*(k + np1*l)).

    ``zvec_to_grid`` gives, for each zvec column, the (x, y) grid position.
    """

    def __init__(self, basis, np0, np1, np2):
        self.basis = basis
        self.np0, self.np1, self.np2 = np0, np1, np2
        self.N = np0 * np1 * np2
        b = basis
        ngw = b.size
        if b.real:
            self.nvec = 2 * b.nrods - 1
            ifftp = np.empty(ngw, dtype=np.int64)
            ifftm = np.empty(ngw, dtype=np.int64)
            ig = 0
            ifftp[0] = 0
            ifftm[0] = 0
            ig = 1
            for l in range(1, int(b.rod_size[0])):
                ifftp[ig] = l
                ifftm[ig] = np2 - l
                ig += 1
            for irod in range(1, b.nrods):
                sz = int(b.rod_size[irod])
                l = np.arange(sz) + int(b.rod_lmin[irod])
                izp = np.where(l < 0, l + np2, l)
                izm = np.where(-l < 0, -l + np2, -l)
                ifftp[ig:ig + sz] = (2 * irod - 1) * np2 + izp
                ifftm[ig:ig + sz] = (2 * irod) * np2 + izm
                ig += sz
            self.ifftp, self.ifftm = ifftp, ifftm
            # column (x, y) positions: col 0 = rod(0,0); col 2r-1 = +rod r; 2r = -rod r
            colx = np.empty(self.nvec, dtype=np.int64)
            coly = np.empty(self.nvec, dtype=np.int64)
            colx[0] = 0
            coly[0] = 0
            for irod in range(1, b.nrods):
                hp, kp = int(b.rod_h[irod]), int(b.rod_k[irod])
                hp = hp + np0 if hp < 0 else hp
                kp = kp + np1 if kp < 0 else kp
                hm, km = -hp, -kp
                hm = hm + np0 if hm < 0 else hm
                km = km + np1 if km < 0 else km
                colx[2 * irod - 1], coly[2 * irod - 1] = hp, kp
                colx[2 * irod], coly[2 * irod] = hm, km
        else:
            self.nvec = b.nrods
            ifftp = np.empty(ngw, dtype=np.int64)
            ig = 0
            for irod in range(b.nrods):
                sz = int(b.rod_size[irod])
                l = np.arange(sz) + int(b.rod_lmin[irod])
                iz = np.where(l < 0, l + np2, l)
                ifftp[ig:ig + sz] = irod * np2 + iz
                ig += sz
            self.ifftp, self.ifftm = ifftp, None
            colx = np.where(b.rod_h < 0, b.rod_h + np0, b.rod_h).astype(np.int64)
            coly = np.where(b.rod_k < 0, b.rod_k + np1, b.rod_k).astype(np.int64)
        self.colx, self.coly = colx, coly
        # FourierTransform.cc:202
        self.ntrans0 = max(abs(b.idxmax[1]), abs(b.idxmin[1])) + 1

    def zvec_index_to_grid(self, izv):
        """Flat zvec index (col*np2 + z) -> flat grid index x + np0*(y + np1*z)."""
        col = izv // self.np2
        z = izv % self.np2
        return self.colx[col] + self.np0 * (self.coly[col] + self.np1 * z)
