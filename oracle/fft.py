"""G-sphere <-> grid transforms -- oracle restatement.  Test infrastructure only.

Follows ``src/qball/FourierTransform.cc``:
* ``backward(c, f)`` :529-540 -> ``vector_to_zvec`` :1624-1664 -> ``bwd`` :584-980
* ``backward(c1, c2, f)`` :555-567 -> ``doublevector_to_zvec`` :1684-1720 -> ``bwd``
* ``forward(f, c)`` :542-553 -> ``fwd`` :983-1361 -> ``zvec_to_vector`` :1666-1681
* ``forward(f, c1, c2)`` :570-581 -> ``fwd`` -> ``zvec_to_doublevector`` :1723-1752

Conventions fixed by the reference: backward is the unnormalised sum
f(r) = sum_G z(G) exp(+i G.r) (FFTW sign +1, ``init_lib`` :1607-1612); forward is
z(G) = (1/N) sum_r f(r) exp(-i G.r) (scale at :1338-1342).  Grid index is
x + np0*(y + np1*z) (FourierTransform.h:165-166).  The z / (x,y) pencil split of the
reference is a factorisation of the same 3-D DFT; here the 3-D DFT is evaluated with
scipy's FFT (pocketfft), which is pinned against a brute-force direct sum in
``tests/test_oracle_fft.py``.

Arrays: ``f`` is complex128 of shape (np2, np1, np0) (so ``f.ravel()`` is in the
reference's grid order); coefficient vectors are complex128 of length ngw.

The 3-D transforms use scipy's pocketfft with ``WORKERS`` threads (1 unless a caller
such as the benchmark's CPU baseline raises it); the thread count does not change the
result beyond rounding.
"""
import numpy as np
from scipy import fft as _sfft

WORKERS = 1


def _zvec_single(maps, c):
    """vector_to_zvec (FourierTransform.cc:1624-1664): zvec[ifftp]=c, zvec[ifftm]=conj(c)
    (Γ); the -G write comes second, so it wins at G=0."""
    z = np.zeros(maps.nvec * maps.np2, dtype=np.complex128)
    z[maps.ifftp] = c
    if maps.basis.real:
        z[maps.ifftm] = np.conj(c)
    return z


def _zvec_pair(maps, c1, c2):
    """doublevector_to_zvec (FourierTransform.cc:1684-1720):
    z[ip] = (a-d, b+c), z[im] = (a+d, c-b) for c1=a+ib, c2=c+id; -G write second."""
    a, b = c1.real, c1.imag
    c, d = c2.real, c2.imag
    z = np.zeros(maps.nvec * maps.np2, dtype=np.complex128)
    z[maps.ifftp] = (a - d) + 1j * (b + c)
    z[maps.ifftm] = (a + d) + 1j * (c - b)
    return z


def zvec_to_grid(maps, z):
    g = np.zeros(maps.N, dtype=np.complex128)
    g[maps.zvec_index_to_grid(np.arange(z.size))] = z
    return g.reshape(maps.np2, maps.np1, maps.np0)


def grid_to_zvec(maps, f):
    return f.ravel()[maps.zvec_index_to_grid(np.arange(maps.nvec * maps.np2))]


def bwd3(g):
    """Unnormalised exp(+iGr) 3-D transform of a (np2,np1,np0) array."""
    return _sfft.ifftn(g, workers=WORKERS) * g.size


def fwd3(f):
    """exp(-iGr) 3-D transform with 1/N."""
    return _sfft.fftn(f, workers=WORKERS) / f.size


def backward(maps, c1, c2=None):
    if c2 is None:
        z = _zvec_single(maps, c1)
    else:
        z = _zvec_pair(maps, c1, c2)
    return bwd3(zvec_to_grid(maps, z))


def forward(maps, f, pair=False):
    z = grid_to_zvec(maps, fwd3(f))
    ip = z[maps.ifftp]
    if not pair:
        return ip.copy()
    im = z[maps.ifftm]
    a, b = ip.real, ip.imag
    c, d = im.real, im.imag
    c1 = 0.5 * (a + c) + 1j * (0.5 * (b - d))
    c2 = 0.5 * (b + d) + 1j * (0.5 * (c - a))
    return c1, c2


# --------------------------------------------------------------------------------------
# Pruned, batched formulation (the reference's pencil order: z transforms of the occupied
# columns, y transforms of the occupied x-lines of each z plane, x transforms of the full
# planes; FourierTransform.cc:584-980 backward, :983-1361 forward).  The same 3-D DFT as
# bwd3 / fwd3, evaluated on a batch of transforms at once so that pocketfft's workers
# share the batch; tests/test_oracle_fft.py checks it against the unpruned path.  Used by
# the benchmark's CPU baseline.
class PrunedPlan:
    """Grids are held as (np0, np1, np2, B) -- x outermost, the batch innermost -- so each
    scatter and gather of the pruned passes moves whole contiguous blocks; ``backward`` /
    ``forward`` convert from / to the (B, np2, np1, np0) layout of the functions above."""

    def __init__(self, maps):
        self.maps = maps
        self.xs = np.unique(maps.colx)                    # occupied x positions
        self.col_lx = np.searchsorted(self.xs, maps.colx)  # column -> x-line index

    def _zvecs(self, c, pair):
        """zvec blocks (nvec, np2, B) of the coefficient columns c[:, bands]."""
        m = self.maps
        nb = c.shape[1]
        if pair:
            nt = (nb + 1) // 2
            c2 = np.zeros((c.shape[0], nt), dtype=np.complex128)
            c2[:, :nb // 2] = c[:, 1:2 * (nb // 2):2]
            c1 = c[:, 0::2]
            a, b = c1.real, c1.imag
            cc, d = c2.real, c2.imag
            zp = (a - d) + 1j * (b + cc)
            zm = (a + d) + 1j * (cc - b)
        else:
            nt = nb
            zp = c
            zm = np.conj(c) if m.basis.real else None
        z = np.zeros((m.nvec * m.np2, nt), dtype=np.complex128)
        z[m.ifftp] = zp
        if zm is not None:
            z[m.ifftm] = zm  # -G written second: it wins at G = 0
        return z.reshape(m.nvec, m.np2, nt)

    def bwd(self, c, pair):
        """(np0, np1, np2, B) grids of the bands in c (pairs when pair)."""
        m = self.maps
        z = _sfft.ifft(self._zvecs(c, pair), axis=1, workers=WORKERS, overwrite_x=True)
        nt = z.shape[2]
        q = np.zeros((m.np1, self.xs.size, m.np2, nt), dtype=np.complex128)
        q[m.coly, self.col_lx] = z
        q = _sfft.ifft(q, axis=0, workers=WORKERS, overwrite_x=True)
        f = np.zeros((m.np0, m.np1, m.np2, nt), dtype=np.complex128)
        f[self.xs] = q.transpose(1, 0, 2, 3)
        f = _sfft.ifft(f, axis=0, workers=WORKERS, overwrite_x=True)
        f *= m.N  # the three unnormalised exp(+iGr) sums
        return f

    def fwd(self, f, nb, pair):
        """Coefficients (ngw, nb) of the (np0, np1, np2, B) grids f; pairs when pair."""
        m = self.maps
        g = _sfft.fft(f, axis=0, workers=WORKERS)[self.xs].transpose(1, 0, 2, 3)
        g = _sfft.fft(g, axis=0, workers=WORKERS, overwrite_x=True)
        z = _sfft.fft(g[m.coly, self.col_lx], axis=1, workers=WORKERS, overwrite_x=True)
        z = z.reshape(m.nvec * m.np2, -1)
        ip = z[m.ifftp] / m.N
        if not pair:
            return ip
        im = z[m.ifftm] / m.N
        a, b = ip.real, ip.imag
        cc, d = im.real, im.imag
        out = np.empty((ip.shape[0], nb), dtype=np.complex128)
        out[:, 0::2] = (0.5 * (a + cc) + 1j * (0.5 * (b - d)))[:, :(nb + 1) // 2]
        out[:, 1::2] = (0.5 * (b + d) + 1j * (0.5 * (cc - a)))[:, :nb // 2]
        return out

    def backward(self, c, pair):
        """Grids (B, np2, np1, np0) of the bands in c."""
        return self.bwd(c, pair).transpose(3, 2, 1, 0)

    def forward(self, f, nb, pair):
        """Coefficients (ngw, nb) of grids f (B, np2, np1, np0)."""
        return self.fwd(np.ascontiguousarray(f.transpose(3, 2, 1, 0)), nb, pair)
