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
numpy's FFT (pocketfft), which is pinned against a brute-force direct sum in
``tests/test_oracle_fft.py``.

Arrays: ``f`` is complex128 of shape (np2, np1, np0) (so ``f.ravel()`` is in the
reference's grid order); coefficient vectors are complex128 of length ngw.
"""
import numpy as np


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
    return np.fft.ifftn(g) * g.size


def fwd3(f):
    """exp(-iGr) 3-D transform with 1/N."""
    return np.fft.fftn(f) / f.size


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
