"""Shared helpers for the parity tests: host<->device transfer and oracle set-up."""
import numpy as np

from oracle.basis import Basis as OBasis, FFTMaps, UnitCell
from oracle.hpsi import Projectors


def to_dev(c):
    """complex (ngw, nst) -> device float64 tensor in band-major interleaved layout."""
    import torch
    arr = np.ascontiguousarray(np.asarray(c, dtype=np.complex128).T)
    return torch.from_numpy(arr.view(np.float64).ravel().copy()).cuda()


def from_dev(t, ngw, nst, ld=None):
    ld = ngw if ld is None else ld
    a = t.detach().cpu().numpy().view(np.complex128).reshape(nst, ld)[:, :ngw]
    return np.ascontiguousarray(a.T)


def real_to_dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).copy()).cuda()


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def max_rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (den if den > 0 else 1.0))


def oracle_for(sysm):
    """Oracle basis/maps/projectors for a workload.System (same cell, cutoff, k=0)."""
    ob = OBasis(UnitCell(sysm.a), sysm.ecut, force_complex=not sysm.real)
    maps = FFTMaps(ob, *sysm.np)
    sp = Projectors(sysm.lproj, sysm.wt, sysm.twnl, sysm.tau)
    return ob, maps, sp


def make_gpu_basis(ctx, sysm):
    from qball_amd import Basis
    r = sysm.rods
    return Basis(ctx, sysm.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np,
                 sysm.b)


def make_gpu_ham(ctx, basis, sysm, with_proj=True):
    from qball_amd import Hamiltonian
    h = Hamiltonian(ctx, basis)
    vd = real_to_dev(sysm.v)
    h.set_vloc(vd.data_ptr())
    h._v = vd
    if with_proj:
        h.set_projectors([sysm.na], [len(sysm.lproj)], sysm.lproj, sysm.wt, sysm.twnl.ravel(),
                         sysm.tau.ravel(), sysm.omega)
    return h


def synthetic_spdf():
    """Radial data of a KB species with s... f channels (lmax 3, local channel l = 0, so p,
    d and f projectors), smooth functions on a 0.01 mesh: the arguments of
    qball_amd.Species / oracle.species.Species."""
    n = 900
    r = np.arange(n) * 0.01
    zv = 3.0
    lmax, llocal = 3, 0
    vps, phi = [], []
    for l in range(lmax + 1):
        v = -zv * np.where(r > 0, 1.0, 0.0) * np.minimum(1.0 / np.maximum(r, 1e-9), 3.0)
        v = v + (1.5 - 0.2 * l) * np.exp(-(r / (1.0 + 0.1 * l)) ** 2)
        vps.append(v)
        phi.append((r ** (l + 1)) * np.exp(-0.8 * r * r))
    return {"zval": zv, "lmax": lmax, "llocal": llocal, "deltar": 0.01, "vps": vps,
            "phi": phi, "nquad": 0}
