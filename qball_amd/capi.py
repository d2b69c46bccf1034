"""ctypes binding of the qbh C ABI (include/qbh.h).

Device buffers are passed as raw addresses (e.g. ``tensor.data_ptr()`` of a torch tensor
on ``cuda:0``); torch is only the allocator and stream provider here.  No function in
this module computes anything itself: every numerical result comes from libqbh.so.
"""
import ctypes
import os

import numpy as np

QBH_NONLOCAL = 0x1
QBH_KINETIC = 0x2
QBH_LOCAL = 0x4
QBH_ALL = 0x7
QBH_ACCUMULATE = 0x8
QBH_FP32 = 0x10

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqbh.so")

# exported symbol -> (restype, argtypes)
_c_int, _c_long, _c_double, _vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "qbh_ctx_create": (_c_int, [_c_int, _vp, ctypes.POINTER(_vp)]),
    "qbh_ctx_destroy": (_c_int, [_vp]),
    "qbh_last_error": (ctypes.c_char_p, [_vp]),
    "qbh_sync": (_c_int, [_vp]),
    "qbh_ctx_set_workspace_cap": (_c_int, [_vp, _c_double]),
    "qbh_ctx_set_timing": (_c_int, [_vp, _c_int]),
    "qbh_ctx_set_overlap": (_c_int, [_vp, _c_int]),
    "qbh_ctx_timing_report": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "qbh_ctx_library_info": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "qbh_ctx_stream_bandwidth": (_c_int, [_vp, _c_double, _c_int, _dp]),
    "qbh_basis_create": (_c_int, [_vp, _c_int, _c_int, _ip, _ip, _ip, _ip, _c_int, _c_int,
                                  _c_int, _dp, _dp, ctypes.POINTER(_vp)]),
    "qbh_basis_destroy": (_c_int, [_vp]),
    "qbh_basis_info": (_c_int, [_vp, _ip, _ip, _ip]),
    "qbh_rods_build": (_c_int, [_dp, _c_double, _c_int, _dp, _c_int, _ip, _ip, _ip, _ip, _ip,
                                _ip]),
    "qbh_factorizable": (_c_int, [_c_int]),
    "qbh_fft_backward": (_c_int, [_vp, _vp, _c_long, _c_int, _c_int, _vp, _c_int]),
    "qbh_fft_backward_noalloc": (_c_int, [_vp, _vp, _c_long, _c_int, _c_int, _vp, _c_int]),
    "qbh_fft_forward": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_long, _c_int]),
    "qbh_fft_forward_real": (_c_int, [_vp, _vp, _c_double, _vp, _c_int]),
    "qbh_ham_create": (_c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "qbh_ham_destroy": (_c_int, [_vp]),
    "qbh_ham_set_vloc": (_c_int, [_vp, _vp]),
    "qbh_ham_set_projectors": (_c_int, [_vp, _c_int, _ip, _ip, _ip, _dp, _dp, _dp, _c_double]),
    "qbh_ham_set_vector_potential": (_c_int, [_vp, _dp]),
    "qbh_hpsi": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _vp, _c_long, _c_int, _dp]),
    "qbh_nl_forces": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _dp, _c_int, _dp]),
    "qbh_density": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _c_double, _c_double, _vp, _c_int]),
    "qbh_gram": (_c_int, [_vp, _vp, _c_long, _c_int]),
    "qbh_subspace_diag": (_c_int, [_vp, _vp, _c_long, _vp, _c_long, _c_int, _dp, _c_int]),
    "qbh_update_vhxc": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _c_double, _vp, _dp]),
    "qbh_exponential": (_c_int, [_vp, _vp, _c_long, _c_int, _c_double, _c_int, _c_int]),
    "qbh_current": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _c_double, _c_double, _dp, _c_double,
                             _vp, _dp]),
    "qbh_wf_write_raw": (_c_int, [ctypes.c_char_p, _c_int, _vp, _c_long, _c_int, _dp, _dp]),
    "qbh_wf_read_raw": (_c_int, [ctypes.c_char_p, _c_long, _vp, _c_long, _c_int, _dp, _dp,
                                 ctypes.POINTER(_c_long)]),
    "qbh_sd_write_xml": (_c_int, [_vp, ctypes.c_char_p, _c_int, _vp, _c_long, _c_int, _dp,
                                  _c_double, _c_int, _c_int]),
    "qbh_sd_read_xml": (_c_int, [_vp, ctypes.c_char_p, _c_long, _vp, _c_long, _c_int, _dp,
                                 ctypes.POINTER(_c_long)]),
    "qbh_psda_create": (_c_int, [_vp, _c_int, _c_double, ctypes.POINTER(_vp)]),
    "qbh_psda_destroy": (_c_int, [_vp]),
    "qbh_psda_preprocess": (_c_int, [_vp]),
    "qbh_psda_update": (_c_int, [_vp, _vp, _c_long, _vp, _c_long, _dp, _dp]),
    "qbh_kinetic_energy": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _dp]),
    "qbh_mixer_create": (_c_int, [_vp, _c_long, _c_int, ctypes.POINTER(_vp)]),
    "qbh_mixer_destroy": (_c_int, [_vp]),
    "qbh_mixer_restart": (_c_int, [_vp]),
    "qbh_mixer_update": (_c_int, [_vp, _vp, _vp, _vp, _vp, _dp]),
    "qbh_scf_run": (_c_int, [_vp, _vp, _vp, _c_long, _c_int, _dp, _c_double, _vp, _vp, _c_int,
                             _c_int, _c_double, _c_double, _c_int, _c_double, _c_int, _c_int,
                             _dp]),
    "qbh_scf_run_smeared": (_c_int, [_vp, _vp, _vp, _c_long, _c_int, _dp, _c_double, _c_double,
                                     _c_double, _vp, _vp, _c_int, _c_int, _c_double, _c_double,
                                     _c_int, _c_double, _c_int, _c_int, _dp, _dp, _dp]),
    "qbh_ham_set_tuning": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "qbh_ham_get_tuning": (_c_int, [_vp, ctypes.c_char_p, _ip]),
    "qbh_ham_autotune": (_c_int, [_vp, _vp, _c_long, _c_int, _dp, _vp, _c_long, _c_int, _c_int,
                                  _c_int, _c_int, ctypes.c_char_p, _c_int]),
    "qbh_batch_create": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                  ctypes.POINTER(_c_long), _ip, ctypes.POINTER(_vp),
                                  ctypes.POINTER(_c_long), _c_int, _c_int, ctypes.POINTER(_vp)]),
    "qbh_batch_launch": (_c_int, [_vp]),
    "qbh_species_create": (_c_int, [_c_double, _c_int, _c_int, _c_double, _c_int, _dp, _dp,
                                    _c_double, ctypes.POINTER(_vp)]),
    "qbh_species_create_xml": (_c_int, [ctypes.c_char_p, _c_double, ctypes.POINTER(_vp)]),
    "qbh_species_destroy": (_c_int, [_vp]),
    "qbh_species_auto_rcps": (_c_int, [_c_double, _dp, _c_double, _dp]),
    "qbh_species_info": (_c_int, [_vp, _dp, _dp, _dp, _ip, _ip, _dp]),
    "qbh_species_eval": (_c_int, [_vp, _c_int, _c_int, _dp, _dp, _dp]),
    "qbh_species_twnl": (_c_int, [_vp, _vp, _dp, _vp]),
    "qbh_species_dtwnl": (_c_int, [_vp, _vp, _dp, _vp]),
    "qbh_ham_set_species": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), _ip, _dp, _c_double]),
    "qbh_local_tables": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), _ip, _vp, _c_double, _vp,
                                  _vp]),
    "qbh_ewald_esr": (_c_int, [_dp, _c_int, ctypes.POINTER(_vp), _ip, _dp, _dp]),
    "qbh_batch_destroy": (_c_int, [_vp]),
    "qbh_td_create": (_c_int, [_vp, _vp, _c_int, _dp, _c_double, _c_double, _vp, _vp,
                               ctypes.POINTER(_vp)]),
    "qbh_td_destroy": (_c_int, [_vp]),
    "qbh_td_forward": (_c_int, [_vp, _vp, _c_long, _c_int, _c_double, _c_int, _dp, _dp, _dp, _vp,
                                _c_int, _c_int, _dp, _vp, _c_long]),
    "qbh_td_backward": (_c_int, [_vp, _vp, _dp]),
    "qbh_etrs_step": (_c_int, [_vp, _vp, _vp, _c_long, _c_int, _dp, _c_double, _c_double, _c_int,
                               _vp, _vp, _c_double, _dp, _c_int]),
}

_lib = None


class QbhError(RuntimeError):
    pass


def lib():
    """Load libqbh.so (raises if it has not been built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QbhError(f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ia(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _da(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def factorizable(n):
    return bool(lib().qbh_factorizable(int(n)))


def rods_build(a, ecut, real=True, kpoint=(0.0, 0.0, 0.0)):
    """Basis::resize restated in the library (qbh_rods_build).  Returns dict with
    rod_h, rod_k, rod_lmin, rod_size (int32 arrays) and np (3 ints)."""
    L = lib()
    a_, ap = _da(np.asarray(a, dtype=np.float64).reshape(9))
    k_, kp = _da(np.asarray(kpoint, dtype=np.float64))
    n = ctypes.c_int(0)
    npv = (ctypes.c_int * 3)()
    rc = L.qbh_rods_build(ap, float(ecut), int(real), kp, 0, ctypes.byref(n), None, None, None,
                          None, npv)
    if rc != 0:
        raise QbhError(f"qbh_rods_build failed ({rc})")
    arrs = [np.zeros(n.value, dtype=np.int32) for _ in range(4)]
    ptrs = [x.ctypes.data_as(_ip) for x in arrs]
    rc = L.qbh_rods_build(ap, float(ecut), int(real), kp, n.value, ctypes.byref(n), *ptrs, npv)
    if rc != 0:
        raise QbhError(f"qbh_rods_build failed ({rc})")
    return dict(rod_h=arrs[0], rod_k=arrs[1], rod_lmin=arrs[2], rod_size=arrs[3],
                np=[npv[0], npv[1], npv[2]])


class Context:
    def __init__(self, device=0, stream=None):
        self.h = _vp()
        rc = lib().qbh_ctx_create(int(device), _vp(stream) if stream else None,
                                  ctypes.byref(self.h))
        if rc != 0:
            raise QbhError(f"qbh_ctx_create failed ({rc})")

    def check(self, rc, what):
        if rc != 0:
            msg = lib().qbh_last_error(self.h)
            raise QbhError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def sync(self):
        self.check(lib().qbh_sync(self.h), "qbh_sync")

    def set_workspace_cap(self, nbytes):
        self.check(lib().qbh_ctx_set_workspace_cap(self.h, float(nbytes)), "workspace cap")

    def set_overlap(self, on):
        self.check(lib().qbh_ctx_set_overlap(self.h, int(bool(on))), "set_overlap")

    def set_timing(self, on):
        self.check(lib().qbh_ctx_set_timing(self.h, int(bool(on))), "set_timing")

    def timing_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        lib().qbh_ctx_timing_report(self.h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, cnt = line.split()
            out[name] = (float(ms), int(cnt))
        return out

    def library_info(self):
        """HIP runtime, rocBLAS and hipBLASLt versions loaded in this process, and the
        library the built-in projector GEMM choice uses with them."""
        buf = ctypes.create_string_buffer(512)
        self.check(lib().qbh_ctx_library_info(self.h, buf, len(buf)), "library_info")
        return buf.value.decode()

    def stream_bandwidth(self, gib=2.0, reps=3):
        """GB/s of grid-stride read-only, write-only and copy (read + written bytes) kernels
        over `gib` GiB buffers on this context's device."""
        out = (_c_double * 3)()
        self.check(lib().qbh_ctx_stream_bandwidth(self.h, float(gib) * (1 << 30), int(reps), out),
                   "stream_bandwidth")
        return {"read": out[0], "write": out[1], "copy": out[2]}

    def close(self):
        if self.h:
            lib().qbh_ctx_destroy(self.h)
            self.h = _vp()


class Basis:
    def __init__(self, ctx, real, rod_h, rod_k, rod_lmin, rod_size, np_grid, b, kpoint=(0, 0, 0)):
        self.ctx = ctx
        self.h = _vp()
        self._keep = [_ia(rod_h), _ia(rod_k), _ia(rod_lmin), _ia(rod_size)]
        b_, bp = _da(np.asarray(b, dtype=np.float64).reshape(9))
        k_, kp = _da(np.asarray(kpoint, dtype=np.float64))
        self.np = tuple(int(x) for x in np_grid)
        rc = lib().qbh_basis_create(ctx.h, int(real), len(self._keep[0][0]),
                                    *[k[1] for k in self._keep], *self.np, bp, kp,
                                    ctypes.byref(self.h))
        ctx.check(rc, "qbh_basis_create")
        ngw, ncol, nlx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().qbh_basis_info(self.h, ctypes.byref(ngw), ctypes.byref(ncol), ctypes.byref(nlx))
        self.ngw, self.ncol, self.nlx = ngw.value, ncol.value, nlx.value
        self.real = bool(real)
        self.N = self.np[0] * self.np[1] * self.np[2]

    def backward(self, c_ptr, ld, nbands, pair, f_ptr, flags=0):
        self.ctx.check(lib().qbh_fft_backward(self.h, c_ptr, ld, nbands, int(pair), f_ptr, flags),
                       "qbh_fft_backward")

    def backward_noalloc(self, c_ptr, ld, nbands, pair, f_ptr, flags=0):
        """backward() through the entry point that allocates nothing (qbh_fft_backward_noalloc)."""
        self.ctx.check(lib().qbh_fft_backward_noalloc(self.h, c_ptr, ld, nbands, int(pair), f_ptr,
                                                      flags), "qbh_fft_backward_noalloc")

    def forward(self, f_ptr, nbands, pair, c_ptr, ld, flags=0):
        self.ctx.check(lib().qbh_fft_forward(self.h, f_ptr, nbands, int(pair), c_ptr, ld, flags),
                       "qbh_fft_forward")

    def forward_real(self, rho_ptr, scale, c_ptr, flags=0):
        self.ctx.check(lib().qbh_fft_forward_real(self.h, rho_ptr, float(scale), c_ptr, flags),
                       "qbh_fft_forward_real")

    def density(self, psi_ptr, ld, nst, occ, weight, omega, rho_ptr, flags=0):
        o_, op = _da(occ)
        self.ctx.check(lib().qbh_density(self.h, psi_ptr, ld, nst, op, float(weight), float(omega),
                                         rho_ptr, flags), "qbh_density")

    def update_vhxc(self, rhor_ptr, rhog_ptr, vion_g_ptr, rhops_g_ptr, omega, v_r_ptr,
                    hamil_rhog_ptr=None):
        """EnergyFunctional::update_vhxc on this (density) basis; returns (eps, ehart, exc)."""
        e = np.zeros(3)
        self.ctx.check(lib().qbh_update_vhxc(self.h, rhor_ptr, rhog_ptr, hamil_rhog_ptr, vion_g_ptr,
                                             rhops_g_ptr, float(omega), v_r_ptr,
                                             e.ctypes.data_as(_dp)), "qbh_update_vhxc")
        return tuple(e)

    def current(self, psi_ptr, ld, nst, occ, weight, omega, A=None, nelectrons=0.0, j_ptr=None):
        """CurrentDensity::update_current; returns the total current (3,) and fills the
        current-density grids at j_ptr (3 x N doubles) when given."""
        o = np.ascontiguousarray(occ, dtype=np.float64)
        a = None if A is None else np.ascontiguousarray(A, dtype=np.float64)
        t = np.zeros(3)
        self.ctx.check(lib().qbh_current(self.h, psi_ptr, ld, nst, o.ctypes.data_as(_dp),
                                         float(weight), float(omega),
                                         None if a is None else a.ctypes.data_as(_dp),
                                         float(nelectrons), j_ptr, t.ctypes.data_as(_dp)),
                       "qbh_current")
        return t

    def gram(self, psi_ptr, ld, nst):
        self.ctx.check(lib().qbh_gram(self.h, psi_ptr, ld, nst), "qbh_gram")

    def subspace_diag(self, psi_ptr, ld, hpsi_ptr, ldh, nst, eigvec=True):
        w = np.zeros(nst)
        self.ctx.check(lib().qbh_subspace_diag(self.h, psi_ptr, ld, hpsi_ptr, ldh, nst,
                                               w.ctypes.data_as(_dp), int(eigvec)),
                       "qbh_subspace_diag")
        return w

    def close(self):
        if self.h:
            lib().qbh_basis_destroy(self.h)
            self.h = _vp()


class Species:
    """Radial tables of a norm-conserving KB species (qbh_species_*), from a species XML
    file (path) or from radial arrays (zval, lmax, llocal, deltar, vps, phi); rcps is the
    Gaussian width of the ion pseudo-charge."""

    def __init__(self, path=None, rcps=1.5, zval=None, lmax=None, llocal=None, deltar=None,
                 vps=None, phi=None):
        self.h = _vp()
        if path is not None:
            rc = lib().qbh_species_create_xml(str(path).encode(), float(rcps), ctypes.byref(self.h))
        else:
            v = np.ascontiguousarray(vps, dtype=np.float64)
            f = np.ascontiguousarray(phi, dtype=np.float64)
            npts = v.shape[1]
            rc = lib().qbh_species_create(float(zval), int(lmax), int(llocal), float(deltar), npts,
                                          v.ctypes.data_as(_dp), f.ctypes.data_as(_dp),
                                          float(rcps), ctypes.byref(self.h))
        if rc != 0:
            raise QbhError(f"qbh_species_create failed ({rc})")
        z, rc_, es = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int()
        lib().qbh_species_info(self.h, ctypes.byref(z), ctypes.byref(rc_), ctypes.byref(es),
                               ctypes.byref(n), None, None)
        self.zval, self.rcps, self.eself, self.nproj = z.value, rc_.value, es.value, n.value
        self.lproj = np.zeros(max(n.value, 1), dtype=np.int32)
        self.wt = np.zeros(max(n.value, 1))
        lib().qbh_species_info(self.h, None, None, None, None, self.lproj.ctypes.data_as(_ip),
                               self.wt.ctypes.data_as(_dp))
        self.lproj, self.wt = self.lproj[:n.value], self.wt[:n.value]

    def eval(self, which, g):
        """(v, dv) of Vloc(G) (which = -1) or of the l = which projector at |G| = g."""
        g_, gp = _da(g)
        v = np.zeros(g_.size)
        dv = np.zeros(g_.size)
        rc = lib().qbh_species_eval(self.h, int(which), g_.size, gp, v.ctypes.data_as(_dp),
                                    dv.ctypes.data_as(_dp))
        if rc != 0:
            raise QbhError(f"qbh_species_eval failed ({rc})")
        return v, dv

    def twnl(self, basis, twnl_ptr, A=None):
        """Projector tables (nproj rows of ngw) at |k+G-A| into device memory twnl_ptr."""
        a_ = None if A is None else np.ascontiguousarray(A, dtype=np.float64)
        basis.ctx.check(lib().qbh_species_twnl(self.h, basis.h,
                                               None if a_ is None else a_.ctypes.data_as(_dp),
                                               twnl_ptr), "qbh_species_twnl")

    def dtwnl(self, basis, dtwnl_ptr, A=None):
        """d twnl / d(k+G-A) (3 blocks of nproj rows of ngw) into device memory dtwnl_ptr."""
        a_ = None if A is None else np.ascontiguousarray(A, dtype=np.float64)
        basis.ctx.check(lib().qbh_species_dtwnl(self.h, basis.h,
                                                None if a_ is None else a_.ctypes.data_as(_dp),
                                                dtwnl_ptr), "qbh_species_dtwnl")

    def close(self):
        if self.h:
            lib().qbh_species_destroy(self.h)
            self.h = _vp()


def auto_rcps(zval, a, ecut):
    """AtomSet::set_rcps's Gaussian width for valence zval, cell a (rows), cutoff ecut (Ha)."""
    a_, ap = _da(np.asarray(a, dtype=np.float64).reshape(9))
    out = ctypes.c_double(0.0)
    if lib().qbh_species_auto_rcps(float(zval), ap, float(ecut), ctypes.byref(out)) != 0:
        raise QbhError("qbh_species_auto_rcps failed")
    return out.value


def local_tables(vbasis, species, na, tau_ptr, omega, vion_ptr, rhops_ptr):
    """v_ion(G) and rho_ps(G) on the density basis (qbh_local_tables); tau_ptr: device,
    3 doubles per atom, species after species."""
    sp = (_vp * max(len(species), 1))(*[s.h for s in species])
    na_, nap = _ia(na)
    vbasis.ctx.check(lib().qbh_local_tables(vbasis.h, len(species), sp, nap, tau_ptr,
                                            float(omega), vion_ptr, rhops_ptr),
                     "qbh_local_tables")


def ewald_esr(a, species, na, tau):
    """esr of EnergyFunctional::atoms_moved (qbh_ewald_esr); tau host (natoms, 3)."""
    a_, ap = _da(np.asarray(a, dtype=np.float64).reshape(9))
    sp = (_vp * max(len(species), 1))(*[s.h for s in species])
    na_, nap = _ia(na)
    t_, tp = _da(np.asarray(tau, dtype=np.float64).reshape(-1))
    out = ctypes.c_double(0.0)
    rc = lib().qbh_ewald_esr(ap, len(species), sp, nap, tp, ctypes.byref(out))
    if rc != 0:
        raise QbhError(f"qbh_ewald_esr failed ({rc})")
    return out.value


class Hamiltonian:
    def __init__(self, ctx, basis):
        self.ctx, self.basis = ctx, basis
        self.h = _vp()
        self._natoms = 0
        ctx.check(lib().qbh_ham_create(ctx.h, basis.h, ctypes.byref(self.h)), "qbh_ham_create")

    def set_vloc(self, v_ptr):
        self.ctx.check(lib().qbh_ham_set_vloc(self.h, v_ptr), "qbh_ham_set_vloc")

    def set_projectors(self, na, npr, lproj, wt, twnl, tau, omega):
        keep = [_ia(na), _ia(npr), _ia(lproj), _da(wt), _da(twnl), _da(tau)]
        self._natoms = int(np.sum(keep[0][0]))
        self.ctx.check(lib().qbh_ham_set_projectors(self.h, len(keep[0][0]),
                                                    *[k[1] for k in keep], float(omega)),
                       "qbh_ham_set_projectors")

    def set_species(self, species, na, tau, omega):
        """Projector tables from Species objects (qbh_ham_set_species), re-tabled at |k+G-A|
        whenever the vector potential changes; the species must stay open meanwhile."""
        sp = (_vp * max(len(species), 1))(*[s.h for s in species])
        na_, nap = _ia(na)
        t_, tp = _da(np.asarray(tau, dtype=np.float64).reshape(-1))
        self._natoms = int(np.sum(na_))
        self._species = list(species)
        self.ctx.check(lib().qbh_ham_set_species(self.h, len(species), sp, nap, tp, float(omega)),
                       "qbh_ham_set_species")

    def set_vector_potential(self, A):
        a_, ap = _da(np.asarray(A, dtype=np.float64))
        self.ctx.check(lib().qbh_ham_set_vector_potential(self.h, ap), "set_vector_potential")

    def hpsi(self, psi_ptr, ld, nst, occ, hpsi_ptr, ldh, flags=QBH_ALL, want_enl=False):
        o_, op = _da(occ)
        enl = ctypes.c_double(0.0)
        self.ctx.check(lib().qbh_hpsi(self.h, psi_ptr, ld, nst, op, hpsi_ptr, ldh, flags,
                                      ctypes.byref(enl) if want_enl else None), "qbh_hpsi")
        return enl.value if want_enl else None

    def nl_forces(self, psi_ptr, ld, nst, occ, fp32=False, want_enl=False):
        """Non-local ionic forces (natoms, 3), atoms in species order; with want_enl also
        Enl: (fion, enl)."""
        o_, op = _da(occ)
        natoms = int(self._natoms)
        f = np.zeros(3 * max(natoms, 1))
        enl = ctypes.c_double(0.0)
        self.ctx.check(lib().qbh_nl_forces(self.h, psi_ptr, ld, nst, op, f.ctypes.data_as(_dp),
                                           QBH_FP32 if fp32 else 0,
                                           ctypes.byref(enl) if want_enl else None),
                       "qbh_nl_forces")
        f = f[:3 * natoms].reshape(natoms, 3)
        return (f, enl.value) if want_enl else f

    def scf_run(self, vbasis, psi_ptr, ld, nst, occ, omega, vion_g_ptr, rhops_g_ptr, nitscf,
                nite, ecutprec, alpha=0.5, ndim=3, rc_kerker=10.0, flags_inner=QBH_ALL,
                flags_last=QBH_ALL):
        """BOSampleStepper SCF loop (PSDA + Anderson/Kerker mixing); returns the history
        (nitscf, max(nite,1), 7): etotal, ekin, enl, eps, ehart, exc, theta."""
        o = np.ascontiguousarray(occ, dtype=np.float64)
        nin = max(nite, 1)
        hist = np.zeros((nitscf, nin, 7))
        self.ctx.check(lib().qbh_scf_run(self.h, vbasis.h, psi_ptr, ld, nst, o.ctypes.data_as(_dp),
                                         float(omega), vion_g_ptr, rhops_g_ptr, int(nitscf),
                                         int(nite), float(ecutprec), float(alpha), int(ndim),
                                         float(rc_kerker), int(flags_inner), int(flags_last),
                                         hist.ctypes.data_as(_dp)), "qbh_scf_run")
        return hist

    def scf_run_smeared(self, vbasis, psi_ptr, ld, nst, occ, nel, smearing, omega, vion_g_ptr,
                        rhops_g_ptr, nitscf, nite, ecutprec, alpha=0.5, ndim=3, rc_kerker=10.0,
                        flags_inner=QBH_ALL, flags_last=QBH_ALL):
        """scf_run with Fermi smearing of width `smearing` (Rydberg) for nel electrons;
        returns (history, final occupations, eigenvalues per cycle (nitscf, nst), mu per
        cycle)."""
        o = np.array(occ, dtype=np.float64)
        nin = max(nite, 1)
        hist = np.zeros((nitscf, nin, 7))
        eig = np.zeros((nitscf, nst))
        mu = np.zeros(nitscf)
        self.ctx.check(lib().qbh_scf_run_smeared(
            self.h, vbasis.h, psi_ptr, ld, nst, o.ctypes.data_as(_dp), float(nel),
            float(smearing), float(omega), vion_g_ptr, rhops_g_ptr, int(nitscf), int(nite),
            float(ecutprec), float(alpha), int(ndim), float(rc_kerker), int(flags_inner),
            int(flags_last), hist.ctypes.data_as(_dp), eig.ctypes.data_as(_dp),
            mu.ctypes.data_as(_dp)), "qbh_scf_run_smeared")
        return hist, o, eig, mu

    def set_tuning(self, knob, value):
        self.ctx.check(lib().qbh_ham_set_tuning(self.h, knob.encode(), int(value)),
                       "qbh_ham_set_tuning")

    def get_tuning(self, knob):
        v = ctypes.c_int(0)
        self.ctx.check(lib().qbh_ham_get_tuning(self.h, knob.encode(), ctypes.byref(v)),
                       "qbh_ham_get_tuning")
        return v.value

    KNOBS = ("zy_reg", "x_reg", "x_spw", "radix", "nsplit", "overlap", "fnl_algo", "enl_algo")

    def autotune(self, psi_ptr, ld, nst, occ, hpsi_ptr, ldh, flags=QBH_ALL, reps=3, force=False,
                 knobs=None):
        """Times the launch choices for this shape and keeps the fastest; `knobs` (names)
        restricts the search (default: all).  Returns the report."""
        o_, op = _da(occ)
        mask = 0 if knobs is None else sum(1 << self.KNOBS.index(k) for k in knobs)
        buf = ctypes.create_string_buffer(8192)
        self.ctx.check(lib().qbh_ham_autotune(self.h, psi_ptr, ld, nst, op, hpsi_ptr, ldh, flags,
                                              int(reps), int(bool(force)), int(mask), buf,
                                              len(buf)), "qbh_ham_autotune")
        return buf.value.decode()

    def exponential(self, psi_ptr, ld, nst, dt, order=4, flags=QBH_ALL):
        self.ctx.check(lib().qbh_exponential(self.h, psi_ptr, ld, nst, float(dt), int(order),
                                             flags), "qbh_exponential")

    def etrs_step(self, vbasis, psi_ptr, ld, nst, occ, weight, dt, vion_g_ptr, rhops_g_ptr,
                  omega, order=4, flags=QBH_ALL):
        """One self-consistent ETRS step (preupdate + update); returns the {eps, ehart, exc}
        of the H(t+dt) density.  v_r of this Hamiltonian becomes H(t+dt)'s."""
        o = np.ascontiguousarray(occ, dtype=np.float64)
        e = np.zeros(3)
        self.ctx.check(lib().qbh_etrs_step(self.h, vbasis.h, psi_ptr, ld, nst, o.ctypes.data_as(_dp),
                                           float(weight), float(dt), int(order), vion_g_ptr,
                                           rhops_g_ptr, float(omega), e.ctypes.data_as(_dp),
                                           flags), "qbh_etrs_step")
        return tuple(e)

    def close(self):
        if self.h:
            lib().qbh_ham_destroy(self.h)
            self.h = _vp()


class TDPropagator:
    """Real-time TDDFT over many ETRS iterations with the objective J = sum_n w_n e.j(psi_n,
    A_n) and its adjoint (qbh_td_*).  ham and vbasis must outlive it; vion/rhops device
    pointers are borrowed."""

    def __init__(self, ham, vbasis, nst, occ, weight, omega, vion_g_ptr, rhops_g_ptr):
        self.ctx = ham.ctx
        self.h = _vp()
        self._occ = np.ascontiguousarray(occ, dtype=np.float64)
        self.ctx.check(lib().qbh_td_create(ham.h, vbasis.h, int(nst), self._occ.ctypes.data_as(_dp),
                                           float(weight), float(omega), vion_g_ptr, rhops_g_ptr,
                                           ctypes.byref(self.h)), "qbh_td_create")
        self.nst = nst
        self.nsteps = None

    def forward(self, psi0_ptr, ld, nsteps, dt, A, w, e=(1.0, 0.0, 0.0), vext_ptr=None,
                checkpoint=20, order=4, flags=0, psi_out_ptr=None, ldo=None):
        A = np.ascontiguousarray(A, dtype=np.float64).reshape(nsteps + 1, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(nsteps + 1)
        e = np.ascontiguousarray(e, dtype=np.float64).reshape(3)
        J = ctypes.c_double(0.0)
        self.ctx.check(lib().qbh_td_forward(self.h, psi0_ptr, int(ld), int(nsteps), float(dt),
                                            int(order), A.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                                            e.ctypes.data_as(_dp), vext_ptr, int(checkpoint),
                                            int(flags), ctypes.byref(J), psi_out_ptr,
                                            int(ldo if ldo is not None else ld)), "qbh_td_forward")
        self.nsteps = nsteps
        return J.value

    def backward(self, dvext_ptr=None):
        """Returns dJ/dA (nsteps+1, 3); fills dJ/dv_ext at dvext_ptr when given."""
        dA = np.zeros((self.nsteps + 1, 3))
        self.ctx.check(lib().qbh_td_backward(self.h, dvext_ptr, dA.ctypes.data_as(_dp)),
                       "qbh_td_backward")
        return dA

    def close(self):
        if self.h:
            lib().qbh_td_destroy(self.h)
            self.h = _vp()


class Batch:
    """Hψ of several cells captured into one hipGraph (qbh_batch_*).  cells: list of
    (Hamiltonian, psi_ptr, ld, nst, hpsi_ptr, ldh); all on ctx, whose stream must be a
    non-default stream."""

    def __init__(self, ctx, cells, flags=QBH_ALL, nbranch=4):
        n = len(cells)
        self.ctx = ctx
        self.h = _vp()
        hams = (_vp * n)(*[c[0].h for c in cells])
        psis = (_vp * n)(*[c[1] for c in cells])
        lds = (_c_long * n)(*[c[2] for c in cells])
        nsts = (ctypes.c_int * n)(*[c[3] for c in cells])
        outs = (_vp * n)(*[c[4] for c in cells])
        ldhs = (_c_long * n)(*[c[5] for c in cells])
        ctx.check(lib().qbh_batch_create(ctx.h, n, hams, psis, lds, nsts, outs, ldhs, flags,
                                         int(nbranch), ctypes.byref(self.h)), "qbh_batch_create")

    def launch(self):
        self.ctx.check(lib().qbh_batch_launch(self.h), "qbh_batch_launch")

    def close(self):
        if self.h:
            lib().qbh_batch_destroy(self.h)
            self.h = _vp()


class PSDA:
    """PSDA wavefunction stepper (qbh_psda_*) on a basis."""

    def __init__(self, basis, nst, ecutprec):
        self.ctx = basis.ctx
        self.h = _vp()
        self.ctx.check(lib().qbh_psda_create(basis.h, int(nst), float(ecutprec), ctypes.byref(self.h)),
                       "qbh_psda_create")

    def preprocess(self):
        self.ctx.check(lib().qbh_psda_preprocess(self.h), "qbh_psda_preprocess")

    def update(self, psi_ptr, ld, hpsi_ptr, ldh, occ):
        o = np.ascontiguousarray(occ, dtype=np.float64)
        th = ctypes.c_double(0.0)
        self.ctx.check(lib().qbh_psda_update(self.h, psi_ptr, ld, hpsi_ptr, ldh,
                                             o.ctypes.data_as(_dp), ctypes.byref(th)),
                       "qbh_psda_update")
        return th.value

    def close(self):
        if self.h:
            lib().qbh_psda_destroy(self.h)
            self.h = _vp()


class Mixer:
    """Anderson mixer (qbh_mixer_*) over m doubles with nmax previous pairs."""

    def __init__(self, ctx, m, nmax):
        self.ctx = ctx
        self.h = _vp()
        self.nmax = nmax
        self.updates = 0
        ctx.check(lib().qbh_mixer_create(ctx.h, int(m), int(nmax), ctypes.byref(self.h)),
                  "qbh_mixer_create")

    def restart(self):
        self.updates = 0
        self.ctx.check(lib().qbh_mixer_restart(self.h), "qbh_mixer_restart")

    def update(self, x_ptr, f_ptr, xbar_ptr, fbar_ptr):
        th = np.zeros(max(self.nmax, 1))
        self.ctx.check(lib().qbh_mixer_update(self.h, x_ptr, f_ptr, xbar_ptr, fbar_ptr,
                                              th.ctypes.data_as(_dp)), "qbh_mixer_update")
        n = min(self.updates, self.nmax)
        self.updates += 1
        return th[:n]

    def close(self):
        if self.h:
            lib().qbh_mixer_destroy(self.h)
            self.h = _vp()


def kinetic_energy(basis, psi_ptr, ld, nst, occ):
    o = np.ascontiguousarray(occ, dtype=np.float64)
    e = ctypes.c_double(0.0)
    basis.ctx.check(lib().qbh_kinetic_energy(basis.h, psi_ptr, ld, int(nst), o.ctypes.data_as(_dp),
                                             ctypes.byref(e)), "qbh_kinetic_energy")
    return e.value


def wf_write_raw(path, psi_ptr, ld, nst, eig=None, occ=None, append=False):
    """Wavefunction::write_fast per-task block (qbh_wf_write_raw)."""
    e = None if eig is None else np.ascontiguousarray(eig, dtype=np.float64)
    o = None if occ is None else np.ascontiguousarray(occ, dtype=np.float64)
    rc = lib().qbh_wf_write_raw(str(path).encode(), int(bool(append)), psi_ptr, int(ld), int(nst),
                                None if e is None else e.ctypes.data_as(_dp),
                                None if o is None else o.ctypes.data_as(_dp))
    if rc != 0:
        raise QbhError(f"qbh_wf_write_raw failed ({rc})")


def wf_read_raw(path, psi_ptr, ld, nst, offset=0, with_eig_occ=False):
    """Reads one block; returns (next_offset, eig, occ)."""
    e = np.zeros(nst) if with_eig_occ else None
    o = np.zeros(nst) if with_eig_occ else None
    nxt = _c_long(0)
    rc = lib().qbh_wf_read_raw(str(path).encode(), int(offset), psi_ptr, int(ld), int(nst),
                               None if e is None else e.ctypes.data_as(_dp),
                               None if o is None else o.ctypes.data_as(_dp), ctypes.byref(nxt))
    if rc != 0:
        raise QbhError(f"qbh_wf_read_raw failed ({rc})")
    return nxt.value, e, o


def sd_write_xml(basis, path, psi_ptr, ld, nst, occ, weight=1.0, ispin=0, nspin=1, append=False):
    """SlaterDet::write with base64 grid functions (qbh_sd_write_xml)."""
    o = np.ascontiguousarray(occ, dtype=np.float64)
    basis.ctx.check(lib().qbh_sd_write_xml(basis.h, str(path).encode(), int(bool(append)), psi_ptr,
                                           int(ld), int(nst), o.ctypes.data_as(_dp), float(weight),
                                           int(ispin), int(nspin)), "qbh_sd_write_xml")


def sd_read_xml(basis, path, psi_ptr, ld, nst, offset=0):
    """Reads one <slater_determinant>; returns (next_offset, occupations)."""
    o = np.zeros(nst)
    nxt = _c_long(0)
    basis.ctx.check(lib().qbh_sd_read_xml(basis.h, str(path).encode(), int(offset), psi_ptr,
                                          int(ld), int(nst), o.ctypes.data_as(_dp),
                                          ctypes.byref(nxt)), "qbh_sd_read_xml")
    return nxt.value, o
