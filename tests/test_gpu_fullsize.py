"""Hψ at the full benchmark sizes: Si512 Γ (72919 plane waves, 1024 bands, 140^3 grid,
2048 s+p projectors) and fcc Al500 Γ (77666 plane waves, 900 bands, 144^3 grid, 4500
s+p+d projectors).  All bands go through the GPU path at once; the oracle checks a band
subset at the same size (band results do not depend on the other bands), and the
size-independent properties are checked on all of them: Hermiticity of the band matrix
2 Re(Ψ^H HΨ) - ψ(0)Hψ(0) and the FP32 form against the FP64 one."""
import numpy as np
import pytest

from helpers import from_dev, oracle_for, real_to_dev, rel_err, to_dev

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-12
FP32_TOL = 2e-6


def _system(which):
    from qball_amd.workload import make_fcc_al, make_system
    if which == "si512":
        return make_system(ncell=4, ecut_ry=25.0, seed=0)
    return make_fcc_al(seed=1)


@pytest.mark.parametrize("which", ["si512", "al500"])
def test_full_size_hpsi(gpu_ctx, which):
    import torch
    from oracle import hpsi as oh
    from qball_amd import Basis, Hamiltonian
    s = _system(which)
    print(which, "ngw", s.ngw, "nst", s.nst, "grid", s.np, "projectors", s.na * len(s.lproj),
          flush=True)
    r = s.rods
    basis = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np,
                  s.b)
    ham = Hamiltonian(gpu_ctx, basis)
    vd = real_to_dev(s.v)
    ham.set_vloc(vd.data_ptr())
    ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(),
                       s.omega)
    ngw, nst = s.ngw, s.nst
    cd = to_dev(s.psi)
    h64 = torch.empty_like(cd)
    h32 = torch.empty_like(cd)
    ham.hpsi(cd.data_ptr(), ngw, nst, s.occ, h64.data_ptr(), ngw, flags=0x7)
    ham.hpsi(cd.data_ptr(), ngw, nst, s.occ, h32.data_ptr(), ngw, flags=0x17)
    gpu_ctx.sync()
    hp = from_dev(h64, ngw, nst)

    # oracle on the first and the last band pair
    sub = [0, 1, nst - 2, nst - 1]
    ob, maps, sp = oracle_for(s)
    ref, enl_ref = oh.hpsi(maps, s.psi[:, sub], s.v, [sp], s.occ[sub], s.omega)
    assert rel_err(hp[:, sub], ref) <= FP64_TOL
    # the same four bands alone through the GPU path, with Enl
    c4 = to_dev(s.psi[:, sub])
    h4 = torch.empty_like(c4)
    enl = ham.hpsi(c4.data_ptr(), ngw, 4, s.occ[sub], h4.data_ptr(), ngw, want_enl=True)
    gpu_ctx.sync()
    assert rel_err(from_dev(h4, ngw, 4), ref) <= FP64_TOL
    assert abs(enl - enl_ref) <= 1e-10 * max(1.0, abs(enl_ref))

    # FP32-bulk form against FP64, all bands
    assert rel_err(from_dev(h32, ngw, nst), hp) <= FP32_TOL

    # Hermiticity over 128 bands (Γ metric, G=0 once): M = 2 Re(Ψ^H HΨ) - ψ(0) Hψ(0)^T
    q = np.asarray(s.psi[:, :128])
    hq = hp[:, :128]
    m = 2.0 * (q.conj().T @ hq).real - np.outer(q[0].real, hq[0].real)
    assert np.linalg.norm(m - m.T) <= 1e-12 * np.linalg.norm(m)
    ham.close()
    basis.close()


@pytest.mark.parametrize("which", ["si512", "al500"])
def test_full_size_nl_forces(gpu_ctx, which):
    """Non-local forces at the benchmark sizes: additivity over bands, the FP32 form against
    FP64, -dEnl/dtau of the library's own Enl by central differences for one atom, and (Si512)
    the oracle on a band subset."""
    from oracle import hpsi as oh
    from qball_amd import Basis, Hamiltonian
    s = _system(which)
    r = s.rods
    basis = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np,
                  s.b)
    ham = Hamiltonian(gpu_ctx, basis)

    def place(tau):
        ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), tau.ravel(),
                           s.omega)

    place(s.tau)
    ngw, nst = s.ngw, s.nst
    cd = to_dev(s.psi)
    f, enl = ham.nl_forces(cd.data_ptr(), ngw, nst, s.occ, want_enl=True)
    f32 = ham.nl_forces(cd.data_ptr(), ngw, nst, s.occ, fp32=True)
    fmax = np.max(np.abs(f))
    print(which, "Enl", enl, "max|F|", fmax, "fp32 vs fp64", rel_err(f32, f), flush=True)
    assert rel_err(f32, f) <= 1e-5
    # forces are sums over bands: two halves add up to the whole
    h = nst // 2
    fa = ham.nl_forces(cd.data_ptr(), ngw, h, s.occ[:h])
    fb = ham.nl_forces(cd.data_ptr() + 16 * ngw * h, ngw, nst - h, s.occ[h:])
    assert rel_err(fa + fb, f) <= 1e-12
    # F = -dEnl/dtau, central differences of the library's Enl (step 2e-4 bohr)
    step = 2e-4
    ia = 1
    for j in range(3):
        e = []
        for sg in (1, -1):
            tau = s.tau.copy()
            tau[ia, j] += sg * step
            place(tau)
            e.append(ham.nl_forces(cd.data_ptr(), ngw, nst, s.occ, want_enl=True)[1])
        fd = -(e[0] - e[1]) / (2 * step)
        print(which, "atom", ia, "dir", j, "F", f[ia, j], "fd", fd, flush=True)
        assert abs(fd - f[ia, j]) <= 1e-5 * fmax
    place(s.tau)
    if which == "si512":
        sub = [0, 1, nst - 2, nst - 1]
        ob, _, sp = oracle_for(s)
        enl_ref, fref = oh.nonlocal_forces(ob, [sp], s.psi[:, sub], s.occ[sub], s.omega)
        c4 = to_dev(s.psi[:, sub])
        f4, enl4 = ham.nl_forces(c4.data_ptr(), ngw, 4, s.occ[sub], want_enl=True)
        assert rel_err(f4, fref[0]) <= 1e-12
        assert abs(enl4 - enl_ref) <= 1e-10 * max(1.0, abs(enl_ref))
    ham.close()
    basis.close()
