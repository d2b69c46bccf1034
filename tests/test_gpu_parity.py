"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

FP64 bar: relative 2-norm error <= 1e-12 for transforms, Hψ and density; energies
<= 1e-10 Ha absolute.  FP32-bulk bar: relative error <= 2e-6 (float rounding of the
transforms and projector GEMMs, accumulated into FP64).
"""
import numpy as np
import pytest

from helpers import (from_dev, make_gpu_basis, make_gpu_ham, max_rel, oracle_for, real_to_dev,
                     rel_err, to_dev)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


@pytest.fixture(scope="module")
def si8():
    from qball_amd.workload import make_system
    return make_system(ncell=1, ecut_ry=25.0, nbands=15, seed=11)


@pytest.fixture(scope="module")
def si8c():
    from qball_amd.workload import make_system
    return make_system(ncell=1, ecut_ry=25.0, nbands=8, seed=12, complex_basis=True)


@pytest.fixture(scope="module")
def si64():
    from qball_amd.workload import make_system
    return make_system(ncell=2, ecut_ry=25.0, seed=0)


def _backward_check(ctx, sysm, pair, flags=0, tol=FP64_TOL):
    import torch
    from oracle import fft as offt
    ob, maps, _ = oracle_for(sysm)
    basis = make_gpu_basis(ctx, sysm)
    assert basis.ngw == ob.size
    nb = sysm.nst
    c = sysm.psi
    cd = to_dev(c)
    nslots = (nb + 1) // 2 if pair else nb
    f = torch.zeros(2 * nslots * basis.N, dtype=torch.float64, device="cuda")
    basis.backward(cd.data_ptr(), basis.ngw, nb, pair, f.data_ptr(), flags)
    ctx.sync()
    fg = f.cpu().numpy().view(np.complex128).reshape(nslots, -1)
    errs = []
    for s in range(nslots):
        if pair:
            n1, n2 = 2 * s, 2 * s + 1
            ref = offt.backward(maps, c[:, n1], c[:, n2] if n2 < nb else None).ravel()
        else:
            ref = offt.backward(maps, c[:, s]).ravel()
        errs.append(rel_err(fg[s], ref))
    basis.close()
    assert max(errs) <= tol, errs


def test_fft_backward_pair_si8(gpu_ctx, si8):
    _backward_check(gpu_ctx, si8, pair=True)


def test_fft_backward_single_si8(gpu_ctx, si8):
    _backward_check(gpu_ctx, si8, pair=False)


def test_fft_backward_complex_si8(gpu_ctx, si8c):
    _backward_check(gpu_ctx, si8c, pair=False)


def test_fft_backward_fp32_si8(gpu_ctx, si8):
    _backward_check(gpu_ctx, si8, pair=True, flags=0x10, tol=FP32_TOL)


@pytest.mark.parametrize("pair", [True, False])
def test_fft_forward_si8(gpu_ctx, si8, pair):
    import torch
    from oracle import fft as offt
    ob, maps, _ = oracle_for(si8)
    basis = make_gpu_basis(gpu_ctx, si8)
    rng = np.random.default_rng(5)
    nb = 5
    nslots = (nb + 1) // 2 if pair else nb
    f = rng.standard_normal((nslots, basis.N)) + 1j * rng.standard_normal((nslots, basis.N))
    fd = torch.from_numpy(f.view(np.float64).ravel().copy()).cuda()
    out = torch.zeros(2 * basis.ngw * nb, dtype=torch.float64, device="cuda")
    basis.forward(fd.data_ptr(), nb, pair, out.data_ptr(), basis.ngw)
    gpu_ctx.sync()
    cg = from_dev(out, basis.ngw, nb)
    for s in range(nslots):
        fs = f[s].reshape(maps.np2, maps.np1, maps.np0)
        if pair:
            c1, c2 = offt.forward(maps, fs, pair=True)
            assert rel_err(cg[:, 2 * s], c1) <= FP64_TOL
            if 2 * s + 1 < nb:
                assert rel_err(cg[:, 2 * s + 1], c2) <= FP64_TOL
        else:
            assert rel_err(cg[:, s], offt.forward(maps, fs)) <= FP64_TOL
    basis.close()


def _hpsi_check(ctx, sysm, flags, tol=FP64_TOL, with_proj=True):
    import torch
    from oracle import hpsi as oh
    ob, maps, sp = oracle_for(sysm)
    basis = make_gpu_basis(ctx, sysm)
    ham = make_gpu_ham(ctx, basis, sysm, with_proj=with_proj)
    c = sysm.psi
    cd = to_dev(c)
    hd = torch.zeros_like(cd)
    enl = ham.hpsi(cd.data_ptr(), basis.ngw, sysm.nst, sysm.occ, hd.data_ptr(), basis.ngw,
                   flags=flags, want_enl=True)
    ctx.sync()
    hg = from_dev(hd, basis.ngw, sysm.nst)
    ref = np.zeros_like(c)
    enl_ref = 0.0
    if flags & 0x1:
        enl_ref = oh.nonlocal_energy(ob, [sp], c, sysm.occ, sysm.omega, cp=ref)
    if flags & 0x2:
        oh.kinetic_add(ob.kpg2, c, ref)
    if flags & 0x4:
        oh.rs_mul_add(maps, c, sysm.v, ref)
    ham.close()
    basis.close()
    e = rel_err(hg, ref)
    assert e <= tol, e
    if flags & 0x1:
        etol = 1e-10 if tol == FP64_TOL else 1e-5 * max(1.0, abs(enl_ref))
        assert abs(enl - enl_ref) <= etol, (enl, enl_ref)
    return e


def test_hpsi_local_si8(gpu_ctx, si8):
    _hpsi_check(gpu_ctx, si8, 0x4)


def test_hpsi_kinetic_only_si8(gpu_ctx, si8):
    _hpsi_check(gpu_ctx, si8, 0x2)


def test_hpsi_nonlocal_si8(gpu_ctx, si8):
    _hpsi_check(gpu_ctx, si8, 0x1)


def test_hpsi_full_si8(gpu_ctx, si8):
    _hpsi_check(gpu_ctx, si8, 0x7)


def test_hpsi_full_complex_si8(gpu_ctx, si8c):
    _hpsi_check(gpu_ctx, si8c, 0x7)


def test_hpsi_full_si64(gpu_ctx, si64):
    _hpsi_check(gpu_ctx, si64, 0x7)


def test_hpsi_fp32_si64(gpu_ctx, si64):
    _hpsi_check(gpu_ctx, si64, 0x7 | 0x10, tol=FP32_TOL)


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_density(gpu_ctx, which, request):
    import torch
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    ob, maps, _ = oracle_for(sysm)
    basis = make_gpu_basis(gpu_ctx, sysm)
    cd = to_dev(sysm.psi)
    occ = sysm.occ.copy()
    occ[-1] = 0.0  # a band with zero factor is skipped
    rho = torch.zeros(basis.N, dtype=torch.float64, device="cuda")
    basis.density(cd.data_ptr(), basis.ngw, sysm.nst, occ, 1.0, sysm.omega, rho.data_ptr())
    gpu_ctx.sync()
    ref = oh.compute_density(maps, sysm.psi, occ, 1.0, sysm.omega)
    basis.close()
    assert rel_err(rho.cpu().numpy(), ref) <= FP64_TOL


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_density_fp32(gpu_ctx, which, request):
    """ρ from the FP32 transforms (FP64 band sum) against the oracle at the FP32 bar."""
    import torch
    from oracle import hpsi as oh
    from qball_amd import QBH_FP32
    sysm = request.getfixturevalue(which)
    ob, maps, _ = oracle_for(sysm)
    basis = make_gpu_basis(gpu_ctx, sysm)
    cd = to_dev(sysm.psi)
    rho = torch.zeros(basis.N, dtype=torch.float64, device="cuda")
    basis.density(cd.data_ptr(), basis.ngw, sysm.nst, sysm.occ, 1.0, sysm.omega, rho.data_ptr(),
                  flags=QBH_FP32)
    gpu_ctx.sync()
    ref = oh.compute_density(maps, sysm.psi, sysm.occ, 1.0, sysm.omega)
    basis.close()
    assert rel_err(rho.cpu().numpy(), ref) <= FP32_TOL


def test_density_rhog_si8(gpu_ctx, si8):
    """update_density's charge_vft: rhog = forward(Omega*rho) on the density basis."""
    import torch
    from oracle import hpsi as oh
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from qball_amd import Basis
    from qball_amd import capi
    ob, maps, _ = oracle_for(si8)
    vob = OB(UnitCell(si8.a), 4.0 * si8.ecut)
    vmaps = FFTMaps(vob, *si8.np)
    rho, nel, rhog = oh.update_density(maps, vmaps, si8.psi, si8.occ, 1.0, si8.omega)
    r = capi.rods_build(si8.a, 4.0 * si8.ecut, True)
    vb = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], si8.np, si8.b)
    rd = real_to_dev(rho)
    out = torch.zeros(2 * vb.ngw, dtype=torch.float64, device="cuda")
    vb.forward_real(rd.data_ptr(), si8.omega, out.data_ptr())
    gpu_ctx.sync()
    g = out.cpu().numpy().view(np.complex128)
    vb.close()
    assert rel_err(g, rhog) <= FP64_TOL


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_gram(gpu_ctx, which, request):
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    basis = make_gpu_basis(gpu_ctx, sysm)
    cd = to_dev(sysm.psi)
    basis.gram(cd.data_ptr(), basis.ngw, sysm.nst)
    gpu_ctx.sync()
    got = from_dev(cd, basis.ngw, sysm.nst)
    ref = oh.gram(sysm.psi, sysm.real)
    basis.close()
    assert rel_err(got, ref) <= 1e-11


@pytest.mark.parametrize("order", [1, 2, 4, 6, 8])
def test_exponential_complex_si8(gpu_ctx, si8c, order):
    """The Taylor exponential at every order the entry point takes the ends and middle of
    (ExponentialWavefunctionStepper.cc:51-148 uses 4)."""
    from oracle import hpsi as oh
    ob, maps, sp = oracle_for(si8c)
    basis = make_gpu_basis(gpu_ctx, si8c)
    ham = make_gpu_ham(gpu_ctx, basis, si8c)
    cd = to_dev(si8c.psi)
    dt = 0.05
    ham.exponential(cd.data_ptr(), basis.ngw, si8c.nst, dt, order)
    gpu_ctx.sync()
    got = from_dev(cd, basis.ngw, si8c.nst)
    occ0 = np.zeros(si8c.nst)

    def apply_h(w):
        return oh.hpsi(maps, w, si8c.v, [sp], occ0, si8c.omega)[0]
    ref = oh.exponential(apply_h, si8c.psi, dt, order)
    ham.close()
    basis.close()
    assert rel_err(got, ref) <= FP64_TOL


def test_vector_potential_complex_si8(gpu_ctx, si8c):
    """Kinetic |k+G-A|^2 and projector phases at k+G-A (twnl kept fixed: the caller
    re-tables it)."""
    import torch
    from oracle import hpsi as oh
    ob, maps, sp = oracle_for(si8c)
    basis = make_gpu_basis(gpu_ctx, si8c)
    ham = make_gpu_ham(gpu_ctx, basis, si8c)
    A = np.array([0.01, -0.02, 0.005])
    ham.set_vector_potential(A)
    cd = to_dev(si8c.psi)
    hd = torch.zeros_like(cd)
    ham.hpsi(cd.data_ptr(), basis.ngw, si8c.nst, si8c.occ, hd.data_ptr(), basis.ngw)
    gpu_ctx.sync()
    got = from_dev(hd, basis.ngw, si8c.nst)
    ref = oh.hpsi(maps, si8c.psi, si8c.v, [sp], si8c.occ, si8c.omega,
                  kpg2=oh.kpgpa2(ob, A), kpgx=ob.kpgx - A[None, :])[0]
    assert rel_err(got, ref) <= FP64_TOL
    # the FP32 form builds its own FP32 projector table at k+G-A
    ham.hpsi(cd.data_ptr(), basis.ngw, si8c.nst, si8c.occ, hd.data_ptr(), basis.ngw, flags=0x17)
    gpu_ctx.sync()
    assert rel_err(from_dev(hd, basis.ngw, si8c.nst), ref) <= FP32_TOL
    ham.close()
    basis.close()


_REG_EDGES = (32, 40, 48, 50, 54, 80, 100, 108, 128)


@pytest.mark.parametrize("grid", [(40, 30, 42), (28, 36, 72), (30, 33, 44), (60, 64, 90),
                                  (64, 90, 96), (90, 96, 120), (96, 120, 60), (120, 60, 64)]
                         + [(_REG_EDGES[i], _REG_EDGES[(i + 3) % 9], _REG_EDGES[(i + 6) % 9])
                            for i in range(9)]
                         + [(150, 162, 196)]              # edges past 144 on the Stockham passes
                         + [g for L in (160, 168, 180, 192, 210, 224, 240, 256)
                            for g in ((L, 36, 40), (36, L, 40), (36, 40, L))])
def test_noncubic_grids_generic_and_register_paths(gpu_ctx, si8, grid):
    """Non-cubic grids mixing edges with and without a register factorisation: transforms
    and local Hψ through both the LDS Stockham and the two-factor register passes."""
    import torch
    from oracle import fft as offt
    from oracle import hpsi as oh
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from qball_amd import Basis, Hamiltonian
    ob = OB(UnitCell(si8.a), si8.ecut)
    maps = FFTMaps(ob, *grid)
    r = si8.rods
    basis = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], grid, si8.b)
    nb = si8.nst
    cd = to_dev(si8.psi)
    N = grid[0] * grid[1] * grid[2]
    nslots = (nb + 1) // 2
    f = torch.zeros(2 * nslots * N, dtype=torch.float64, device="cuda")
    basis.backward(cd.data_ptr(), basis.ngw, nb, True, f.data_ptr())
    gpu_ctx.sync()
    fg = f.cpu().numpy().view(np.complex128).reshape(nslots, -1)
    for s in range(nslots):
        n2 = 2 * s + 1
        ref = offt.backward(maps, si8.psi[:, 2 * s], si8.psi[:, n2] if n2 < nb else None).ravel()
        assert rel_err(fg[s], ref) <= FP64_TOL
    rng = np.random.default_rng(7)
    v = rng.standard_normal(N)
    vd = real_to_dev(v)
    ham = Hamiltonian(gpu_ctx, basis)
    ham.set_vloc(vd.data_ptr())
    hd = torch.zeros_like(cd)
    ham.hpsi(cd.data_ptr(), basis.ngw, nb, si8.occ, hd.data_ptr(), basis.ngw, flags=0x4)
    gpu_ctx.sync()
    ref = oh.rs_mul_add(maps, si8.psi, v, np.zeros_like(si8.psi))
    assert rel_err(from_dev(hd, basis.ngw, nb), ref) <= FP64_TOL
    ham.hpsi(cd.data_ptr(), basis.ngw, nb, si8.occ, hd.data_ptr(), basis.ngw, flags=0x14)
    gpu_ctx.sync()
    assert rel_err(from_dev(hd, basis.ngw, nb), ref) <= FP32_TOL
    rho = torch.zeros(N, dtype=torch.float64, device="cuda")
    basis.density(cd.data_ptr(), basis.ngw, nb, si8.occ, 1.0, si8.omega, rho.data_ptr())
    gpu_ctx.sync()
    assert rel_err(rho.cpu().numpy(), oh.compute_density(maps, si8.psi, si8.occ, 1.0, si8.omega)) <= FP64_TOL
    ham.close()
    basis.close()


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_subspace_diag(gpu_ctx, which, request):
    import torch
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    ob, maps, sp = oracle_for(sysm)
    c = oh.gram(sysm.psi, sysm.real)
    hc, _ = oh.hpsi(maps, c, sysm.v, [sp], sysm.occ, sysm.omega)
    w_ref, cz_ref = oh.subspace_diag(c, hc, sysm.real)
    basis = make_gpu_basis(gpu_ctx, sysm)
    cd = to_dev(c)
    hd = to_dev(hc)
    w = basis.subspace_diag(cd.data_ptr(), basis.ngw, hd.data_ptr(), basis.ngw, sysm.nst)
    gpu_ctx.sync()
    cz = from_dev(cd, basis.ngw, sysm.nst)
    basis.close()
    assert np.max(np.abs(w - w_ref)) <= 1e-12 * np.max(np.abs(w_ref))
    # eigenvectors are fixed up to a phase per column (sign for Γ)
    for n in range(sysm.nst):
        ph = np.vdot(cz[:, n], cz_ref[:, n])
        ph /= abs(ph)
        assert rel_err(cz[:, n] * ph, cz_ref[:, n]) <= 1e-10


def test_gram_si64_split_k(gpu_ctx, si64):
    """Si64 (18278-row real view): the overlap product runs K-split."""
    from oracle import hpsi as oh
    basis = make_gpu_basis(gpu_ctx, si64)
    cd = to_dev(si64.psi)
    basis.gram(cd.data_ptr(), basis.ngw, si64.nst)
    gpu_ctx.sync()
    got = from_dev(cd, basis.ngw, si64.nst)
    basis.close()
    assert rel_err(got, oh.gram(si64.psi, True)) <= 1e-11


@pytest.mark.parametrize("which,tddft", [("si8", False), ("si8c", False), ("si8c", True)])
def test_update_vhxc(gpu_ctx, which, tddft, request):
    import torch
    from oracle import hpsi as oh
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from oracle.vhxc import update_vhxc
    from qball_amd import Basis, capi
    sysm = request.getfixturevalue(which)
    ob, maps, _ = oracle_for(sysm)
    vob = OB(UnitCell(sysm.a), 4.0 * sysm.ecut, force_complex=not sysm.real)
    vmaps = FFTMaps(vob, *sysm.np)
    c = oh.gram(sysm.psi, sysm.real)
    rho, nel, rhog = oh.update_density(maps, vmaps, c, sysm.occ, 1.0, sysm.omega)
    rng = np.random.default_rng(21)
    vion = (rng.standard_normal(vob.size) + 1j * rng.standard_normal(vob.size)) * np.exp(-0.3 * vob.g2)
    rhops = (0.05 * np.exp(-0.5 * vob.g2)).astype(np.complex128)
    if sysm.real:
        vion[0] = vion[0].real
    hrhog = rhog * (1.0 + 0.01 * rng.standard_normal(vob.size)) if tddft else None
    v_ref, eps, eh, exc = update_vhxc(vmaps, rho, rhog, vion, rhops, sysm.omega, hamil_rhog=hrhog)
    r = capi.rods_build(sysm.a, 4.0 * sysm.ecut, sysm.real)
    vb = Basis(gpu_ctx, sysm.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np,
               sysm.b)
    cplx = lambda z: torch.from_numpy(np.ascontiguousarray(z).view(np.float64).copy()).cuda()
    rd, gd, vd, pd = real_to_dev(rho), cplx(rhog), cplx(vion), cplx(rhops)
    hd = cplx(hrhog) if tddft else None
    out = torch.zeros(vb.N, dtype=torch.float64, device="cuda")
    e = vb.update_vhxc(rd.data_ptr(), gd.data_ptr(), vd.data_ptr(), pd.data_ptr(), sysm.omega,
                       out.data_ptr(), hamil_rhog_ptr=hd.data_ptr() if tddft else None)
    gpu_ctx.sync()
    vb.close()
    assert rel_err(out.cpu().numpy(), v_ref) <= FP64_TOL
    assert abs(e[0] - eps) <= 1e-10 and abs(e[1] - eh) <= 1e-10 and abs(e[2] - exc) <= 1e-10


def test_etrs_step_selfconsistent_si8c(gpu_ctx, si8c):
    """qbh_etrs_step against the oracle's ETRS step with H(t+dt) rebuilt from the density
    of psi~(t+dt) (update_density + update_vhxc), two steps in a row."""
    import torch
    from oracle import hpsi as oh
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from oracle.vhxc import update_vhxc
    from qball_amd import Basis, capi
    sysm = si8c
    ob, maps, sp = oracle_for(sysm)
    vob = OB(UnitCell(sysm.a), 4.0 * sysm.ecut, force_complex=True)
    vmaps = FFTMaps(vob, *sysm.np)
    rng = np.random.default_rng(5)
    vion = (rng.standard_normal(vob.size) + 1j * rng.standard_normal(vob.size)) * 0.5 * np.exp(-0.3 * vob.g2)
    rhops = (0.05 * np.exp(-0.5 * vob.g2)).astype(np.complex128)
    occ0 = np.zeros(sysm.nst)
    psi = oh.gram(sysm.psi, False)
    dt = 0.05

    def oracle_step(c, v):
        def apply_t(w):
            return oh.hpsi(maps, w, v, [sp], occ0, sysm.omega)[0]

        state = {}

        def h_tdt(tilde):
            rho, _, rhog = oh.update_density(maps, vmaps, tilde, sysm.occ, 1.0, sysm.omega)
            v2, eps, eh, exc = update_vhxc(vmaps, rho, rhog, vion, rhops, sysm.omega, hamil_rhog=rhog)
            state.update(v=v2, e=(eps, eh, exc))
            return lambda w: oh.hpsi(maps, w, v2, [sp], occ0, sysm.omega)[0]
        out, _ = oh.etrs_step(apply_t, h_tdt, c, dt)
        return out, state["v"], state["e"]

    r = capi.rods_build(sysm.a, 4.0 * sysm.ecut, False)
    vb = Basis(gpu_ctx, False, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np, sysm.b)
    basis = make_gpu_basis(gpu_ctx, sysm)
    ham = make_gpu_ham(gpu_ctx, basis, sysm)
    cplx = lambda z: torch.from_numpy(np.ascontiguousarray(z).view(np.float64).copy()).cuda()
    vd, pd = cplx(vion), cplx(rhops)
    cd = to_dev(psi)
    ref, v = psi, sysm.v
    for step in range(2):
        e = ham.etrs_step(vb, cd.data_ptr(), basis.ngw, sysm.nst, sysm.occ, 1.0, dt, vd.data_ptr(),
                          pd.data_ptr(), sysm.omega)
        gpu_ctx.sync()
        ref, v, eref = oracle_step(ref, v)
        got = from_dev(cd, basis.ngw, sysm.nst)
        assert rel_err(got, ref) <= FP64_TOL, step
        assert all(abs(a - b) <= 1e-10 for a, b in zip(e, eref)), (step, e, eref)
    # the FP32-bulk step (QBH_FP32 alone selects the whole H) from the same start, against
    # the FP64 result of the first oracle step
    from qball_amd import QBH_ALL, QBH_FP32
    ref1 = oracle_step(psi, sysm.v)[0]
    outs = []
    for fl in (QBH_FP32, QBH_ALL | QBH_FP32):
        v0 = real_to_dev(sysm.v)
        ham.set_vloc(v0.data_ptr())
        gpu_ctx.sync()
        c32 = to_dev(psi)
        ham.etrs_step(vb, c32.data_ptr(), basis.ngw, sysm.nst, sysm.occ, 1.0, dt, vd.data_ptr(),
                      pd.data_ptr(), sysm.omega, flags=fl)
        gpu_ctx.sync()
        outs.append(from_dev(c32, basis.ngw, sysm.nst))
    print("FP32 ETRS step vs FP64 oracle", rel_err(outs[0], ref1))
    assert np.array_equal(outs[0], outs[1])
    assert rel_err(outs[0], ref1) <= 1e-5
    ham.close()
    basis.close()
    vb.close()


@pytest.mark.parametrize("which", ["si8c", "si8"])
def test_current(gpu_ctx, which, request):
    import torch
    from oracle import hpsi as oh
    from oracle.observables import current
    sysm = request.getfixturevalue(which)
    ob, maps, _ = oracle_for(sysm)
    psi = oh.gram(sysm.psi, sysm.real)
    occ = np.array(sysm.occ, dtype=np.float64)
    occ[-1] = 0.0
    A = np.array([0.01, -0.02, 0.03])
    jref, tref = current(maps, psi, occ, 1.0, sysm.omega, A=A, nelectrons=float(occ.sum()))
    basis = make_gpu_basis(gpu_ctx, sysm)
    cd = to_dev(psi)
    jd = torch.zeros(3 * maps.N, dtype=torch.float64, device="cuda")
    tot = basis.current(cd.data_ptr(), basis.ngw, sysm.nst, occ, 1.0, sysm.omega, A=A,
                        nelectrons=float(occ.sum()), j_ptr=jd.data_ptr())
    gpu_ctx.sync()
    basis.close()
    j = jd.cpu().numpy().reshape(3, -1)
    if sysm.real:   # Γ: real orbitals carry no paramagnetic current
        assert np.max(np.abs(j)) <= 1e-12 and np.max(np.abs(tot - A * occ.sum())) <= 1e-12
    else:
        assert rel_err(j, jref) <= FP64_TOL
        assert np.max(np.abs(tot - tref)) <= 1e-12 * np.max(np.abs(tref))


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_hpsi_spd_projectors(gpu_ctx, which, request):
    """s + p + d channels (9 KB projectors per atom, as the Al500 form uses): the (-i)^l
    phases of l = 2 and the 9-column atom blocks."""
    import math
    import torch
    from oracle import hpsi as oh
    from qball_amd import Hamiltonian
    sysm = request.getfixturevalue(which)
    ob, maps, _ = oracle_for(sysm)
    k = ob.kpgx
    g = np.sqrt(ob.kpg2)
    gi = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    x, y, z = (k[:, d] * gi for d in range(3))
    vd = 0.7 * g * g * np.exp(-0.45 * g * g)
    s15 = math.sqrt(15.0 / (4 * math.pi))
    s5 = math.sqrt(5.0 / (16 * math.pi))
    dtab = [s15 * x * y * vd, s15 * y * z * vd, s15 * z * x * vd,
            0.5 * s15 * (x * x - y * y) * vd, s5 * (3 * z * z - 1.0) * vd]
    twnl = np.concatenate([sysm.twnl, np.stack(dtab)])
    lproj = np.array(list(sysm.lproj) + [2] * 5, dtype=np.int32)
    wt = np.array(list(sysm.wt) + [0.8] * 5)
    sp = oh.Projectors(lproj, wt, twnl, sysm.tau)
    basis = make_gpu_basis(gpu_ctx, sysm)
    ham = Hamiltonian(gpu_ctx, basis)
    vdv = real_to_dev(sysm.v)
    ham.set_vloc(vdv.data_ptr())
    ham.set_projectors([sysm.na], [9], lproj, wt, twnl.ravel(), sysm.tau.ravel(), sysm.omega)
    cd = to_dev(sysm.psi)
    hd = torch.zeros_like(cd)
    enl = ham.hpsi(cd.data_ptr(), basis.ngw, sysm.nst, sysm.occ, hd.data_ptr(), basis.ngw,
                   want_enl=True)
    gpu_ctx.sync()
    ref, enl_ref = oh.hpsi(maps, sysm.psi, sysm.v, [sp], sysm.occ, sysm.omega)
    ham.close()
    basis.close()
    assert rel_err(from_dev(hd, basis.ngw, sysm.nst), ref) <= FP64_TOL
    assert abs(enl - enl_ref) <= 1e-10


@pytest.mark.parametrize("which", ["si8", "si8c"])
@pytest.mark.parametrize("nb", [1, 2, 3])
def test_few_bands(gpu_ctx, which, nb, request):
    """One to three bands (at Γ one band is an odd last band alone in its transform): Hψ in
    FP64 and FP32, the density and the non-local forces against the oracle."""
    import torch
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    ob, maps, sp = oracle_for(sysm)
    c = np.ascontiguousarray(sysm.psi[:, :nb])
    occ = sysm.occ[:nb]
    basis = make_gpu_basis(gpu_ctx, sysm)
    ham = make_gpu_ham(gpu_ctx, basis, sysm)
    cd = to_dev(c)
    ref, enl_ref = oh.hpsi(maps, c, sysm.v, [sp], occ, sysm.omega)
    for flags, tol in ((0x7, FP64_TOL), (0x17, FP32_TOL)):
        hd = torch.full_like(cd, float("nan"))
        ham.hpsi(cd.data_ptr(), basis.ngw, nb, occ, hd.data_ptr(), basis.ngw, flags=flags)
        gpu_ctx.sync()
        assert rel_err(from_dev(hd, basis.ngw, nb), ref) <= tol, flags
    rho = torch.zeros(basis.N, dtype=torch.float64, device="cuda")
    basis.density(cd.data_ptr(), basis.ngw, nb, occ, 1.0, sysm.omega, rho.data_ptr())
    gpu_ctx.sync()
    assert rel_err(rho.cpu().numpy(), oh.compute_density(maps, c, occ, 1.0, sysm.omega)) <= FP64_TOL
    f, enl = ham.nl_forces(cd.data_ptr(), basis.ngw, nb, occ, want_enl=True)
    enl_f, fref = oh.nonlocal_forces(ob, [sp], c, occ, sysm.omega)
    assert rel_err(f, fref[0]) <= FP64_TOL and abs(enl - enl_ref) <= 1e-10
    ham.close()
    basis.close()


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_padded_leading_dimensions(gpu_ctx, which, request):
    """qball's ComplexMatrix has mloc >= ngw rows per band (zero padding rows): ψ with
    ld = ngw + 3 and Hψ with ldh = ngw + 7.  Hψ in FP64 and FP32, cleared (padding rows of Hψ
    come out zero) and accumulating (padding rows untouched); density and Gram on the same
    padded ψ."""
    import torch
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    ob, maps, sp = oracle_for(sysm)
    basis = make_gpu_basis(gpu_ctx, sysm)
    ham = make_gpu_ham(gpu_ctx, basis, sysm)
    ngw, nst = basis.ngw, sysm.nst
    ld, ldh = ngw + 3, ngw + 7
    pad = np.zeros((ld, nst), dtype=np.complex128)
    pad[:ngw] = sysm.psi
    cd = to_dev(pad)
    ref = oh.hpsi(maps, sysm.psi, sysm.v, [sp], sysm.occ, sysm.omega)[0]
    for flags, tol in ((0x7, FP64_TOL), (0x17, FP32_TOL)):
        hd = to_dev(np.full((ldh, nst), 5.0 + 3.0j))
        ham.hpsi(cd.data_ptr(), ld, nst, sysm.occ, hd.data_ptr(), ldh, flags=flags)
        gpu_ctx.sync()
        h = hd.cpu().numpy().view(np.complex128).reshape(nst, ldh).T
        assert rel_err(h[:ngw], ref) <= tol, flags
        assert np.all(h[ngw:] == 0), flags                      # cleared output
        start = np.full((ldh, nst), 1.0 - 2.0j)
        hd = to_dev(start)
        ham.hpsi(cd.data_ptr(), ld, nst, sysm.occ, hd.data_ptr(), ldh, flags=flags | 0x8)
        gpu_ctx.sync()
        h = hd.cpu().numpy().view(np.complex128).reshape(nst, ldh).T
        assert rel_err(h[:ngw] - start[:ngw], ref) <= tol, flags
        assert np.all(h[ngw:] == start[ngw:]), flags            # accumulate: untouched
    rho = torch.zeros(basis.N, dtype=torch.float64, device="cuda")
    basis.density(cd.data_ptr(), ld, nst, sysm.occ, 1.0, sysm.omega, rho.data_ptr())
    gpu_ctx.sync()
    assert rel_err(rho.cpu().numpy(), oh.compute_density(maps, sysm.psi, sysm.occ, 1.0,
                                                         sysm.omega)) <= FP64_TOL
    basis.gram(cd.data_ptr(), ld, nst)
    gpu_ctx.sync()
    g = cd.cpu().numpy().view(np.complex128).reshape(nst, ld).T
    assert rel_err(g[:ngw], oh.gram(sysm.psi, sysm.real)) <= FP64_TOL
    assert np.all(g[ngw:] == 0)
    ham.close()
    basis.close()


def test_dense_complex_128_bands(gpu_ctx):
    """Gram and subspace diagonalisation at 128 bands of Si64 in a complex basis (K-split
    band products, rocSOLVER potrf / heevd): C^H C = 1 after Gram; the eigenvalues equal
    numpy's of the projected matrix C^H (HC) built from the GPU's Hψ; the rotated bands stay
    orthonormal."""
    import torch
    from qball_amd.workload import make_system
    sysm = make_system(ncell=2, ecut_ry=25.0, nbands=128, seed=17, complex_basis=True)
    basis = make_gpu_basis(gpu_ctx, sysm)
    ham = make_gpu_ham(gpu_ctx, basis, sysm)
    ngw, nst = basis.ngw, sysm.nst
    cd = to_dev(sysm.psi)
    basis.gram(cd.data_ptr(), ngw, nst)
    gpu_ctx.sync()
    c = from_dev(cd, ngw, nst)
    assert np.max(np.abs(c.conj().T @ c - np.eye(nst))) <= 1e-12
    hd = torch.zeros_like(cd)
    ham.hpsi(cd.data_ptr(), ngw, nst, sysm.occ, hd.data_ptr(), ngw)
    gpu_ctx.sync()
    h = from_dev(hd, ngw, nst)
    hs = c.conj().T @ h
    ref = np.linalg.eigvalsh(0.5 * (hs + hs.conj().T))
    eig = basis.subspace_diag(cd.data_ptr(), ngw, hd.data_ptr(), ngw, nst, eigvec=True)
    gpu_ctx.sync()
    assert np.max(np.abs(np.sort(eig) - ref)) <= 1e-10 * np.max(np.abs(ref))
    c2 = from_dev(cd, ngw, nst)
    assert np.max(np.abs(c2.conj().T @ c2 - np.eye(nst))) <= 1e-11
    ham.close()
    basis.close()
