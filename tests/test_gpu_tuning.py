"""Launch knobs, the autotuner and the multi-cell Hψ graph: every choice keeps parity with
the oracle, and graph replays match eager Hψ."""
import numpy as np
import pytest

from helpers import from_dev, oracle_for, real_to_dev, rel_err, to_dev

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _cell(ctx, which, grid=None, seed=0):
    """(system, oracle maps, projectors, gpu basis, gpu ham) for si8 (Γ) or si8c (complex)."""
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from qball_amd import Basis, Hamiltonian
    from qball_amd.workload import make_system
    sysm = make_system(ncell=1, ecut_ry=25.0, nbands=7 if which == "si8" else 6, seed=seed,
                       complex_basis=(which == "si8c"))
    grid = tuple(grid or sysm.np)
    ob, _, sp = oracle_for(sysm)
    maps = FFTMaps(ob, *grid)
    r = sysm.rods
    basis = Basis(ctx, sysm.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], grid, sysm.b)
    rng = np.random.default_rng(seed + 100)
    v = rng.standard_normal(maps.N) * 0.3
    ham = Hamiltonian(ctx, basis)
    vd = real_to_dev(v)
    ham.set_vloc(vd.data_ptr())
    ham._v = vd
    ham.set_projectors([sysm.na], [len(sysm.lproj)], sysm.lproj, sysm.wt, sysm.twnl.ravel(),
                       sysm.tau.ravel(), sysm.omega)
    return dict(sysm=sysm, maps=maps, sp=sp, basis=basis, ham=ham, v=v)


def _ref(c):
    from oracle import hpsi as oh
    s = c["sysm"]
    return oh.hpsi(c["maps"], s.psi, c["v"], [c["sp"]], s.occ, s.omega)[0]


def _run(ctx, c, flags=0x7):
    import torch
    s = c["sysm"]
    cd = to_dev(s.psi)
    hd = torch.zeros_like(cd)
    c["ham"].hpsi(cd.data_ptr(), c["basis"].ngw, s.nst, s.occ, hd.data_ptr(), c["basis"].ngw,
                  flags=flags)
    ctx.sync()
    return from_dev(hd, c["basis"].ngw, s.nst)


@pytest.mark.parametrize("which,grid", [("si8", None), ("si8", (36, 36, 72)), ("si8c", (36, 28, 40))])
def test_every_knob_value_keeps_parity(gpu_ctx, which, grid):
    c = _cell(gpu_ctx, which, grid)
    ref = _ref(c)
    ham = c["ham"]
    # the K split applies to the rocBLAS product (fnl_algo 0)
    cases = [("radix", 1), ("radix", 2), ("radix", 0), ("x_spw", 1), ("x_spw", 3), ("x_spw", 16),
             ("x_spw", 0), ("zy_reg", 0), ("x_reg", 0), ("zy_reg", 1), ("x_reg", 1),
             ("fnl_algo", 0), ("enl_algo", 0),
             ("nsplit", 1), ("nsplit", 3), ("nsplit", 7), ("nsplit", 32), ("nsplit", 0),
             ("overlap", 0), ("overlap", 1), ("overlap", -1),
             ("fnl_algo", 1), ("enl_algo", 1), ("fnl_algo", 16), ("enl_algo", 16),
             ("fnl_algo", 2), ("enl_algo", 2),
             # the library's matrix-core kernels, with their own K slices
             ("fnl_algo", 17), ("enl_algo", 17), ("nsplit", 1), ("nsplit", 5), ("nsplit", 64),
             ("nsplit", 0)]
    for knob, val in cases:
        ham.set_tuning(knob, val)
        assert ham.get_tuning(knob) == val
        got = _run(gpu_ctx, c)
        assert rel_err(got, ref) <= TOL, (knob, val)
    ham.close()
    c["basis"].close()


@pytest.mark.parametrize("flags,tol", [(0x7, TOL), (0x17, 2e-6)])
def test_every_gemm_algorithm_keeps_parity(gpu_ctx, flags, tol):
    """Each projector GEMM algorithm hipBLASLt lists for the shape (F over the full K), and
    the library's own matrix-core kernels (17), give the oracle's Hψ, in FP64 and in the
    FP32 form."""
    c = _cell(gpu_ctx, "si8", None, seed=3)
    ref = _ref(c)
    ham = c["ham"]
    for knob in ("fnl_algo", "enl_algo"):
        for v in range(1, 18):
            ham.set_tuning(knob, v)
            assert rel_err(_run(gpu_ctx, c, flags), ref) <= tol, (knob, v)
        ham.set_tuning(knob, 0)
    ham.set_tuning("fnl_algo", 3)
    ham.set_tuning("enl_algo", 5)
    assert rel_err(_run(gpu_ctx, c, flags), ref) <= tol
    ham.close()
    c["basis"].close()


def test_autotune_keeps_parity_and_caches(gpu_ctx):
    import torch
    c = _cell(gpu_ctx, "si8", (36, 36, 72))
    s = c["sysm"]
    cd = to_dev(s.psi)
    hd = torch.zeros_like(cd)
    rep = c["ham"].autotune(cd.data_ptr(), c["basis"].ngw, s.nst, s.occ, hd.data_ptr(),
                            c["basis"].ngw, reps=2)
    print(rep)
    assert "->" in rep
    # a Γ basis: hipBLASLt lists algorithms for both projector products, and they are tuned
    assert "fnl_algo ->" in rep and "enl_algo ->" in rep
    assert rel_err(_run(gpu_ctx, c), _ref(c)) <= TOL
    rep2 = c["ham"].autotune(cd.data_ptr(), c["basis"].ngw, s.nst, s.occ, hd.data_ptr(),
                             c["basis"].ngw, reps=2)
    assert rep2.startswith("cached")
    c["ham"].close()
    c["basis"].close()


@pytest.mark.skip(reason="qbh_batch_create is switched off: its first GPU run ended in a host "
                         "segmentation fault inside the call (eager warm-up or stream capture); "
                         "the cause is not identified")
@pytest.mark.parametrize("flags", [0x7, 0x17])
def test_batch_graph_replays_match_eager(flags):
    """Three uneven cells in one graph over two branches; each replay (with eager work and
    synchronisation in between) equals the eager Hψ of every cell."""
    import torch
    from qball_amd import Batch, Context
    ctx = Context(0, None)           # a stream of its own: capturable
    cells = [_cell(ctx, "si8", None, 1), _cell(ctx, "si8c", (36, 28, 40), 2),
             _cell(ctx, "si8", (36, 36, 72), 3)]
    eager, ins, outs = [], [], []
    for c in cells:
        s = c["sysm"]
        cd = to_dev(s.psi)
        torch.cuda.synchronize()
        hd = torch.zeros_like(cd)
        c["ham"].hpsi(cd.data_ptr(), c["basis"].ngw, s.nst, s.occ, hd.data_ptr(), c["basis"].ngw,
                      flags=flags)
        ctx.sync()
        eager.append(from_dev(hd, c["basis"].ngw, s.nst))
        ins.append(cd)
        outs.append(torch.zeros_like(cd))
    if flags == 0x7:
        for c, e in zip(cells, eager):
            assert rel_err(e, _ref(c)) <= TOL
    torch.cuda.synchronize()
    batch = Batch(ctx, [(c["ham"], i.data_ptr(), c["basis"].ngw, c["sysm"].nst, o.data_ptr(),
                         c["basis"].ngw) for c, i, o in zip(cells, ins, outs)], flags=flags,
                  nbranch=2)
    ctx.sync()
    scratch = torch.zeros_like(ins[0])
    for rep in range(3):
        for o in outs:
            o.fill_(7.0)
        torch.cuda.synchronize()
        batch.launch()
        ctx.sync()
        for c, o, e in zip(cells, outs, eager):
            got = from_dev(o, c["basis"].ngw, c["sysm"].nst)
            assert np.max(np.abs(got - e)) <= 1e-14 * np.max(np.abs(e)), rep
        # eager work between replays
        c0 = cells[0]
        c0["ham"].hpsi(ins[0].data_ptr(), c0["basis"].ngw, c0["sysm"].nst, c0["sysm"].occ,
                       scratch.data_ptr(), c0["basis"].ngw, flags=flags)
        ctx.sync()
    batch.close()
    for c in cells:
        c["ham"].close()
        c["basis"].close()
    ctx.close()


def test_batch_create_reports_disabled(gpu_ctx):
    """The switched-off graph entry point fails with a clear error instead of running."""
    import ctypes
    from types import SimpleNamespace
    from qball_amd import Batch, QbhError
    with pytest.raises(QbhError, match="disabled"):
        Batch(gpu_ctx, [(SimpleNamespace(h=ctypes.c_void_p()), 16, 1, 1, 32, 1)])


@pytest.mark.parametrize("flags,tol", [(0x7, TOL), (0x17, 2e-6), (0x5, TOL), (0x3, TOL),
                                       (0x13, 2e-6)])
def test_cleared_output_is_never_read(gpu_ctx, flags, tol):
    """Without QBH_ACCUMULATE the output starts cleared: an hpsi full of NaN gives the same
    Hψ as a zeroed one, whichever term writes it first (projector product, forward z pass or
    the clear itself), for both projector GEMM libraries."""
    import torch
    c = _cell(gpu_ctx, "si8", None, seed=6)
    s = c["sysm"]
    ham = c["ham"]
    cd = to_dev(s.psi)
    ngw = c["basis"].ngw
    for algo in (0, 2):
        ham.set_tuning("fnl_algo", algo)
        ham.set_tuning("enl_algo", algo)
        ref = torch.zeros_like(cd)
        ham.hpsi(cd.data_ptr(), ngw, s.nst, s.occ, ref.data_ptr(), ngw, flags=flags)
        got = torch.full_like(cd, float("nan"))
        ham.hpsi(cd.data_ptr(), ngw, s.nst, s.occ, got.data_ptr(), ngw, flags=flags)
        gpu_ctx.sync()
        assert bool(torch.isfinite(got).all()), (flags, algo)
        a, b = from_dev(got, ngw, s.nst), from_dev(ref, ngw, s.nst)
        assert rel_err(a, b) <= 1e-15, (flags, algo)
        if flags & 0x4 and flags & 0x2 and flags & 0x1:
            assert rel_err(a, _ref(c)) <= tol, (flags, algo)
    ham.close()
    c["basis"].close()


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_small_workspace_chunks_give_identical_results(which):
    """A transform workspace that holds only two FP64 slots splits every pass into many
    chunks (and changes the slots per x-pass workgroup); each slot's arithmetic is
    unchanged, so Hψ (FP64 and FP32 form) and the transforms are bitwise the same as with
    one chunk, and the density (summed over chunks) equal to rounding."""
    import torch
    from qball_amd import Context
    ctx = Context(0, torch.cuda.current_stream().cuda_stream)
    c = _cell(ctx, which, None, seed=8)
    s = c["sysm"]
    basis = c["basis"]
    ngw = basis.ngw
    n0, n1, n2 = (int(x) for x in s.np)
    N = n0 * n1 * n2
    cd = to_dev(s.psi)

    def run_all():
        out = []
        for fl in (0x7, 0x17):
            hd = torch.zeros_like(cd)
            c["ham"].hpsi(cd.data_ptr(), ngw, s.nst, s.occ, hd.data_ptr(), ngw, flags=fl)
            out.append(hd)
        rho = torch.zeros(N, dtype=torch.float64, device="cuda")
        basis.density(cd.data_ptr(), ngw, s.nst, s.occ, 1.0, s.omega, rho.data_ptr())
        f = torch.zeros(2 * N * s.nst, dtype=torch.float64, device="cuda")
        basis.backward(cd.data_ptr(), ngw, s.nst, 0, f.data_ptr())
        back = torch.zeros_like(cd)
        basis.forward(f.data_ptr(), s.nst, 0, back.data_ptr(), ngw)
        ctx.sync()
        return out + [rho, f, back]

    ref = run_all()
    slot = 16 * (basis.ncol * n2 + n2 * basis.nlx * n1)   # FP64 T1 + T2 of one slot
    ctx.set_workspace_cap(2 * slot + 1)
    got = run_all()
    for k, (a, b) in enumerate(zip(got, ref)):
        if k == 2:  # the density sums the chunks in a different grouping
            assert float((a - b).abs().max()) <= 1e-14 * float(b.abs().max()), k
        else:
            assert torch.equal(a, b), k
    c["ham"].close()
    basis.close()
    ctx.close()


@pytest.mark.parametrize("flags,tol", [(0x7, TOL), (0x17, 2e-6)])
def test_wide_sphere_uses_full_line_count(gpu_ctx, flags, tol):
    """A sphere spanning most of the x edge (8-atom Si at 64 Ry on a 36^3 grid: 27 of 36 x
    positions hold lines) takes the x-pass variant with the full line-element count per
    thread; Hψ still matches the oracle."""
    import torch
    from oracle import hpsi as oh
    from oracle.basis import FFTMaps
    from qball_amd import Basis, Hamiltonian
    from qball_amd.workload import make_system
    s = make_system(ncell=1, ecut_ry=64.0, nbands=6, seed=9)
    grid = (36, 36, 36)
    ob, _, sp = oracle_for(s)
    maps = FFTMaps(ob, *grid)
    r = s.rods
    basis = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], grid, s.b)
    assert basis.nlx > 36 // 2 + 8, basis.nlx
    ham = Hamiltonian(gpu_ctx, basis)
    v = np.random.default_rng(19).standard_normal(maps.N) * 0.3
    vd = real_to_dev(v)
    ham.set_vloc(vd.data_ptr())
    ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(),
                       s.omega)
    cd = to_dev(s.psi)
    hd = torch.zeros_like(cd)
    ham.hpsi(cd.data_ptr(), basis.ngw, s.nst, s.occ, hd.data_ptr(), basis.ngw, flags=flags)
    gpu_ctx.sync()
    ref = oh.hpsi(maps, s.psi, v, [sp], s.occ, s.omega)[0]
    assert rel_err(from_dev(hd, basis.ngw, s.nst), ref) <= tol
    ham.close()
    basis.close()


def test_stream_bandwidth_is_plausible(gpu_ctx):
    """The stream-bandwidth diagnostic returns read/write/copy rates between 1 TB/s and the
    8 TB/s HBM peak, and rejects sizes outside its range."""
    from qball_amd import QbhError
    bw = gpu_ctx.stream_bandwidth(0.5, 2)
    print(bw)
    for k in ("read", "write", "copy"):
        assert 1000.0 < bw[k] < 8000.0 * 1.02, (k, bw)
    with pytest.raises(QbhError):
        gpu_ctx.stream_bandwidth(1e-4, 1)
