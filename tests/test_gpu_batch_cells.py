"""The multi-cell workload (12 uneven Γ cells: cubic boxes of 12-28 bohr, 32-48 atoms at
jittered diamond sites, 64-320 bands, ecut 30 Ry, 60^3-120^3 grids) through the library
with every cell's Hψ issued on one of four contexts (streams) before one synchronisation,
as the benchmark's batching section runs them; each cell against the oracle on a band
pair (the first two bands; FP64 1e-12, FP32-bulk 2e-6 relative) and, over all its bands,
FP32-bulk against FP64 (2e-6) -- the cells share the GPU while they run.

The one-graph launch of the same cells (qbh_batch_create) is switched off: see
test_gpu_tuning.py.
"""
import numpy as np
import pytest

from helpers import from_dev, oracle_for, real_to_dev, rel_err, to_dev

pytestmark = pytest.mark.gpu


def test_twelve_uneven_cells_concurrently():
    import torch
    from oracle import hpsi as oh
    from qball_amd import QBH_ALL, QBH_FP32, Basis, Context, Hamiltonian
    from qball_amd.workload import batch_workload, make_batch_cell
    ctxs = [Context(0, None) for _ in range(4)]
    specs = [make_batch_cell(**p) for p in batch_workload()]
    assert len(specs) == 12
    cells = []
    for i, s in enumerate(specs):
        ctx = ctxs[i % 4]
        r = s.rods
        basis = Basis(ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
        ham = Hamiltonian(ctx, basis)
        vd = real_to_dev(s.v)
        ham.set_vloc(vd.data_ptr())
        ham.set_projectors([s.na], [4], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(), s.omega)
        cd = to_dev(s.psi)
        out = {fl: torch.zeros_like(cd) for fl in (QBH_ALL, QBH_ALL | QBH_FP32)}
        cells.append(dict(s=s, ctx=ctx, basis=basis, ham=ham, vd=vd, cd=cd, out=out))
    torch.cuda.synchronize()
    for fl in (QBH_ALL, QBH_ALL | QBH_FP32):
        for c in cells:  # all issued before one synchronisation of every stream
            s = c["s"]
            c["ham"].hpsi(c["cd"].data_ptr(), s.ngw, s.nst, s.occ, c["out"][fl].data_ptr(),
                          s.ngw, flags=fl)
        for ctx in ctxs:
            ctx.sync()
    grids = set()
    for c in cells:
        s = c["s"]
        grids.add(s.np[0])
        h64 = from_dev(c["out"][QBH_ALL], s.ngw, s.nst)
        h32 = from_dev(c["out"][QBH_ALL | QBH_FP32], s.ngw, s.nst)
        assert rel_err(h32, h64) <= 2e-6, s.np
        ob, maps, sp = oracle_for(s)
        ref, _ = oh.hpsi(maps, s.psi[:, :2], s.v, [sp], s.occ[:2], s.omega)
        assert rel_err(h64[:, :2], ref) <= 1e-12, s.np
        assert rel_err(h32[:, :2], ref) <= 2e-6, s.np
        c["ham"].close()
        c["basis"].close()
    assert len(grids) >= 4  # uneven: several grid sizes among the cells
    for ctx in ctxs:
        ctx.close()
