"""qbh_fft_backward_noalloc (the backward transform that allocates nothing: its workspace is
a file-scope device array) against qbh_fft_backward through the C ABI, on two calls in a
row where the second has a smaller input, in FP64 and FP32, pair and single-band modes, Γ
and complex bases; and the rejection of a grid whose single transform slot exceeds the
array.

Bar: the same kernels run on the same slots, so the results must agree to 1e-14 of the
largest magnitude (they are bitwise equal in practice); qbh_fft_backward itself is checked
against the oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

from helpers import make_gpu_basis, to_dev

pytestmark = pytest.mark.gpu


def _grid(basis, nf):
    import torch
    return torch.full((2 * basis.N * nf,), float("nan"), dtype=torch.float64, device="cuda")


def _compare(ctx, basis, psi, nbands, pair, flags):
    cd = to_dev(psi[:, :nbands])
    nf = (nbands + 1) // 2 if pair else nbands
    ref, got = _grid(basis, nf), _grid(basis, nf)
    basis.backward(cd.data_ptr(), basis.ngw, nbands, pair, ref.data_ptr(), flags)
    basis.backward_noalloc(cd.data_ptr(), basis.ngw, nbands, pair, got.data_ptr(), flags)
    ctx.sync()
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    assert np.isfinite(r).all() and np.isfinite(g).all()
    return float(np.max(np.abs(g - r)) / np.max(np.abs(r)))


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_noalloc_backward_matches_allocating_entry_point(gpu_ctx, which):
    from qball_amd.workload import make_system
    s = make_system(ncell=1, ecut_ry=25.0, nbands=15, seed=51, complex_basis=(which == "si8c"))
    basis = make_gpu_basis(gpu_ctx, s)
    modes = [True, False] if s.real else [False]
    for pair in modes:
        for flags in (0x0, 0x10):
            # two calls in a row on the shared scratch, the second with fewer bands
            for nb in (15, 4):
                err = _compare(gpu_ctx, basis, s.psi, nb, pair, flags)
                assert err <= 1e-14, (which, pair, flags, nb, err)
    basis.close()


def test_noalloc_backward_in_chunks(gpu_ctx):
    """A band count whose slots exceed the scratch runs in several chunks (Si512 grid,
    28 MB per FP64 slot: 9 slots per chunk), then a smaller one."""
    from qball_amd.workload import make_system
    s = make_system(ncell=4, ecut_ry=25.0, nbands=40, seed=52)
    basis = make_gpu_basis(gpu_ctx, s)
    for nb, flags in ((40, 0x0), (7, 0x0), (40, 0x10)):
        err = _compare(gpu_ctx, basis, s.psi, nb, True, flags)
        assert err <= 1e-14, (nb, flags, err)
    basis.close()


def test_noalloc_backward_rejects_a_slot_larger_than_the_scratch(gpu_ctx):
    """A 256^3 grid filled by the sphere (4.1 M plane waves): one FP64 slot needs 445 MB,
    more than the 256 MiB array."""
    import torch
    from qball_amd import Basis, QbhError, capi
    from qball_amd.workload import cell_vectors
    a = np.diag([40.0] * 3)
    r = capi.rods_build(a, 195.0, True)
    _, b = cell_vectors(a)
    basis = Basis(gpu_ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"],
                  tuple(r["np"]), b)
    # valid buffers of the right sizes, although the call must refuse before any launch
    cd = torch.zeros(2 * basis.ngw * 2, dtype=torch.float64, device="cuda")
    fd = torch.zeros(2 * basis.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(QbhError, match="no-allocation scratch"):
        basis.backward_noalloc(cd.data_ptr(), basis.ngw, 2, True, fd.data_ptr(), 0)
    basis.close()
