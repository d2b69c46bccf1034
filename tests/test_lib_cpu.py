"""libqbh.so on the CPU: loads, exports the header's API, host-side logic matches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle.basis import Basis, UnitCell, density_grid, factorizable

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "qbh.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qbh_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_function():
    from qball_amd import capi
    L = capi.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in capi.SIGNATURES, n


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from qball_amd import capi
    h = ctypes.c_void_p()
    rc = capi.lib().qbh_ctx_create(0, None, ctypes.byref(h))
    assert rc != 0 and not h.value


@pytest.mark.parametrize("a,ecut,real", [
    (np.diag([20.52] * 3), 12.5, True),
    (np.diag([10.26] * 3), 12.5, False),
    (np.array([[7.1, 0.3, -0.2], [0.9, 6.4, 0.5], [-0.4, 1.1, 8.3]]), 9.0, True),
    (np.array([[7.1, 0.3, -0.2], [0.9, 6.4, 0.5], [-0.4, 1.1, 8.3]]), 9.0, False),
    (np.array([[0.0, 3.825, 3.825], [3.825, 0.0, 3.825], [3.825, 3.825, 0.0]]), 15.0, True),
])
def test_rods_build_matches_oracle(a, ecut, real):
    from qball_amd import capi
    r = capi.rods_build(a, ecut, real=real)
    ob = Basis(UnitCell(a), ecut, force_complex=not real)
    for k in ("rod_h", "rod_k", "rod_lmin", "rod_size"):
        assert np.array_equal(r[k], getattr(ob, k)), k
    assert r["np"] == ob.np


def test_factorizable_matches_oracle():
    from qball_amd import capi
    for n in range(1, 3000):
        assert capi.factorizable(n) == factorizable(n)


def test_workload_density_grid_matches_oracle():
    from qball_amd.workload import density_grid_np
    a = np.diag([20.52] * 3)
    assert density_grid_np(a, 12.5) == density_grid(UnitCell(a), 12.5)[1]
