"""Parity of Hψ on the system ROCm libraries (HIP runtime, rocBLAS, hipBLASLt as a C++ host
links them), checked in a child process that never imports torch: every projector GEMM
choice (the library's matrix-core kernels, rocBLAS and each hipBLASLt algorithm) gives the
oracle's Hψ."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_hpsi_parity_on_system_rocm_libraries():
    p = subprocess.run([sys.executable, "-u", os.path.join(HERE, "native_stack_check.py")],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-3000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    print(out["libraries"])
    assert "hipblaslt" in out["libraries"]
    # the built-in projector GEMMs are the library's own matrix-core kernels (Γ)
    assert "projector-gemm-builtin mfma" in out["libraries"], out["libraries"]
    assert len(out["cases"]) == 2 * 2 * (1 + 2 * 18)
    for c in out["cases"]:
        tol = 1e-12 if c["flags"] == 0x7 else 2e-6
        assert c["err"] <= tol, c


def test_fullsize_parity_on_system_rocm_libraries():
    """The projector products at the benchmark's own shapes (Si512: 2048 projectors x 72919
    plane waves x 1024 bands; Al500: 4500 x 77666 x 900), torch-free: built-in, rocBLAS,
    hipBLASLt's second algorithm and the autotuned choice against the oracle."""
    p = subprocess.run([sys.executable, "-u", os.path.join(HERE, "native_stack_check.py"),
                        "--fullsize"], capture_output=True, text=True, timeout=900)
    assert p.returncode == 0, p.stderr[-3000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    for c in out["cases"]:
        print(c)
    assert len(out["cases"]) == 2 * (4 + 2)
    for c in out["cases"]:
        tol = 1e-12 if c["flags"] == 0x7 else 2e-6
        assert c["err"] <= tol, c
        if c["flags"] == 0x17:
            assert c["err_vs_fp64"] <= 2e-6, c
    # the built-in choice (-1) of a Γ cell is the library's matrix-core kernels
    assert "projector-gemm-builtin mfma" in out["libraries"], out["libraries"]
    builtin = [c for c in out["cases"] if c["choice"] == "builtin"]
    assert all(c["knobs"]["fnl_algo"] == -1 and c["knobs"]["enl_algo"] == -1 for c in builtin)


def test_library_info_in_a_torch_process(gpu_ctx):
    """In a process that imported torch first the library runs on torch's bundled HIP/BLAS
    libraries; the report names their versions and the built-in GEMM library for them."""
    info = gpu_ctx.library_info()
    print(info)
    assert info.startswith("hip ") and " rocblas " in info and " hipblaslt " in info
    assert "projector-gemm-builtin mfma" in info
