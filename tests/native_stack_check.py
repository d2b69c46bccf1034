"""Hψ parity on the system ROCm libraries, for test_gpu_native_stack.py.

A process that never imports torch, so libqbh.so runs on the HIP runtime, rocBLAS and
hipBLASLt it was linked against (as in a C++ host), not on the copies bundled with torch.
Checks Hψ of an 8-atom Si cell (Γ, and a complex basis) against the oracle for the built-in
projector GEMM choice (the library's matrix-core kernels), rocBLAS, and every algorithm
hipBLASLt lists, in FP64 and in the FP32
form, into an output buffer full of NaN (which Hψ without accumulation must not read).
Prints one JSON line with the errors.

With --fullsize: Hψ of the benchmark's two full-size cells (Si512 Γ, 1024 bands; fcc Al500
Γ, 900 bands) with all bands at once, so the projector products run at the shapes the
benchmark times, for the built-in launch choices, the autotuner's choice, rocBLAS and
hipBLASLt's second listed algorithm (FP64) and the built-in and autotuned FP32 form; the
first and last band pairs against the oracle (1e-12; FP32 2e-6) and the FP32 form against
FP64 over all bands.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def main():
    from helpers import oracle_for, rel_err
    from oracle import hpsi as oh
    from oracle.basis import FFTMaps
    from qball_amd import Basis, Context, Hamiltonian
    from qball_amd.devmem import DeviceArray
    from qball_amd.workload import make_system

    print("native stack check: start", file=sys.stderr, flush=True)
    ctx = Context(0, None)
    out = {"libraries": ctx.library_info(), "cases": []}
    for complex_basis in (False, True):
        s = make_system(ncell=1, ecut_ry=25.0, nbands=7, seed=4, complex_basis=complex_basis)
        ob, _, sp = oracle_for(s)
        maps = FFTMaps(ob, *s.np)
        r = s.rods
        basis = Basis(ctx, s.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
        ham = Hamiltonian(ctx, basis)
        v = np.random.default_rng(11).standard_normal(maps.N) * 0.3
        vd = DeviceArray.from_numpy(v)
        ham.set_vloc(vd.ptr)
        ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(),
                           s.omega)
        ref = oh.hpsi(maps, s.psi, v, [sp], s.occ, s.omega)[0]
        ngw = basis.ngw
        psi = DeviceArray.from_numpy(np.ascontiguousarray(s.psi.T).view(np.float64).ravel())
        hp = DeviceArray(psi.nbytes)

        nan = np.full(psi.nbytes // 8, np.nan)

        def run(flags):
            hp.upload(nan)  # a cleared output is never read: NaN must not reach the result
            ham.hpsi(psi.ptr, ngw, s.nst, s.occ, hp.ptr, ngw, flags=flags)
            ctx.sync()
            got = hp.to_numpy().view(np.complex128).reshape(s.nst, ngw).T
            return rel_err(got, ref)

        for flags in (0x7, 0x17):
            out["cases"].append({"complex": complex_basis, "flags": flags, "knob": "builtin",
                                 "value": -1, "err": run(flags)})
            for knob in ("fnl_algo", "enl_algo"):
                for val in range(0, 18):
                    ham.set_tuning(knob, val)
                    out["cases"].append({"complex": complex_basis, "flags": flags, "knob": knob,
                                         "value": val, "err": run(flags)})
                ham.set_tuning(knob, -1)
        ham.close()
        basis.close()
        for d in (psi, hp, vd):
            d.close()
    ctx.close()
    print("native stack check: done", file=sys.stderr, flush=True)
    print(json.dumps(out))
    assert "torch" not in sys.modules


def fullsize():
    from helpers import oracle_for, rel_err
    from oracle import hpsi as oh
    from qball_amd import Basis, Context, Hamiltonian
    from qball_amd.devmem import DeviceArray
    from qball_amd.workload import make_fcc_al, make_system

    ctx = Context(0, None)
    out = {"libraries": ctx.library_info(), "cases": []}
    knobs = ("zy_reg", "x_reg", "x_spw", "radix", "nsplit", "overlap", "fnl_algo", "enl_algo")
    for which in ("si512", "al500"):
        s = make_system(ncell=4, ecut_ry=25.0, seed=0) if which == "si512" else make_fcc_al(seed=1)
        print(f"native stack check: {which} ngw {s.ngw} nst {s.nst}", file=sys.stderr, flush=True)
        r = s.rods
        basis = Basis(ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
        ham = Hamiltonian(ctx, basis)
        vd = DeviceArray.from_numpy(np.ascontiguousarray(s.v))
        ham.set_vloc(vd.ptr)
        ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(),
                           s.omega)
        ngw, nst = s.ngw, s.nst
        sub = [0, 1, nst - 2, nst - 1]
        _, maps, sp = oracle_for(s)
        ref = oh.hpsi(maps, s.psi[:, sub], s.v, [sp], s.occ[sub], s.omega)[0]
        psi = DeviceArray.from_numpy(np.ascontiguousarray(s.psi.T).view(np.float64).ravel())
        hp = DeviceArray(psi.nbytes)
        h64 = None
        for flags in (0x7, 0x17):
            settings = [("builtin", {"zy_reg": 1, "x_reg": 1, "x_spw": 0, "radix": 0,
                                     "nsplit": 0, "overlap": 0, "fnl_algo": -1,
                                     "enl_algo": -1})]
            if flags == 0x7:
                settings += [("rocblas", {"fnl_algo": 0, "enl_algo": 0}),
                             ("hipblaslt-2", {"fnl_algo": 2, "enl_algo": 2})]
            settings.append(("autotuned", None))
            for name, choice in settings:
                if choice is None:
                    ham.autotune(psi.ptr, ngw, nst, s.occ, hp.ptr, ngw, flags=flags, reps=2,
                                 force=True)
                else:
                    for k, v in choice.items():
                        ham.set_tuning(k, v)
                hp.upload(np.full(psi.nbytes // 8, np.nan))
                ham.hpsi(psi.ptr, ngw, nst, s.occ, hp.ptr, ngw, flags=flags)
                ctx.sync()
                got = hp.to_numpy().view(np.complex128).reshape(nst, ngw).T
                case = {"cell": which, "flags": flags, "choice": name,
                        "knobs": {k: ham.get_tuning(k) for k in knobs},
                        "err": rel_err(got[:, sub], ref)}
                if flags == 0x7 and name == "builtin":
                    h64 = got.copy()
                elif flags == 0x17:
                    case["err_vs_fp64"] = rel_err(got, h64)
                out["cases"].append(case)
                print(f"native stack check: {json.dumps(case)}", file=sys.stderr, flush=True)
        ham.close()
        basis.close()
        for d in (psi, hp, vd):
            d.close()
    ctx.close()
    print(json.dumps(out))
    assert "torch" not in sys.modules


if __name__ == "__main__":
    if "--fullsize" in sys.argv[1:]:
        fullsize()
    else:
        main()
