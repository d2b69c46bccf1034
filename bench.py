#!/usr/bin/env python3
"""Hψ throughput benchmark: band-applications/s of a full Hamiltonian application
(non-local KB projectors + kinetic + local potential via the G<->r transforms) on diamond
Si512 at Γ (a = 41.04 bohr, ecut 25 Ry: 72919 plane waves, 1024 bands, 140^3 grid,
512 x 4 projectors of the Si HSCV LDA pseudopotential), one MI355X.

One step = one Hψ over all 1024 bands (qbh_hpsi with QBH_ALL), inputs resident in HBM.
The projector table is generated on the first (warm-up) application and reused, as the
ions do not move.  Prints one JSON line.

The process does not import torch: it drives libqbh.so through its C ABI with buffers from
the HIP runtime libqbh.so links (qball_amd.devmem), so the library runs on the system's
ROCm libraries exactly as a C++ host linking it does (a process that imports torch first
binds it to the HIP runtime and BLAS libraries bundled with torch instead).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SI_XML = os.path.join(ROOT, "tests", "golden", "Si_HSCV_LDA-1.0.xml")
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ncell", type=int, default=4, help="diamond cells per edge (4 = Si512)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-local-bands", type=int, default=1024,
                   help="bands of the CPU baseline's local + kinetic term (>= nst: all; "
                        "fewer are timed and scaled to nst)")
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-batch", action="store_true", help="skip the 12-cell batch section")
    p.add_argument("--batch-reps", type=int, default=10)
    p.add_argument("--no-large", action="store_true", help="skip the Al500 large-cell section")
    p.add_argument("--no-td", action="store_true",
                   help="skip the Si64 TDDFT forward/adjoint section (differentiation form)")
    p.add_argument("--no-scf", action="store_true",
                   help="skip the Si512 SCF section (precision form against FP64)")
    p.add_argument("--scf-cycles", type=int, default=120)
    p.add_argument("--smeared-cycles", type=int, default=8,
                   help="Fermi-smeared SCF cycles of the Al500 section")
    p.add_argument("--scf-nite", type=int, default=3)
    p.add_argument("--scf-no-diag", dest="scf_diag", action="store_false",
                   help="no subspace diagonalisation after the SCF cycles (default: wf_diag T)")
    p.add_argument("--scf-alpha", type=float, default=0.5, help="charge mixing coefficient")
    p.add_argument("--scf-ndim", type=int, default=3, help="Anderson history (<= 8)")
    p.add_argument("--scf-kerker", type=float, default=10.0, help="Kerker rc (bohr)")
    p.add_argument("--scf-ecutprec", type=float, default=1.5,
                   help="PSDA preconditioner cutoff in Hartree (0: the wavefunction cutoff)")
    p.add_argument("--batch-streams", type=int, default=4,
                   help="contexts (HIP streams) the 12 batch cells are spread over")
    p.add_argument("--no-tune", action="store_true",
                   help="built-in launch choices, sequential schedule (short profiled runs)")
    p.add_argument("--calibrate-gib", type=float, default=0.0,
                   help="only fill, then sum, a buffer of this many GiB with torch and exit: "
                        "kernels of known HBM bytes that calibrate the byte counters")
    return p.parse_args()


def calibrate(gib):
    """A fill (writes exactly `gib` GiB) and a sum (reads exactly `gib` GiB) for the
    rocprofv3 WRITE_SIZE / FETCH_SIZE calibration; a process of its own, with torch."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    s = float(a.sum())
    torch.cuda.synchronize()
    print(json.dumps({"calibration_gib": gib, "sum": s}))


def kernel_bytes(sysm, nslots, prec_bytes):
    """Algorithmic HBM bytes per launch of each transform kernel for a chunk of nslots
    band pairs (complex element size prec_bytes; psi / hpsi stay FP64)."""
    n0, n1, n2 = sysm.np
    ngw, ncol, nlx = sysm.ngw, sysm.ncol, sysm.nlx
    s = prec_bytes
    t1 = ncol * n2 * s
    t2 = n2 * nlx * n1 * s
    psi_pair = 2 * ngw * 16
    N = n0 * n1 * n2
    return {
        "zpass_bwd": nslots * (psi_pair + t1),
        "ypass_bwd": nslots * (t1 + t2),
        "xpass_vmul": nslots * (2 * t2) + N * 8,
        "ypass_fwd": nslots * (t2 + t1),
        "zpass_fwd": nslots * (t1 + psi_pair + 2 * psi_pair),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the Hψ path shards by bands; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_gib > 0:
        calibrate(args.calibrate_gib)
        return
    from qball_amd import Basis, Context, Hamiltonian, QBH_ALL, QBH_FP32, Species, auto_rcps
    from qball_amd.devmem import DeviceArray, synchronize
    from qball_amd.workload import make_system

    t_setup = time.time()
    sysm = make_system(ncell=args.ncell, ecut_ry=25.0, seed=0)
    ctx = Context(0, None)
    libinfo = ctx.library_info()
    log(libinfo)
    r = sysm.rods
    basis = Basis(ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np, sysm.b)
    sysm.ncol, sysm.nlx = basis.ncol, basis.nlx
    ham = Hamiltonian(ctx, basis)
    v = DeviceArray.from_numpy(sysm.v)
    ham.set_vloc(v.ptr)
    # the Si HSCV LDA pseudopotential with the Ewald width AtomSet::set_rcps picks; the
    # projector tables are made on the device (update_twnl)
    species = Species(SI_XML, rcps=auto_rcps(4.0, sysm.a, sysm.ecut))
    ham.set_species([species], [sysm.na], sysm.tau, sysm.omega)
    ngw, nst = sysm.ngw, sysm.nst
    psi = DeviceArray.from_numpy(np.ascontiguousarray(sysm.psi.T).view(np.float64).ravel())
    del sysm.psi
    basis.gram(psi.ptr, ngw, nst)              # orthonormalise (SlaterDet::gram on the GPU)
    hpsi = DeviceArray(psi.nbytes)
    ctx.sync()
    t_setup = time.time() - t_setup
    log(f"setup done ({t_setup:.1f} s)")

    def step(flags):
        ham.hpsi(psi.ptr, ngw, nst, sysm.occ, hpsi.ptr, ngw, flags=flags)

    for _ in range(args.warmup):
        step(QBH_ALL)
    synchronize()

    # launch choices for this shape from the library's autotuner (untimed): transform
    # variants, x-pass blocking, GEMM/transform overlap and the projector K-split count
    knobs = ("zy_reg", "x_reg", "x_spw", "radix", "nsplit", "overlap", "fnl_algo", "enl_algo")
    builtin = {"zy_reg": 1, "x_reg": 1, "x_spw": 0, "radix": 0, "nsplit": 0, "overlap": 0,
               "fnl_algo": -1, "enl_algo": -1}

    def apply(choice):
        for k, v in choice.items():
            ham.set_tuning(k, v)

    def tune(flags):
        rep = ham.autotune(psi.ptr, ngw, nst, sysm.occ, hpsi.ptr, ngw, flags=flags,
                           reps=2, force=True)
        return {"choice": {k: ham.get_tuning(k) for k in knobs},
                "report": [ln for ln in rep.splitlines() if "->" in ln or ln.startswith("start")]}

    def timed(flags, steps, kernel_events=False):
        """(s per step, per-kernel (total ms, launches)) over `steps` Hψ applications.  The
        per-kernel HIP events are recorded only when asked: the step time of a run that
        reports them includes their cost, so the headline time comes from a run without."""
        ctx.timing_report()                    # discard earlier totals
        ctx.set_timing(kernel_events)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(flags)
        synchronize()
        t1 = time.perf_counter()
        ctx.set_timing(False)
        return (t1 - t0) / steps, ctx.timing_report()

    if args.no_tune:
        apply(builtin)
        sched64 = {"choice": dict(builtin)}
    else:
        sched64 = tune(QBH_ALL)
    log("fp64 launch choices: " + json.dumps(sched64["choice"]))
    dt, _ = timed(QBH_ALL, args.steps)
    log(f"fp64 timed: {dt * 1e3:.2f} ms/step")
    value = nst / dt

    # energy-like check of this application: band sum of <psi|H|psi> (Γ: 2 Re - G=0 term)
    psi_h = psi.to_numpy().reshape(nst, 2 * ngw)

    def band_energy(h):
        d = 2.0 * np.einsum("ij,ij->i", psi_h, h) - psi_h[:, 0] * h[:, 0]
        return float(np.dot(sysm.occ, d))
    h64 = hpsi.to_numpy().reshape(nst, 2 * ngw)
    e64 = band_energy(h64)

    # kernel characterisation: with the GEMMs concurrent to the transforms each kernel's
    # duration includes the other's share of the GPU, so the per-kernel rooflines come from
    # the built-in launch choices on the sequential schedule (what `--no-tune` times, and
    # what the committed rocprof summary of `bench.py --no-tune` shows)
    char_steps = max(2, args.steps)
    apply(builtin)
    _, kchar = timed(QBH_ALL, char_steps, kernel_events=True)
    apply(sched64["choice"])
    traffic = load_traffic()
    nslots_total = (nst + 1) // 2
    chunk_bytes = kernel_bytes(sysm, 1, 16)
    flops_gemm = 2.0 * (2 * ngw) * nst * (4 * sysm.na)  # one GEMM: 2*M*N*K
    per_kernel = {}
    for name, (ms, cnt) in kchar.items():
        calls = cnt / char_steps
        avg_ms = ms / cnt
        entry = {"avg_ms": avg_ms, "launches_per_step": calls}
        if name in chunk_bytes:
            # launches per step = chunks; bytes scale with the slots in each chunk
            b_step = kernel_bytes(sysm, nslots_total, 16)[name]
            b_launch = b_step / calls
            entry["bytes_per_launch"] = b_launch
            entry["GBps"] = b_launch / (avg_ms * 1e-3) / 1e9
        if name in ("fnl_gemm", "enl_gemm"):
            entry["flops_per_launch"] = flops_gemm
            entry["TFLOPs"] = flops_gemm / (avg_ms * 1e-3) / 1e12
        if name in traffic:
            entry["hbm_bytes_per_launch"] = traffic[name]
        per_kernel[name] = entry
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"] * per_kernel[k]["launches_per_step"])
    d = per_kernel[dom]
    if "TFLOPs" in d:
        roof = {"kernel": dom, "bound": "mfma", "achieved": d["TFLOPs"], "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": d["TFLOPs"] / PEAK_FP64_TFLOPS}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": d.get("GBps"), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": (d.get("GBps") or 0) / PEAK_HBM_GBS}
    roof["traffic"] = d.get("hbm_bytes_per_launch")
    # FFT stages (the five transform passes) as one: time per Hψ, the bandwidth on this
    # pipeline's algorithmic bytes, on the counted DRAM bytes, and the "equivalent" one on
    # the unfused pruned-pencil model of qball's FourierTransform (per Γ pair:
    # 2(2 nvec np2 s) z + 2(4 ntrans0 np0 np2 s) pruned x + 2(2 N s) y + 2(nvec np2 + N) s
    # zvec<->grid + (2 N s + N r) xv + 6 ngw s coefficients in / Hψ RMW)
    fft_names = ("zpass_bwd", "ypass_bwd", "xpass_vmul", "ypass_fwd", "zpass_fwd")
    t_fft = sum(per_kernel[k]["avg_ms"] * per_kernel[k]["launches_per_step"] for k in fft_names)
    b_fft = sum(per_kernel[k]["bytes_per_launch"] * per_kernel[k]["launches_per_step"]
                for k in fft_names)
    b_dram = (sum(traffic[k] * per_kernel[k]["launches_per_step"] for k in fft_names)
              if all(k in traffic for k in fft_names) else None)
    n0, n1, n2 = sysm.np
    N = n0 * n1 * n2
    nvec, ntrans0 = sysm.ncol, (sysm.nlx + 1) // 2
    b_pair = (2 * (2 * nvec * n2 * 16) + 2 * (4 * ntrans0 * n0 * n2 * 16) + 2 * (2 * N * 16)
              + 2 * (nvec * n2 + N) * 16 + (2 * N * 16 + N * 8) + 6 * ngw * 16)
    b_model = b_pair * nslots_total
    fft = {"ms_per_hpsi": t_fft,
           "algorithmic_bytes": b_fft, "GBps": b_fft / t_fft / 1e6,
           "frac_of_peak": b_fft / t_fft / 1e6 / PEAK_HBM_GBS,
           "dram_bytes": b_dram,
           "dram_GBps": (b_dram / t_fft / 1e6) if b_dram else None,
           "pruned_pencil_model_bytes": b_model,
           "equivalent_GBps": b_model / t_fft / 1e6,
           "equivalent_frac_of_peak": b_model / t_fft / 1e6 / PEAK_HBM_GBS,
           "note": "five transform passes of one FP64 Hψ, sequential characterisation pass; "
                   "dram_bytes from the committed FETCH/WRITE counter runs"}
    # what plain streaming kernels reach on this GPU (2 GiB buffers): the practical ceiling
    # for the HBM rate of a pass that reads and writes
    sb = ctx.stream_bandwidth(2.0, 3)
    fft["stream_kernels_GBps"] = sb
    fft["frac_of_stream_copy"] = fft["GBps"] / sb["copy"]
    fft["dram_frac_of_stream_copy"] = (fft["dram_GBps"] / sb["copy"]) if b_dram else None
    log(f"stream kernels: read {sb['read']:.0f} write {sb['write']:.0f} copy {sb['copy']:.0f} GB/s")
    roof["measured_in"] = ("sequential characterisation pass at the built-in launch choices, "
                           f"{char_steps} steps, HIP events on the library stream")

    # FP32-bulk form
    fp32 = None
    if not args.no_fp32:
        for _ in range(args.warmup):
            step(QBH_ALL | QBH_FP32)
        if args.no_tune:
            apply(builtin)
            sched32 = {"choice": dict(builtin)}
        else:
            sched32 = tune(QBH_ALL | QBH_FP32)
        dt32, _ = timed(QBH_ALL | QBH_FP32, args.steps)
        log(f"fp32 timed: {dt32 * 1e3:.2f} ms/step")
        h32 = hpsi.to_numpy().reshape(nst, 2 * ngw)
        e32 = band_energy(h32)
        rel = float(np.linalg.norm(h32 - h64) / np.linalg.norm(h64))
        del h32
        # density L1 error of the FP32-bulk transforms (FP64 band sum) against FP64:
        # (Omega/N) sum |rho32 - rho64| in electrons, and relative to N_e
        N = int(np.prod(sysm.np))
        rho = DeviceArray(8 * N)
        dens = {}
        for name, fl in (("fp64", 0), ("fp32", QBH_FP32)):
            rho.zero()
            basis.density(psi.ptr, ngw, nst, sysm.occ, 1.0, sysm.omega, rho.ptr, flags=fl)
            ctx.sync()
            dens[name] = rho.to_numpy()
        rho.close()
        dl1 = float(np.abs(dens["fp32"] - dens["fp64"]).sum() * sysm.omega / N)
        ne = float(dens["fp64"].sum() * sysm.omega / N)
        del dens
        # non-local ionic forces (the force term the projector products feed), FP32 form
        # against FP64 over all bands; one untimed call each, then one timed
        forces = {}
        for name, f32 in (("fp64", False), ("fp32", True)):
            ham.nl_forces(psi.ptr, ngw, nst, sysm.occ, fp32=f32)
            t0 = time.perf_counter()
            forces[name] = ham.nl_forces(psi.ptr, ngw, nst, sysm.occ, fp32=f32)
            forces[name + "_ms"] = (time.perf_counter() - t0) * 1e3
        fdiff = forces["fp32"] - forces["fp64"]
        force_err = {
            "max_abs_err_Ha_per_bohr": float(np.max(np.abs(fdiff))),
            "rel_err": float(np.linalg.norm(fdiff) / np.linalg.norm(forces["fp64"])),
            "max_abs_force_Ha_per_bohr": float(np.max(np.abs(forces["fp64"]))),
            "call_ms": {"fp64": forces["fp64_ms"], "fp32": forces["fp32_ms"]},
            "note": "non-local ionic forces (qbh_nl_forces, all atoms and bands); this path "
                    "computes no local or ionic force terms"}
        log(f"nl forces fp32 vs fp64: max {force_err['max_abs_err_Ha_per_bohr']:.3e} Ha/bohr, "
            f"rel {force_err['rel_err']:.3e}")
        apply(builtin)
        _, k32c = timed(QBH_ALL | QBH_FP32, char_steps, kernel_events=True)
        k32 = {k: {"avg_ms": ms / cnt, "launches_per_step": cnt / char_steps}
               for k, (ms, cnt) in k32c.items()}
        fp32 = {"value": nst / dt32, "ms_per_step": dt32 * 1e3, "speedup_vs_fp64": dt / dt32,
                "hpsi_rel_err_vs_fp64": rel,
                "band_energy_diff_per_atom_Ha": abs(e32 - e64) / sysm.na,
                "density_L1_vs_fp64_electrons": dl1, "density_L1_rel": dl1 / ne,
                "nelectrons": ne, "nl_force_error": force_err,
                "tuning": sched32, "kernels_sequential_builtin": k32}
    del h64

    scf = None
    if not args.no_scf:
        scf = scf_bench(args, ctx, basis, ham, sysm, species, psi)
        log("scf section done")
    batch = None
    if not args.no_batch:
        batch = batch_bench(args)
        log("batch section done")
    large = None
    if not args.no_large:
        large = large_cell_bench(args)
        log("large-cell section done")
    td = None
    if not args.no_td:
        td = td_bench(args)
        log("differentiation-form section done")

    # CPU baseline: the oracle (numpy restatement) on the host cores, bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sysm, args.cpu_local_bands, psi, ngw, nst)
        log("cpu baseline done")

    out = {
        "metric": "Hψ band-applications/s at Si512 Γ (FP64 plain form)",
        "value": value,
        "unit": "band-applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded psi, smooth random v_loc; KB projectors of the Si HSCV LDA "
                "pseudopotential)",
        "config": {"workload": f"Si{sysm.na} diamond Γ, a={10.26 * args.ncell:.2f} bohr, ecut 25 Ry, "
                               f"{ngw} PWs, {nst} bands, grid {sysm.np[0]}^3, {4 * sysm.na} KB projectors",
                   "global_batch": nst, "parallelism": "single GPU"},
        "roofline": roof,
        "fft_stages": fft,
        "cpu_baseline": cpu,
        "fp32_bulk": fp32,
        "scf": scf,
        "batching": batch,
        "large_cell": large,
        "differentiation": td,
        "tuning": sched64,
        "kernels_sequential_builtin": per_kernel,
        "band_energy_Ha": e64,
        "setup_s": t_setup,
        "libraries": libinfo,
    }
    print(json.dumps(out))
    ham.close()
    basis.close()
    ctx.close()


def scf_bench(args, ctx, basis, ham, sysm, species, psi0):
    """The precision form at its own size: the BOSampleStepper SCF loop (PSDA with the
    preconditioner cutoff --scf-ecutprec, Anderson / Kerker charge mixing, the subspace
    diagonalisation after every cycle) on Si512 with the Si HSCV LDA pseudopotential
    (v_ion, rho_ps from its local part, Ewald esr - eself added to etotal), from the
    benchmark's orthonormal start, run twice: all Hψ in FP64, and FP32-bulk Hψ with the
    last inner iteration of every cycle in FP64.  Reports etotal of both, ΔE per atom, the
    density L1 difference and the non-local force difference of the final states, the last
    cycle's energy change (convergence) and times."""
    from qball_amd import QBH_ALL, QBH_FP32, Basis, capi, ewald_esr, local_tables
    from qball_amd.devmem import DeviceArray, synchronize
    ngw, nst = sysm.ngw, sysm.nst
    vr = capi.rods_build(sysm.a, 4.0 * sysm.ecut, True)
    vb = Basis(ctx, True, vr["rod_h"], vr["rod_k"], vr["rod_lmin"], vr["rod_size"], sysm.np,
               sysm.b)
    tau = DeviceArray.from_numpy(np.ascontiguousarray(sysm.tau).ravel())
    vion = DeviceArray(16 * vb.ngw)
    rhops = DeviceArray(16 * vb.ngw)
    local_tables(vb, [species], [sysm.na], tau.ptr, sysm.omega, vion.ptr, rhops.ptr)
    esr = ewald_esr(sysm.a, [species], [sysm.na], sysm.tau)
    eself = sysm.na * species.eself
    occ = np.full(nst, 2.0)
    ecp = args.scf_ecutprec if args.scf_ecutprec > 0.0 else sysm.ecut
    start = psi0.to_numpy()
    psi = DeviceArray(psi0.nbytes)
    N = int(np.prod(sysm.np))
    rho = DeviceArray(8 * N)
    runs = {}
    for name, fi in (("fp64", QBH_ALL), ("precision", QBH_ALL | QBH_FP32)):
        psi.upload(start)
        synchronize()
        t0 = time.perf_counter()
        if args.scf_diag:
            hist = ham.scf_run_smeared(vb, psi.ptr, ngw, nst, occ, 2.0 * nst, 0.0, sysm.omega,
                                       vion.ptr, rhops.ptr, args.scf_cycles, args.scf_nite,
                                       ecp, alpha=args.scf_alpha, ndim=args.scf_ndim,
                                       rc_kerker=args.scf_kerker, flags_inner=fi,
                                       flags_last=QBH_ALL)[0]
        else:
            hist = ham.scf_run(vb, psi.ptr, ngw, nst, occ, sysm.omega, vion.ptr, rhops.ptr,
                               args.scf_cycles, args.scf_nite, ecp, alpha=args.scf_alpha,
                               ndim=args.scf_ndim, rc_kerker=args.scf_kerker, flags_inner=fi,
                               flags_last=QBH_ALL)
        synchronize()
        t1 = time.perf_counter()
        rho.zero()
        basis.density(psi.ptr, ngw, nst, occ, 1.0, sysm.omega, rho.ptr)
        ctx.sync()
        f = ham.nl_forces(psi.ptr, ngw, nst, occ)
        et = hist[:, -1, 0] + esr - eself
        runs[name] = {"etotal_Ha": float(et[-1]), "etotal_per_cycle_Ha": [float(x) for x in et],
                      "last_cycle_change_Ha": float(et[-1] - et[-2]) if len(et) > 1 else None,
                      "terms_Ha": dict(zip(("ekin", "enl", "eps", "ehart", "exc"),
                                           [float(x) for x in hist[-1, -1, 1:6]])),
                      "seconds": t1 - t0, "rho": rho.to_numpy(), "fnl": f}
        log(f"scf {name}: etotal {et[-1]:.10f} Ha in {t1 - t0:.2f} s "
            f"(last cycle change {runs[name]['last_cycle_change_Ha']:.3e})")
    a, b = runs["fp64"], runs["precision"]
    out = {"workload": f"Si{sysm.na} Γ, Si HSCV LDA (rcps {species.rcps:.2f}), {ngw} PWs, "
                       f"{nst} bands, grid {sysm.np[0]}^3, {args.scf_cycles} SCF cycles x "
                       f"{args.scf_nite} PSDA iterations, Anderson ndim {args.scf_ndim}, "
                       f"mixing {args.scf_alpha}, Kerker rc {args.scf_kerker:g}, "
                       f"preconditioner cutoff {ecp:g} Ha"
                       + (", subspace diagonalisation each cycle" if args.scf_diag else ""),
           "esr_Ha": esr, "eself_Ha": eself,
           "dE_per_atom_Ha": abs(b["etotal_Ha"] - a["etotal_Ha"]) / sysm.na,
           "density_L1_electrons": float(np.abs(b["rho"] - a["rho"]).sum() * sysm.omega / N),
           "nl_force_max_abs_diff_Ha_per_bohr": float(np.max(np.abs(b["fnl"] - a["fnl"]))),
           "nl_force_max_abs_Ha_per_bohr": float(np.max(np.abs(a["fnl"]))),
           "speedup": a["seconds"] / b["seconds"]}
    for k in ("fp64", "precision"):
        out[k] = {x: y for x, y in runs[k].items() if x not in ("rho", "fnl")}
    for d in (psi, rho, tau, vion, rhops):
        d.close()
    vb.close()
    return out


def load_traffic():
    """HBM bytes per launch of each kernel at Si512 FP64 from rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs (scaled by the counters' reading for a fill and a sum of known size),
    committed beside the rocprof summaries; {} when absent."""
    p = os.path.join(ROOT, "profiles", "traffic_si512_fp64.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        d = json.load(f)
    return {k: v["hbm_bytes_per_launch"] for k, v in d.get("kernels", {}).items()}


def batch_bench(args):
    """12 uneven cells (cubic boxes 12-28 bohr, 32-48 atoms, 64-320 bands, ecut 30 Ry, Γ),
    FP32-bulk Hψ after FP64 Gram orthonormalisation.  Each cell lives on one of
    `--batch-streams` contexts (own HIP stream), cells spread over them by estimated work;
    timed one cell after another (each synchronised) and with every cell's Hψ issued before
    one synchronisation of all streams, so small cells share the GPU; both with the built-in
    launch choices, the concurrent one also after the autotuner (the one-graph launch,
    qbh_batch_create, is switched off)."""
    from qball_amd import QBH_ALL, QBH_FP32, Basis, Context, Hamiltonian
    from qball_amd.devmem import DeviceArray, synchronize
    from qball_amd.workload import batch_workload, make_batch_cell
    flags = QBH_ALL | QBH_FP32
    nstream = max(1, args.batch_streams)
    ctxs = [Context(0, None) for _ in range(nstream)]
    load = [0.0] * nstream
    specs = [make_batch_cell(**p) for p in batch_workload()]
    cells = []
    # heaviest cells first, each onto the least loaded stream (work ~ bands * grid points)
    for s in sorted(specs, key=lambda s: -s.nst * float(np.prod(s.np))):
        k = int(np.argmin(load))
        load[k] += s.nst * float(np.prod(s.np))
        ctx = ctxs[k]
        r = s.rods
        basis = Basis(ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
        ham = Hamiltonian(ctx, basis)
        v = DeviceArray.from_numpy(s.v)
        ham.set_vloc(v.ptr)
        ham.set_projectors([s.na], [4], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(), s.omega)
        psi = DeviceArray.from_numpy(np.ascontiguousarray(s.psi.T).view(np.float64).ravel())
        basis.gram(psi.ptr, s.ngw, s.nst)
        out = DeviceArray(psi.nbytes)
        out.zero()
        cells.append(dict(s=s, ctx=ctx, basis=basis, ham=ham, v=v, psi=psi, out=out))
        del s.psi
    synchronize()
    nb_total = sum(c["s"].nst for c in cells)

    def run(c):
        c["ham"].hpsi(c["psi"].data_ptr(), c["s"].ngw, c["s"].nst, c["s"].occ,
                      c["out"].data_ptr(), c["s"].ngw, flags=flags)

    def serial():
        for c in cells:
            run(c)
            c["ctx"].sync()

    def concurrent():
        for c in cells:
            run(c)
        for ctx in ctxs:
            ctx.sync()

    def timed(fn, reps):
        fn()
        synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        synchronize()
        return (time.perf_counter() - t) / reps

    t_serial = timed(serial, args.batch_reps)
    t_conc = timed(concurrent, args.batch_reps)
    tuned = []
    for c in cells:
        c["ham"].autotune(c["psi"].data_ptr(), c["s"].ngw, c["s"].nst, c["s"].occ,
                          c["out"].data_ptr(), c["s"].ngw, flags=flags, reps=3)
        tuned.append({k: c["ham"].get_tuning(k) for k in
                      ("zy_reg", "x_reg", "x_spw", "radix", "nsplit", "overlap", "fnl_algo",
                       "enl_algo")})
    t_tuned = timed(concurrent, args.batch_reps)
    out = {"cells": len(cells), "bands_total": nb_total, "dtype": "fp32-bulk",
           "streams": nstream, "value": nb_total / t_tuned, "unit": "band-applications/s",
           "serial_default_ms": t_serial * 1e3, "concurrent_default_ms": t_conc * 1e3,
           "concurrent_tuned_ms": t_tuned * 1e3,
           "graph": "disabled (qbh_batch_create is switched off)",
           "shapes": [{"box": round(c["s"].a[0, 0], 2), "atoms": c["s"].na, "bands": c["s"].nst,
                       "pw": c["s"].ngw, "grid": c["s"].np[0],
                       "stream": ctxs.index(c["ctx"])} for c in cells],
           "tuning": tuned}
    for c in cells:
        c["ham"].close()
        c["basis"].close()
    for ctx in ctxs:
        ctx.close()
    return out


def large_cell_bench(args):
    """fcc Al500 Γ (a = 38.25 bohr, ecut 30 Ry: 77666 plane waves, 900 bands, 144^3 grid,
    500 x 9 s+p+d projectors): FP64 and FP32-bulk Hψ over all bands with the built-in
    launch choices and then with the autotuner's (unless --no-tune), then the FP64 subspace
    diagonalisation of the FP32-bulk result (the large-cell form: FP32-bulk Hψ + FP64
    subspace), then SCF cycles with Fermi smearing of 0.01 Ha for the 1500 electrons over
    the 900 bands (FP32-bulk inner Hψ, FP64 last iteration, FP64 subspace diagonalisation
    and occupations each cycle) on Gaussian ionic inputs of valence 3."""
    from qball_amd import QBH_ALL, QBH_FP32, Basis, Context, Hamiltonian, capi
    from qball_amd.devmem import DeviceArray, synchronize
    from qball_amd.workload import make_fcc_al, td_ionic_inputs
    s = make_fcc_al(seed=1)
    ctx = Context(0, None)
    r = s.rods
    basis = Basis(ctx, True, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
    ham = Hamiltonian(ctx, basis)
    v = DeviceArray.from_numpy(s.v)
    ham.set_vloc(v.ptr)
    ham.set_projectors([s.na], [len(s.lproj)], s.lproj, s.wt, s.twnl.ravel(), s.tau.ravel(),
                       s.omega)
    psi = DeviceArray.from_numpy(np.ascontiguousarray(s.psi.T).view(np.float64).ravel())
    del s.psi
    basis.gram(psi.ptr, s.ngw, s.nst)
    hpsi = DeviceArray(psi.nbytes)
    out = {"workload": f"fcc Al{s.na} Γ, a=38.25 bohr, ecut 30 Ry, {s.ngw} PWs, {s.nst} bands, "
                       f"grid {s.np[0]}^3, {s.na * len(s.lproj)} KB projectors (s+p+d)"}
    knobs = ("zy_reg", "x_reg", "x_spw", "radix", "nsplit", "overlap", "fnl_algo", "enl_algo")

    def timed():
        for _ in range(2):
            ham.hpsi(psi.ptr, s.ngw, s.nst, s.occ, hpsi.ptr, s.ngw, flags=fl)
        synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            ham.hpsi(psi.ptr, s.ngw, s.nst, s.occ, hpsi.ptr, s.ngw, flags=fl)
        synchronize()
        return (time.perf_counter() - t) / args.steps

    for name, fl in (("fp64", QBH_ALL), ("fp32_bulk", QBH_ALL | QBH_FP32)):
        for k in knobs:
            ham.set_tuning(k, -1 if k in ("overlap", "fnl_algo", "enl_algo") else
                           (1 if k in ("zy_reg", "x_reg") else 0))
        dt = dt_builtin = timed()
        choice = None
        if not args.no_tune:
            # launch choices for this shape from the library's autotuner (untimed)
            ham.autotune(psi.ptr, s.ngw, s.nst, s.occ, hpsi.ptr, s.ngw, flags=fl, reps=2,
                         force=True)
            choice = {k: ham.get_tuning(k) for k in knobs}
            dt = timed()
        out[name] = {"ms_per_hpsi": dt * 1e3, "band_applications_per_s": s.nst / dt,
                     "builtin_ms_per_hpsi": dt_builtin * 1e3, "tuned_choice": choice}
    # FP64 subspace diagonalisation of the FP32-bulk Hψ (eigenvalues only; psi unchanged)
    synchronize()
    t = time.perf_counter()
    eig = basis.subspace_diag(psi.ptr, s.ngw, hpsi.ptr, s.ngw, s.nst, eigvec=False)
    synchronize()
    out["subspace_diag_fp64_ms"] = (time.perf_counter() - t) * 1e3
    out["eig_min_max_Ha"] = [float(eig.min()), float(eig.max())]
    # Fermi-smeared SCF (BOSampleStepper with fractional occupations)
    vr = capi.rods_build(s.a, 4.0 * s.ecut, True)
    vb = Basis(ctx, True, vr["rod_h"], vr["rod_k"], vr["rod_lmin"], vr["rod_size"], s.np, s.b)
    vion, rhops = td_ionic_inputs(s, vr, valence=3.0)
    dvion = DeviceArray.from_numpy(np.ascontiguousarray(vion).view(np.float64))
    drhops = DeviceArray.from_numpy(np.ascontiguousarray(rhops).view(np.float64))
    nel, width, ncyc, nite = 3.0 * s.na, 0.02, args.smeared_cycles, 2
    occ0 = np.full(s.nst, nel / s.nst)
    synchronize()
    t = time.perf_counter()
    hist, occ, eigs, mu = ham.scf_run_smeared(vb, psi.ptr, s.ngw, s.nst, occ0, nel, width,
                                              s.omega, dvion.ptr, drhops.ptr, ncyc, nite,
                                              s.ecut, flags_inner=QBH_ALL | QBH_FP32,
                                              flags_last=QBH_ALL)
    synchronize()
    ts = time.perf_counter() - t
    frac = int(np.sum((occ > 1e-6) & (occ < 2.0 - 1e-6)))
    out["smeared_scf"] = {
        "cycles": ncyc, "psda_iterations_per_cycle": nite, "smearing_Ry": width,
        "nelectrons": nel, "seconds": ts, "ms_per_cycle": ts / ncyc * 1e3,
        "etotal_per_cycle_Ha": [float(x) for x in hist[:, -1, 0]],
        "mu_per_cycle_Ha": [float(x) for x in mu],
        "fractionally_occupied_bands": frac, "occupation_sum": float(occ.sum()),
        "eig_lowest_highest_Ha": [float(eigs[-1].min()), float(eigs[-1].max())],
        "note": "electronic terms only (the Gaussian ionic inputs carry no Ewald energy)"}
    log(f"al500 smeared scf: {ncyc} cycles in {ts:.2f} s, mu {mu[-1]:.5f} Ha, "
        f"{frac} fractional bands")
    for d in (dvion, drhops):
        d.close()
    vb.close()
    ham.close()
    basis.close()
    ctx.close()
    return out


def td_bench(args):
    """The differentiation form: real-time TDDFT on Si64 in a complex basis (18277 plane
    waves, 128 bands, 72^3 grid), 200 ETRS steps of dt = 0.05 under A(t) = A0 sin(0.3 t)
    exp(-(t - t0)^2 / 2 tau^2) along x (A0 = 0.01), objective the time-integrated dipole
    along x, psi stored every 20 steps: forward and adjoint (backward, with the segment
    recomputation) times in FP64 and FP32-bulk, and the FP32-bulk J and dJ/dA0 against FP64.
    The Si HSCV LDA pseudopotential: projectors re-tabled at |k+G-A| at every step, v_ion and
    rho_ps from its local part."""
    from qball_amd import (QBH_FP32, Basis, Context, Hamiltonian, Species, TDPropagator,
                           auto_rcps, capi, local_tables)
    from qball_amd.devmem import DeviceArray, synchronize
    from qball_amd.workload import dipole_weights, field_shape, make_system
    nsteps, dt, ckpt, A0 = 200, 0.05, 20, 0.01
    s = make_system(ncell=2, ecut_ry=25.0, nbands=128, seed=41, complex_basis=True)
    ctx = Context(0, None)
    r = s.rods
    basis = Basis(ctx, False, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], s.np, s.b)
    ham = Hamiltonian(ctx, basis)
    v = DeviceArray.from_numpy(s.v)
    ham.set_vloc(v.ptr)
    species = Species(SI_XML, rcps=auto_rcps(4.0, s.a, s.ecut))
    ham.set_species([species], [s.na], s.tau, s.omega)
    vr = capi.rods_build(s.a, 4.0 * s.ecut, False)
    vb = Basis(ctx, False, vr["rod_h"], vr["rod_k"], vr["rod_lmin"], vr["rod_size"], s.np, s.b)
    tau = DeviceArray.from_numpy(np.ascontiguousarray(s.tau).ravel())
    dvion = DeviceArray(16 * vb.ngw)
    drhops = DeviceArray(16 * vb.ngw)
    local_tables(vb, [species], [s.na], tau.ptr, s.omega, dvion.ptr, drhops.ptr)
    psi = DeviceArray.from_numpy(np.ascontiguousarray(s.psi.T).view(np.float64).ravel())
    basis.gram(psi.ptr, s.ngw, s.nst)
    n0, n1, n2 = s.np
    z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    vext = DeviceArray.from_numpy((0.02 * np.cos(2 * np.pi * x / n0) * np.cos(2 * np.pi * y / n1)
                                   + 0.01 * np.sin(2 * np.pi * z / n2)).ravel())
    gv = DeviceArray(8 * n0 * n1 * n2)
    occ = np.full(s.nst, 2.0)
    td = TDPropagator(ham, vb, s.nst, occ, 1.0, s.omega, dvion.ptr, drhops.ptr)
    shape = field_shape(nsteps, dt, omega=0.3)
    w = dipole_weights(nsteps, dt)
    out = {"workload": f"Si{s.na} complex basis, Si HSCV LDA (projectors re-tabled at "
                       f"|k+G-A| every step), {s.ngw} PWs, {s.nst} bands, grid {n0}^3, "
                       f"{nsteps} ETRS steps of dt={dt}, checkpoint every {ckpt} steps"}
    for name, fl in (("fp64", 0), ("fp32_bulk", QBH_FP32)):
        synchronize()
        t0 = time.perf_counter()
        J = td.forward(psi.ptr, s.ngw, nsteps, dt, A0 * shape, w, vext_ptr=vext.ptr,
                       checkpoint=ckpt, flags=fl)
        synchronize()
        t1 = time.perf_counter()
        gA = td.backward(gv.ptr)
        synchronize()
        t2 = time.perf_counter()
        out[name] = {"J": J, "dJ_dA0": float(gA[:, 0] @ shape[:, 0]), "forward_s": t1 - t0,
                     "backward_s": t2 - t1, "forward_steps_per_s": nsteps / (t1 - t0)}
    out["fp32_bulk"]["J_rel_vs_fp64"] = abs(out["fp32_bulk"]["J"] - out["fp64"]["J"]) / abs(out["fp64"]["J"])
    out["fp32_bulk"]["dJ_dA0_rel_vs_fp64"] = (abs(out["fp32_bulk"]["dJ_dA0"] - out["fp64"]["dJ_dA0"])
                                              / abs(out["fp64"]["dJ_dA0"]))
    td.close()
    ham.close()
    species.close()
    vb.close()
    basis.close()
    ctx.close()
    return out


def cpu_baseline(sysm, local_bands, psi_dev, ngw, nst):
    """The oracle's Hψ on the host cores (at most 16, the GPU's share of the box): the
    pruned pencil transforms of oracle.fft.PrunedPlan on scipy's pocketfft with one worker
    per core, and the projector products as numpy BLAS products with as many threads.  The
    non-local term runs over all nst bands (projector table built per 128-atom block); the
    local + kinetic term over `local_bands` bands (all of them when local_bands >= nst),
    scaled to nst.  One small untimed pass first starts the thread pools and the heap.
    Returns band-applications/s of the full Hψ so composed."""
    import ctypes
    from threadpoolctl import threadpool_limits
    from oracle import fft as offt
    from oracle import hpsi as oh
    from oracle import species as osp
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    # keep large blocks on the heap and freed memory mapped, so the pass-to-pass buffers
    # of the transforms are reused instead of faulted in afresh (glibc M_MMAP_MAX = 0,
    # M_TRIM_THRESHOLD = 2 GiB)
    libc = ctypes.CDLL("libc.so.6")
    libc.mallopt(-4, 0)
    libc.mallopt(-1, 2 ** 31 - 1)
    ob = OB(UnitCell(sysm.a), sysm.ecut)
    maps = FFTMaps(ob, *sysm.np)
    plan = offt.PrunedPlan(maps)
    ora = osp.Species(osp.read_species_xml(SI_XML), osp.auto_rcps(4.0, sysm.a, sysm.ecut))
    rows, ls, wt = osp.twnl_table(ora, ob.kpgx)
    sp = oh.Projectors(ls, wt, rows, sysm.tau)
    c = psi_dev.to_numpy().view(np.complex128).reshape(nst, ngw).T.copy()
    nb = min(nst, max(2, 2 * (local_bands // 2)))
    workers = offt.WORKERS
    offt.WORKERS = cores
    try:
        with threadpool_limits(limits=cores):
            warm = np.zeros((ngw, 4), dtype=np.complex128)
            oh.rs_mul_add_batched(maps, c[:, :4], sysm.v, warm, batch=1, plan=plan)
            oh.nonlocal_energy(ob, [oh.Projectors(ls, wt, rows, sysm.tau[:min(128, sysm.na)])],
                               c[:, :64], sysm.occ[:64], sysm.omega, cp=np.zeros_like(c[:, :64]))
            del warm
            cp = np.zeros((ngw, nb), dtype=np.complex128)
            t0 = time.perf_counter()
            oh.kinetic_add(ob.kpg2, c[:, :nb], cp)
            oh.rs_mul_add_batched(maps, c[:, :nb], sysm.v, cp, batch=1, plan=plan)
            t_loc = time.perf_counter() - t0
            del cp
            hp = np.zeros_like(c)
            t0 = time.perf_counter()
            oh.nonlocal_energy(ob, [sp], c, sysm.occ, sysm.omega, cp=hp)
            t_nl = time.perf_counter() - t0
            del hp
    finally:
        offt.WORKERS = workers
    t_full = t_nl + t_loc * (nst / nb)
    scaled = "" if nb == nst else f" scaled to {nst}"
    return {"value": nst / t_full, "unit": "band-applications/s", "cores": cores, "kind": "port",
            "sample": f"oracle Hψ on {cores} host cores (pruned pencil transforms on scipy "
                      f"pocketfft with {cores} workers, numpy BLAS with {cores} threads): the "
                      f"non-local term over all {nst} bands ({t_nl:.1f} s) plus the local + "
                      f"kinetic term over {nb} bands ({t_loc:.1f} s){scaled}",
            "t_nonlocal_s": t_nl, "t_local_s": t_loc, "local_bands": nb}


if __name__ == "__main__":
    main()
