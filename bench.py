"""Benchmark of the TRANSPORTSE SOR hot path on one MI355X.

Workload: the 512x512x256 pressure problem -- 5 m cells, lognormal
permeability (sigma_lnK 2.5, exponential correlation length 16 cells,
seed 3), porosity 0.2, linear EOS rho = 1000 + 700 C with a salt patch on
top, hydrostatic Dirichlet x faces, dt = 1 day.  One step = one pressure
solve (compute_P's system) from the hydrostatic initial state to relative
residual ||b - A p|| / ||b|| <= 1e-10 in the mixed form: FP32 red/black
sweeps on the correction inside FP64 defect correction (operator assembly,
sweeps, FP64 residual passes and write-back included; inputs resident in HBM).

Reported (one JSON line on stdout):
  value        SOR cell-updates/s over the whole timed steps
  roofline     dominant kernel (FP32 fused red/black sweep, NS sweeps per
               launch), timed with HIP events on the library's stream:
               achieved = algorithmic bytes per launch (one pass of the
               operands: 18 B per cell for the tuned plan's sweep with the
               bfloat16 operator and FP16 rhs, 20 B with the FP32 rhs, 28 B
               with the FP32 operator) / mean launch time, frac = achieved /
               8 TB/s; traffic = HBM bytes per launch measured in this run by
               two rocprofv3 --pmc child runs (FETCH_SIZE, WRITE_SIZE, each
               scaled by its reading for a 1 GiB copy); the per-sweep figure
               (every fused sweep at those bytes per cell) is
               effective_per_sweep_*
  cpu_baseline the oracle's OpenMP red/black sweeps on the host cores, and
               the time to solution of one compute_P call (oracle, faithful
               and mixed GPU forms) on the 64x32x32 problem
  extra keys   time to solution, solves/s, sweeps, the plain FP64 form's time
               to solution on the same problem and the speedup over it, and
               the mixed solution's deviation from the FP64 solution.

Further sections of the same line (each bounded; --no-configs / --no-sia /
--no-ensemble / --no-serve skip them):
  configs_faithful64  the salt 64x32x32 config (100 steps) and the salt +
               heat SIA 256x256x128 config (20 steps) through the SIA driver
               with the faithful FP64 solvers; the 64x32x32 one also with the
               oracle on the host cores (100 steps)
  sia          device-resident SIA steps (heat + salt + tracer) on the
               512x512x256 problem in the mixed form and in plain FP64 solved
               to 1e-13, and the relative L2 / max error of P, T and each C
               of the mixed run against the FP64 run
  ensemble     64 realizations of uneven grids (nx, ny in 32..160, nz in
               32..96, seed 4), heat + salt, 50 SIA steps each, in lockstep
               batched solves with hipGraph-replayed persistent tile-queue
               cycles and autotuned plans: cell-updates/s, s per step
  serving      Poisson request arrivals (grids 64^3 / 128^3 / 256x256x128,
               heat + 3 species, 10 SIA steps each, mixed to 1e-10) over a
               3 s window per rate (8..24 requests) under a 200 GB HBM
               admission budget, plus the top rate under a 12 GB budget that
               makes admission bind: p50 / p99 time to solution split into
               queueing and service, solves/s, concurrency

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
       python bench.py --profile-kernel N   (N launches of the sweep kernel
                                              and a 1 GiB copy; for rocprof)
       python bench.py --profile-faithful N (N red/black iterations of the
                                              drop-in compute_P sweep kernel on
                                              the same inputs; for rocprof)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FP32_SWEEP = 28   # 3 faces + 1/diag + rhs + unknown read + write, FP32
# variant 5 stores the three faces and the storage term as bfloat16 (the
# diagonal is summed in the kernel): 8 + 4 + 4 + 4 bytes; variant 7 the rhs
# as well: 8 + 2 + 4 + 4 bytes
BYTES_FP32_SWEEP_BY_VARIANT = {5: 20, 7: 18}
BYTES_DEFINITION = {
    5: "20 B per cell (16-byte pair record of the 3 face transmissibilities and "
       "the storage term as bfloat16, rhs, unknown in and out FP32), once per "
       "launch",
    7: "18 B per cell (16-byte pair record of the 3 face transmissibilities and "
       "the storage term as bfloat16, rhs FP16, unknown in and out FP32), once "
       "per launch"}
BYTES_DEFINITION_FP32 = ("28 B per cell (3 face transmissibilities, 1/diag, rhs, "
                         "unknown in and out, FP32), once per launch")
DTYPE_BY_VARIANT = {
    5: "f32 sweeps (inner operator stored as bf16) in f64 defect correction",
    7: "f32 sweeps (inner operator stored as bf16, rhs as fp16) in f64 defect "
       "correction"}
BYTES_FP64_SWEEP = 56
# sweep kernel (name as rocprofv3 reports it, without template arguments)
# per plan variant
VARIANT_KERNEL = {0: "rb_zmarch_kernel", 1: "rb_zmarch_xl_kernel",
                  2: "rb_zmarch_xr_kernel", 3: "rb_zmarch_x2_kernel",
                  5: "rb_zmarch_x5_kernel", 7: "rb_zmarch_x5_kernel"}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--grid", type=int, nargs=3, default=[512, 512, 256])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--omega", type=float, default=None,
                    help="SOR factor (default: tuned online over the warm-up solves)")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sweeps", type=int, default=8)
    ap.add_argument("--profile-kernel", type=int, default=0)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc runs that measure the sweep "
                         "kernel's HBM traffic")
    ap.add_argument("--profile-faithful", type=int, default=0)
    ap.add_argument("--plan", type=int, nargs="+", default=None,
                    metavar="N", help="FP32 sweep plan NS BY RY ZCHUNK [VARIANT]"
                    " instead of autotuning")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--a-steps", type=int, default=100)
    ap.add_argument("--a-cpu-steps", type=int, default=100)
    ap.add_argument("--b-steps", type=int, default=20)
    ap.add_argument("--no-sia", action="store_true")
    ap.add_argument("--sia-steps", type=int, default=1)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-members", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=50)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rates", type=float, nargs="+",
                    default=[5.0, 10.0, 20.0, 50.0, 100.0, 200.0])
    ap.add_argument("--serve-window", type=float, default=3.0,
                    help="seconds over which a rate's requests arrive")
    ap.add_argument("--serve-min-requests", type=int, default=8)
    ap.add_argument("--serve-max-requests", type=int, default=24)
    ap.add_argument("--serve-tight-gb", type=float, default=12.0,
                    help="admission budget (GB) of the extra run at the top rate")
    ap.add_argument("--serve-steps", type=int, default=10)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not "
              "supported (the solve shards by z slabs with halo exchange, but "
              "no multi-GPU run is part of this build)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from transportse_amd.pressure import FastPressure
    from transportse_amd.workloads import PressureProblem

    nx, ny, nz = args.grid
    N = nx * ny * nz
    t0 = time.time()
    prob = PressureProblem(nx, ny, nz, spacing=5.0, sigma=2.5, corr_len=16,
                           seed=3)
    log(f"inputs generated in {time.time() - t0:.1f}s")
    dt = 86400.0
    if args.profile_faithful:
        return profile_faithful(args, prob, dt)
    fp = FastPressure(*prob.grid_args())
    fp.load(prob.P, dt, **prob.p_inputs())
    torch.cuda.synchronize()
    log("uploaded and assembled")

    if args.plan:
        fp.plans["mixed"] = tuple(args.plan)
    t_tune = time.perf_counter()
    plan32 = fp.plan("mixed")
    tune_s = time.perf_counter() - t_tune
    log(f"FP32 plan (ns, by, ry, zchunk) = {plan32}"
        + ("" if args.plan else " (autotuned)"))

    if args.profile_kernel:
        fp.solve("mixed", max_sweeps=0, plan=plan32)
        n = args.profile_kernel * plan32[0]
        fp.sweeps(n, plan32, sor=args.omega or 1.996, fp32=True)
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)      # calibration: 1 GiB read + 1 GiB write
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernel_launches": args.profile_kernel,
                          "plan": plan32, "copy_bytes": 2 * src.numel() * 8}))
        return 0

    from transportse_amd.omega import OmegaTuner
    tuner = OmegaTuner(max(nx, ny, nz))
    explicit_omega = args.omega is not None

    def step(form, omega):
        # with the factor from the tuner (no --omega), each solve may also
        # move it inside the solve (adaptive SOR, Young's estimate)
        fp.assemble()
        st = fp.solve(form, tol=args.tol, sor=omega, adaptive=True,
                      adapt_sor=not explicit_omega)
        fp.result_device()  # solution back into a split-layout device field
        return st

    cold = None
    for w in range(args.warmup):
        om = args.omega if args.omega else tuner.omega
        st = step("mixed", om)
        if w == 0:   # the first solve: the tuner's starting factor, untuned
            cold = {"time_to_solution_s": st.seconds, "sweeps": st.sweeps,
                    "omega": om, "omega_final": st.sor, "converged": st.converged,
                    "plan_autotune_s": tune_s if not args.plan else 0.0}
        if not args.omega:
            tuner.report(st.sweeps, st.contraction, st.sor)
        log(f"warmup solve (omega {om:.5f} -> {st.sor:.5f}): {st.sweeps} sweeps, rel "
            f"{st.rel_residual:.2e}, {st.seconds:.2f}s, contraction per sweep "
            f"{st.contraction:.6f}")
    if not args.omega:      # the best factor the warm-up solves found
        tuner.u = tuner.best_u if tuner.best_u is not None else tuner.u
    omega = args.omega if args.omega else tuner.omega
    args.omega = omega
    log(f"SOR factor {omega:.5f}")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step("mixed", omega))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solve_sweeps = sum(s.sweeps for s in stats)
    conv = all(s.converged for s in stats)
    value = float(N) * solve_sweeps / elapsed
    tts = elapsed / args.steps
    log(f"timed {args.steps} mixed solves: {tts:.3f}s each, "
        f"{solve_sweeps // args.steps} sweeps, converged={conv}, "
        f"{value:.3e} cell-updates/s")
    P_mixed = fp.result()

    # dominant kernel: the FP32 fused sweep, HIP events on the library stream
    # (the context issues on torch's current stream, where the events go)
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ns = plan32[0]
    launches = 200
    fp.sweeps(50 * ns, plan32, sor=args.omega, fp32=True)  # steady clocks
    ev0.record(stream)
    fp.sweeps(launches * ns, plan32, sor=args.omega, fp32=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    launch_s = ev0.elapsed_time(ev1) * 1e-3 / launches
    # algorithmic bytes of one launch: every operand once (the plan's bytes
    # per cell); the ns sweeps a launch fuses share that one pass
    variant = plan32[4] if len(plan32) > 4 else 0
    cell_bytes = BYTES_FP32_SWEEP_BY_VARIANT.get(variant, BYTES_FP32_SWEEP)
    bytes_launch = cell_bytes * N
    achieved = bytes_launch / launch_s / 1e9
    per_sweep = cell_bytes * N * ns / launch_s / 1e9
    log(f"sweep kernel: {launch_s * 1e6:.1f} us/launch ({ns} sweeps), "
        f"{achieved:.0f} GB/s one-pass ({achieved / HBM_PEAK_GBS:.3f} of peak), "
        f"{per_sweep:.0f} GB/s counting every sweep at {cell_bytes} B/cell")

    plain = None
    if not args.no_compare:
        st64 = step("plain64", omega)
        P_plain = fp.result()
        change = float(np.max(np.abs(P_plain - prob.P)))
        dev = float(np.max(np.abs(P_mixed - P_plain)))
        plan64 = fp.plan("plain64")
        plain = {"time_to_solution_s": st64.seconds, "sweeps": st64.sweeps,
                 "converged": st64.converged, "rel_residual": st64.rel_residual,
                 "plan": list(plan64),
                 "speedup_mixed_vs_plain64": st64.seconds / tts,
                 "max_abs_dev_mixed_vs_plain64_Pa": dev,
                 "dev_rel_to_change": dev / change if change else None,
                 "dev_rel_to_P": dev / float(np.max(np.abs(P_plain)))}
        log(f"plain64 solve: {st64.seconds:.3f}s, {st64.sweeps} sweeps; speedup "
            f"{st64.seconds / tts:.2f}x; max |P_mixed - P_plain| = {dev:.3e} Pa")

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.build()
        threads = len(os.sched_getaffinity(0))
        orc.set_threads(threads)
        o = orc.Oracle(*prob.grid_args(), frozen_source=prob)
        P = prob.P.copy()
        c0 = time.perf_counter()
        for _ in range(args.cpu_sweeps):
            for colour in (0, 1):
                o.p_half_sweep(P, prob.P, prob.c_f, prob.nue, prob.rho_f,
                               prob.phi, prob.Kx, prob.Ky, prob.Kz, prob.BP,
                               prob.WP, dt, args.omega, colour)
        cpu_s = time.perf_counter() - c0
        cpu = {"value": N * args.cpu_sweeps / cpu_s, "unit": "cell-updates/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_sweeps} red+black SOR sweeps of the full "
                         f"{nx}x{ny}x{nz} grid by the oracle (C restatement of "
                         f"compute_P's sweep, OpenMP {threads} threads, FP64), "
                         f"{cpu_s:.1f}s"}
        log(f"cpu baseline {cpu['value']:.3e} cell-updates/s ({threads} threads)")
        cpu["one_thread"] = cpu_sweep_rate(orc, prob, dt, args.omega, 1, 1)
        cpu["other_grids"] = {}
        for g, sweeps in (((64, 32, 32), 40), ((256, 256, 128), 6)):
            small = PressureProblem(*g, spacing=10.0, sigma=1.0, corr_len=5, seed=1)
            cpu["other_grids"]["x".join(map(str, g))] = cpu_sweep_rate(
                orc, small, dt, 1.8, threads, sweeps)
        orc.set_threads(threads)
        log(f"cpu 1 thread {cpu['one_thread']['value']:.3e}; other grids "
            + ", ".join(f"{k} {v['value']:.3e}" for k, v in cpu["other_grids"].items()))
        cpu["compute_P_time_to_solution"] = compute_p_tts(orc, threads)
        orc.set_threads(threads)

    del fp
    torch.cuda.empty_cache()
    ceiling = stream_ceiling()
    log(f"streaming ceiling (c = a + b, FP32, 3 x 1 GiB): {ceiling:.0f} GB/s")
    traffic = None if args.no_traffic else measure_traffic(args, plan32)
    cfgs = None if args.no_configs else configs_section(args)
    sia = None if args.no_sia else sia_section(args, nx, ny, nz)
    ens = None if args.no_ensemble else ensemble_section(args)
    serving = None if args.no_serve else serve_section(args)

    variant = plan32[4] if len(plan32) > 4 else 0
    tbytes = traffic.get("bytes_per_launch") if traffic else None
    srv0 = serving["runs"][0] if serving and serving.get("runs") else None
    line = {
        "metric": "SOR cell-updates/s & HBM roofline %; solves/s + p99 at rel. "
                  "residual 1e-10",
        "value_is": "SOR cell-updates/s of the 512x512x256 mixed solve",
        "hbm_roofline_pct": 100.0 * achieved / HBM_PEAK_GBS,
        "hbm_roofline_pct_definition": "one pass of the sweep operands "
                                       f"({cell_bytes} B x cells) per launch / launch time "
                                       "/ 8 TB/s",
        "serving_solves_per_s": srv0["solves_per_s"] if srv0 else None,
        "serving_p99_s": srv0["p99_s"] if srv0 else None,
        "serving_rate_per_s": srv0["rate_per_s"] if srv0 else None,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tts * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_BY_VARIANT.get(variant, "f32 sweeps in f64 defect correction"),
        "data": "synthetic",
        "config": {"workload": f"pressure solve {nx}x{ny}x{nz}, 5 m cells, "
                               "sigma_lnK 2.5, corr 16 cells, mixed precision "
                               f"to rel. residual {args.tol:g}",
                   "grid": [nx, ny, nz], "form": "mixed", "omega": omega,
                   "omega_source": "fixed" if explicit_omega else
                                   "tuned online over the warm-up solves",
                   "plan_ns_by_ry_zchunk_variant": list(plan32)},
        "time_to_solution_s": tts,
        # the same solve before any tuning on this problem: the first warm-up
        # solve at the tuner's starting SOR factor (2 - 2/max(nx, ny, nz))
        "cold_start": cold,
        "solves_per_s": 1.0 / tts,
        "sweeps_per_solve": solve_sweeps // args.steps,
        "rel_residual": stats[-1].rel_residual,
        "converged": conv,
        "plain64": plain,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tbytes,
                     "kernel": VARIANT_KERNEL[variant],
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "algorithmic_bytes_definition": BYTES_DEFINITION.get(
                         variant, BYTES_DEFINITION_FP32),
                     "sweeps_per_launch": ns, "launch_us": launch_s * 1e6,
                     # every fused sweep counted at cell_bytes per cell: the rate a
                     # one-sweep-per-launch kernel would need for the same
                     # sweep throughput (temporal blocking puts it above peak)
                     "effective_per_sweep_gbs": per_sweep,
                     "effective_per_sweep_frac": per_sweep / HBM_PEAK_GBS,
                     "traffic_over_algorithmic": (tbytes / bytes_launch
                                                  if tbytes else None),
                     "hbm_measured_gbs": tbytes / launch_s / 1e9 if tbytes else None,
                     "hbm_measured_frac": (tbytes / launch_s / 1e9 / HBM_PEAK_GBS
                                           if tbytes else None),
                     "traffic_measurement": traffic,
                     # what a plain streaming kernel reaches on this device:
                     # the practical ceiling beside the nominal peak
                     "stream_ceiling_gbs": ceiling,
                     "frac_of_stream_ceiling": achieved / ceiling},
        "cpu_baseline": cpu,
        "configs_faithful64": cfgs,
        "sia": sia,
        "ensemble": ens,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


def stream_ceiling(reps=10):
    """HBM rate of a plain streaming kernel on this GPU: c = a + b over
    three 1 GiB FP32 arrays (two reads, one write per element), timed with
    HIP events: the practical bandwidth ceiling beside the nominal 8 TB/s."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    torch.add(a, b, out=c)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.add(a, b, out=c)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b, c
    torch.cuda.empty_cache()
    return 3 * 4 * n / sec / 1e9


def _pmc_per_dispatch(outdir, counter):
    """{kernel name: [value per dispatch]} from a rocprofv3 --pmc CSV run."""
    import csv
    import glob
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"),
                      recursive=True)
    if not files:
        raise RuntimeError(f"no counter_collection.csv under {outdir}")
    per = {}
    for fn in files:
        with open(fn, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                key = (row.get("Dispatch_Id"), row.get("Kernel_Name", ""))
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    out = {}
    for (_, name), v in per.items():
        out.setdefault(name, []).append(v)
    return out


def measure_traffic(args, plan, launches=6):
    """HBM bytes per launch of the FP32 sweep kernel with `plan` on this
    problem, measured in this run: two rocprofv3 --pmc child runs of
    ``bench.py --profile-kernel`` (FETCH_SIZE, then WRITE_SIZE, in KiB per
    dispatch), each counter scaled by what it reads for the 1 GiB device
    copy the same child makes (2^30 bytes read and 2^30 written).  Returns
    the figures, or {"error": ...} (also printed) if a run fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    name = VARIANT_KERNEL[plan[4] if len(plan) > 4 else 0]
    child = [sys.executable, os.path.join(ROOT, "bench.py"), "--profile-kernel",
             str(launches), "--grid", *map(str, args.grid), "--plan", *map(str, plan)]
    res = {"kernel": name, "launches": launches, "plan": list(plan),
           "method": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over "
                     + " ".join(["bench.py"] + child[2:])
                     + "; each scaled by the same counter on a 1 GiB device copy"}
    env = dict(os.environ, TMPDIR="/tmp")
    t0 = time.time()
    try:
        for counter, key in (("FETCH_SIZE", "read"), ("WRITE_SIZE", "write")):
            d = tempfile.mkdtemp(prefix="tse_pmc_", dir="/tmp")
            cmd = ["timeout", "-s", "KILL", "300", prof, "--pmc", counter,
                   "--output-format", "csv", "-d", d, "-o", "p", "--", *child]
            p = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True,
                               text=True)
            try:
                vals = _pmc_per_dispatch(d, counter)
            except RuntimeError:
                raise RuntimeError(f"{counter} run rc {p.returncode}: "
                                   f"{p.stderr[-800:]}")
            finally:
                shutil.rmtree(d, ignore_errors=True)
            sweep = [v for n, v in vals.items() if name + "<" in n or n.endswith(name)]
            copy = [v for n, v in vals.items() if "copy" in n.lower()]
            if not sweep or not copy:
                raise RuntimeError(f"{counter}: kernels seen {sorted(vals)[:6]}")
            k_sweep = [x for v in sweep for x in v]
            k_copy = max(x for v in copy for x in v)
            factor = float(2 ** 30) / (k_copy * 1024.0)
            mean = sum(k_sweep) / len(k_sweep)
            res[key] = {"counter": counter, "kib_per_dispatch": mean,
                        "dispatches": len(k_sweep), "copy_kib": k_copy,
                        "copy_factor": factor, "bytes": mean * 1024.0 * factor}
        res["bytes_per_launch"] = res["read"]["bytes"] + res["write"]["bytes"]
        log(f"traffic of {name}: {res['bytes_per_launch'] / 1e9:.3f} GB per launch "
            f"(read factor {res['read']['copy_factor']:.2f}, write factor "
            f"{res['write']['copy_factor']:.2f}) in {time.time() - t0:.0f}s")
    except Exception as e:   # reported in the JSON line, never silently
        res["error"] = str(e)
        res["bytes_per_launch"] = None
        log(f"TRAFFIC MEASUREMENT FAILED: {e}")
    return res


def cpu_sweep_rate(orc, prob, dt, omega, threads, sweeps):
    """Red+black SOR sweeps of the oracle (reference arithmetic, FP64) on a
    pressure problem: cell-updates/s with `threads` OpenMP threads."""
    orc.set_threads(threads)
    o = orc.Oracle(*prob.grid_args(), frozen_source=prob)
    P = prob.P.copy()
    c0 = time.perf_counter()
    for _ in range(sweeps):
        for colour in (0, 1):
            o.p_half_sweep(P, prob.P, prob.c_f, prob.nue, prob.rho_f, prob.phi,
                           prob.Kx, prob.Ky, prob.Kz, prob.BP, prob.WP, dt, omega,
                           colour)
    sec = time.perf_counter() - c0
    n = prob.nx * prob.ny * prob.nz
    return {"value": n * sweeps / sec, "unit": "cell-updates/s", "cores": threads,
            "sweeps": sweeps, "seconds": sec}


def compute_p_tts(orc, threads):
    """Time to solution of one compute_P call (tse.py:647-742, the drop-in
    signature) on the 64x32x32 salt problem from its hydrostatic state (dt
    1 day, SOR 1.8, inc_it 1.25, P_conv 1e-8): the oracle on the host cores,
    the faithful GPU form (bit-identical, same iterations) and the mixed GPU
    form (P_conv read as relative residual); host arrays in and out, so the
    GPU times include the transfers."""
    import numpy as np
    import torch

    from transportse_amd.solvers import HipSolvers
    from transportse_amd.workloads import PressureProblem
    prob = PressureProblem(64, 32, 32, spacing=10.0, sigma=1.0, corr_len=5, seed=1)
    args_ = lambda: (prob.c_f, prob.nue, prob.rho_f, prob.phi, prob.Kx, prob.Ky,
                     prob.Kz, prob.qx, prob.qy, prob.qz, 86400.0, prob.BP, prob.WP)
    kw = dict(SOR=1.8, inc_it=1.25, P_conv=1e-8, max_it=20000)
    out = {"grid": [64, 32, 32], "P_conv": 1e-8, "SOR": 1.8}
    orc.set_threads(threads)
    o = orc.Oracle(*prob.grid_args(), frozen_source=prob)
    P = prob.P.copy()
    t0 = time.perf_counter()
    r = o.compute_P(P, *args_(), **kw)
    out["oracle"] = {"seconds": time.perf_counter() - t0, "iterations": r[5],
                     "cores": threads}
    P_ref = P
    for form in ("faithful64", "mixed"):
        h = HipSolvers(*prob.grid_args(), frozen_source=prob, verbose=False,
                       p_form=form)
        for rep in range(2):          # the second call is timed
            P = prob.P.copy()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = h.compute_P(P, *args_(), **kw)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
        out[form] = {"seconds": sec, "iterations": r[5], "converged": r[4],
                     "max_abs_dev_from_oracle_Pa": float(np.max(np.abs(P - P_ref)))}
        del h
    log("compute_P time to solution 64x32x32: oracle "
        f"{out['oracle']['seconds']:.3f}s ({out['oracle']['iterations']} it), GPU "
        f"faithful {out['faithful64']['seconds']:.3f}s, mixed {out['mixed']['seconds']:.3f}s")
    return out


def profile_faithful(args, prob, dt):
    """Fixed-dt red/black iterations of the drop-in compute_P half-sweep
    kernel (tse_p_sweeps: assemble once, then sweeps without control
    launches), timed with HIP events on the context's stream.  Algorithmic
    bytes per half-sweep: 129 per slot of the colour (13 record doubles, the
    old and new value, the other colour's value and the mask byte)."""
    import ctypes

    import numpy as np
    import torch

    from transportse_amd import _lib
    from transportse_amd.device import HipContext
    hx = HipContext(*prob.grid_args())
    d = {n: hx.upload(np.ascontiguousarray(a)) for n, a in prob.p_inputs().items()}
    d["P"] = hx.upload(prob.P)
    d["Pn"] = hx.upload(prob.P)
    pf = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
    _lib.check(hx.lib.tse_p_sweeps(hx.ctx, ctypes.byref(pf), dt, 1.8, 1, 1),
               "tse_p_sweeps")
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    _lib.check(hx.lib.tse_p_sweeps(hx.ctx, ctypes.byref(pf), dt, 1.8,
                                   args.profile_faithful, 0), "tse_p_sweeps")
    e1.record(stream)
    torch.cuda.synchronize()
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)      # calibration: 1 GiB read + 1 GiB write
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / (2 * args.profile_faithful)
    slots = hx.F // 2
    gbs = 129.0 * slots / (us * 1e-6) / 1e9
    print(json.dumps({"profile_faithful_iterations": args.profile_faithful,
                      "grid": list(prob.grid_args()[:3]),
                      "us_per_half_sweep": us, "algorithmic_GBps": gbs,
                      "roofline_frac": gbs / 8000.0,
                      "copy_bytes": 2 * src.numel() * 8}))
    hx.close()
    return 0


def configs_section(args):
    """The salt 64x32x32 and salt + heat SIA 256x256x128 configs through the
    SIA driver with the faithful FP64 solvers (bit-identical to the reference
    arithmetic); the first also with the oracle on the host cores for a
    bounded number of steps."""
    import torch

    from transportse_amd.driver import Simulation, make_hip_backend
    cfg_dir = os.path.join(ROOT, "transportse_amd", "configs")
    out = {}
    small, large = "salt_64x32x32", "salt_heat_sia_256x256x128"
    for name, steps in ((small, args.a_steps), (large, args.b_steps)):
        cfg = name + ".py"
        sim = Simulation(os.path.join(cfg_dir, cfg), make_hip_backend,
                         verbose=False, max_steps=steps)
        n = sim.st.nx * sim.st.ny * sim.st.nz
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run()
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        sm = sim.summary
        rec = {"grid": [sim.st.nx, sim.st.ny, sim.st.nz], "steps": sm["n_tsteps"],
               "seconds": sec, "steps_per_s": sm["n_tsteps"] / sec,
               "p_sweeps": sm["numitP"], "c_sweeps": sm["numitC"],
               "t_sweeps": sm["numitT"], "solver_seconds": sim.solve_time,
               "p_cell_updates_per_s_end_to_end": sm["numitP"] * n / sec}
        log(f"{name} faithful FP64: {sm['n_tsteps']} steps in {sec:.2f}s, "
            f"P sweeps {sm['numitP']}")
        out[name] = rec
        del sim
    if args.a_cpu_steps and not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.build()
        threads = len(os.sched_getaffinity(0))
        orc.set_threads(threads)

        def oracle_backend(sim):
            st, c = sim.st, sim.config
            return orc.Oracle(st.nx, st.ny, st.nz, st.dx, st.dy, st.dz, st.g,
                              st.beta_deg, frozen_source=st, maxit_p=sim.MAXIT_P,
                              maxit_c=sim.MAXIT_C, maxit_t=sim.MAXIT_T,
                              dt_max=c.dt_max, c_imblflo=c.C_IMBLFLO,
                              t_imblflo=c.T_IMBLFLO)
        sim = Simulation(os.path.join(cfg_dir, small + ".py"), oracle_backend,
                         verbose=False, max_steps=args.a_cpu_steps)
        t0 = time.perf_counter()
        sim.run()
        sec = time.perf_counter() - t0
        out[small]["cpu_oracle"] = {"steps": sim.summary["n_tsteps"], "seconds": sec,
                                  "steps_per_s": sim.summary["n_tsteps"] / sec,
                                  "cores": threads,
                                  "gpu_speedup": out[small]["steps_per_s"]
                                  / (sim.summary["n_tsteps"] / sec)}
        log(f"{small} oracle on {threads} host threads: "
            f"{sim.summary['n_tsteps']} steps in {sec:.2f}s")
    return out


def sia_section(args, nx, ny, nz):
    """SIA steps with heat + salt + tracer, mixed vs plain FP64 (to 1e-13)."""
    import numpy as np
    import torch

    from transportse_amd.scenario import DeviceScenario, ScenarioSpec, host_fields
    spec = ScenarioSpec(nx, ny, nz, spacing=5.0, sigma=2.5, corr=16, seed=3,
                        species=2, heat=True, salt_patch=4, dt=86400.0,
                        sia_maxit=2, sia_conv=1e-6)
    t0 = time.time()
    hf = host_fields(spec)
    log(f"sia: inputs in {time.time() - t0:.1f}s")
    runs = {}
    ref_tol = 1e-12
    for form, tol in (("mixed", args.tol), ("plain64", ref_tol)):
        scn = DeviceScenario(spec, fields=hf)
        scn.setup_p()
        plan = scn.fp.plan(form)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = scn.run(args.sia_steps, form=form, tol=tol, plan_p=plan,
                      plan_tr=(2, 16, 1, 64, 2), inner_p=99, inner_tr=10,
                      max_sweeps=30000, strict=(form == "mixed"), adaptive=True,
                      progress=lambda k, r: log(
                          f"  sia {form} {'P' if k == 0 else 'C/T'} solve: "
                          f"{r.sweeps} sweeps, rel {r.rel_residual:.2e}, "
                          f"{r.seconds:.2f}s"))
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        runs[form] = (scn.download(), sec, out, plan, scn.omega_p.omega)
        scn.close()
        del scn
        torch.cuda.empty_cache()
        log(f"sia {form}: {args.sia_steps} steps in {sec:.2f}s, P sweeps "
            f"{[r.sweeps for k, r in out if k == 0]}")
    m, ref = runs["mixed"][0], runs["plain64"][0]

    def err(a, b):
        return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                "max_abs": float(np.max(np.abs(a - b)))}
    errors = {"P": err(m["P"], ref["P"]), "T": err(m["T"], ref["T"])}
    for s_ in range(spec.species):
        errors[f"C{s_}"] = err(m["C"][s_], ref["C"][s_])
    log(f"sia errors mixed vs FP64: " + ", ".join(
        f"{k} {v['rel_l2']:.2e}" for k, v in errors.items()))
    return {"grid": [nx, ny, nz], "steps": args.sia_steps,
            "species": spec.species, "heat": True, "sia_maxit": spec.sia_maxit,
            "seconds_mixed": runs["mixed"][1],
            "seconds_plain64": runs["plain64"][1],
            "p_sweeps_mixed": [r.sweeps for k, r in runs["mixed"][2] if k == 0],
            "p_sweeps_plain64": [r.sweeps for k, r in runs["plain64"][2] if k == 0],
            "tr_sweeps_mixed": [r.sweeps for k, r in runs["mixed"][2] if k == 1],
            "reference": f"plain FP64 form solved to rel. residual {ref_tol:g}"
                         " (or 30000 sweeps)",
            "reference_rel_residuals": [r.rel_residual for _, r in runs["plain64"][2]],
            "error_mixed_vs_fp64": errors}


def ensemble_section(args):
    import torch

    from transportse_amd.ensemble import Ensemble, ensemble_specs
    specs = ensemble_specs(args.ens_members)
    cells = sum(s.cells for s in specs)
    t0 = time.time()
    ens = Ensemble(specs, graph=True, inner_p=51, inner_tr=6)
    log(f"ensemble: {len(specs)} realizations, {cells / 1e6:.1f}M cells, set up "
        f"in {time.time() - t0:.1f}s")
    tuned = ens.tune_plans(cycles=2, zchunks=(0, 32))
    log(f"ensemble plans: P {ens.plan_p}, transport {ens.plan_tr}")
    st = ens.run(args.ens_steps)
    ens.close()
    del ens
    torch.cuda.empty_cache()
    log(f"ensemble: {args.ens_steps} steps in {st.seconds:.2f}s, "
        f"{st.cell_updates_per_s:.3e} cell-updates/s")
    return {"realizations": len(specs), "cells": cells, "steps": args.ens_steps,
            "seconds": st.seconds, "seconds_per_step": st.seconds / args.ens_steps,
            "realization_steps_per_s": len(specs) * args.ens_steps / st.seconds,
            "cell_updates_per_s": st.cell_updates_per_s, "solves": st.solves,
            "sweeps_p": st.sweeps_p, "sweeps_tr": st.sweeps_tr,
            "cycles": st.cycles, "graph": True,
            "plan_p": list(tuned[0][0]) if tuned.get(0) else None,
            "plan_tr": list(tuned[1][0]) if tuned.get(1) else None}


def serve_section(args):
    """Serving at each arrival rate: requests arrive over a window of
    --serve-window seconds (at least --serve-min-requests, at most
    --serve-max-requests of them), the same request list (grids, seeds) at
    every rate, with the 200 GB admission budget; one more run at the
    highest rate with a budget of --serve-tight-gb so that admission binds.
    Each run reports latency (arrival to completion) split into queueing
    (arrival to admission) and service (admission to completion)."""
    import math

    import torch

    from transportse_amd.serve import Server, make_requests, prepare
    count = lambda rate: int(min(args.serve_max_requests,
                                 max(args.serve_min_requests,
                                     math.ceil(rate * args.serve_window))))
    cache = {}
    t0 = time.time()
    prepare(make_requests(max(count(r) for r in args.serve_rates), 1.0,
                          steps=args.serve_steps), cache)
    log(f"serving: {len(cache)} request payloads prepared in {time.time() - t0:.1f}s")
    out = []
    runs = [(rate, 200e9) for rate in args.serve_rates]
    runs.append((max(args.serve_rates), args.serve_tight_gb * 1e9))
    for rate, budget in runs:
        reqs = make_requests(count(rate), rate, steps=args.serve_steps)
        prepare(reqs, cache)
        srv = Server(budget_bytes=budget)
        st = srv.run(reqs)
        srv.close()
        del srv, reqs
        torch.cuda.empty_cache()
        log(f"serving {rate}/s budget {budget / 1e9:.0f} GB: {st.completed} requests,"
            f" p50 {st.p50:.2f}s p99 {st.p99:.2f}s (queue p99 {st.queue_p99:.2f}s,"
            f" service p99 {st.service_p99:.2f}s), {st.solves_per_s:.2f} solves/s,"
            f" max {st.max_concurrent} concurrent, {st.waited} waited")
        out.append({"rate_per_s": rate, "budget_bytes": budget,
                    "requests": st.completed,
                    "arrival_window_s": count(rate) / rate,
                    "p50_s": st.p50, "p99_s": st.p99, "mean_s": st.mean,
                    "queue_p50_s": st.queue_p50, "queue_p99_s": st.queue_p99,
                    "service_p50_s": st.service_p50,
                    "service_p99_s": st.service_p99,
                    "waited_for_admission": st.waited,
                    "solves_per_s": st.solves_per_s, "seconds": st.seconds,
                    "peak_admitted_bytes": st.peak_bytes,
                    "max_concurrent": st.max_concurrent})
    cap = max(r["solves_per_s"] for r in out)
    return {"budget_bytes": 200e9, "steps_per_request": args.serve_steps,
            "grids": [[64, 64, 64], [128, 128, 128], [256, 256, 128]],
            "species": 3, "heat": True, "tol": 1e-10, "form": "mixed",
            "capacity_solves_per_s": cap,
            "note": "every offered rate above capacity_solves_per_s overloads the "
                    "server: the queue grows over the arrival window, so latency "
                    "grows with the rate and the window (no steady state)",
            "runs": out}


if __name__ == "__main__":
    sys.exit(main())
