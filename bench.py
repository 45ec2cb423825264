"""Benchmark of the TRANSPORTSE SOR hot path on one MI355X.

Workload (default): the 512x512x256 pressure problem (5 m cells, lognormal
permeability sigma_lnK = 2.5 with exponential correlation length 16 cells,
seed 3, hydrostatic Dirichlet x faces, salt patch on top).  One step = one
compute_P call of `--sweeps` red/black iterations of the faithful FP64 form
at fixed dt (coefficient assembly + sweeps), with inputs resident in HBM.

Reported: SOR cell-updates/s (whole job), the roofline of the dominant
kernel (the colour half-sweep) from HIP events around its launches on the
library's stream, and a CPU baseline: the oracle's OpenMP sweep on the host
cores over the same grid.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--sweeps S]
                       [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell update of the faithful FP64 record sweep:
# 13 record words + own unknown read + write + neighbour-colour read
FAITHFUL_BYTES_PER_UPDATE = 16 * 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--grid", type=int, nargs=3, default=[512, 512, 256])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sweeps", type=int, default=1)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not "
              "supported (the hot path shards by slab, but no multi-GPU run "
              "is part of this build)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from transportse_amd import _lib
    from transportse_amd.device import HipContext
    from transportse_amd.workloads import PressureProblem

    nx, ny, nz = args.grid
    t0 = time.time()
    prob = PressureProblem(nx, ny, nz, spacing=5.0, sigma=2.5, corr_len=16,
                           seed=3)
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")
    N = nx * ny * nz
    hip = HipContext(*prob.grid_args())
    lib = hip.lib
    d = {n: hip.upload(a) for n, a in prob.p_inputs().items()}
    d["P"] = hip.upload(prob.P)
    d["Pn"] = hip.upload(prob.P)
    pf = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
    dt, sor = 86400.0, 1.8
    torch.cuda.synchronize()
    log(f"[bench] uploaded, field {hip.F} doubles")

    def step():
        _lib.check(lib.tse_p_sweeps(hip.ctx, ctypes.byref(pf), dt, sor,
                                    args.sweeps, 1), "tse_p_sweeps")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f}s")
    updates = float(N) * args.sweeps * args.steps
    value = updates / elapsed

    # dominant kernel: half-sweeps only, HIP events on the library's stream
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    n_it = max(args.sweeps, 10)
    ev0.record(stream)
    _lib.check(lib.tse_p_sweeps(hip.ctx, ctypes.byref(pf), dt, sor, n_it, 0),
               "tse_p_sweeps")
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    per_launch_s = ms * 1e-3 / (2 * n_it)
    bytes_per_launch = FAITHFUL_BYTES_PER_UPDATE * N / 2
    achieved = bytes_per_launch / per_launch_s / 1e9
    log(f"[bench] half-sweep {per_launch_s * 1e6:.1f} us, {achieved:.0f} GB/s")

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
                               prob.WP, dt, sor, colour)
        cpu_s = time.perf_counter() - c0
        cpu = {"value": N * args.cpu_sweeps / cpu_s, "unit": "cell-updates/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_sweeps} red+black sweep(s) of the full "
                         f"{nx}x{ny}x{nz} grid, oracle C restatement "
                         f"(OpenMP, {threads} threads), {cpu_s:.1f}s"}
        log(f"[bench] cpu baseline {cpu['value']:.3e} cell-updates/s")

    line = {
        "metric": "SOR cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"pressure SOR {nx}x{ny}x{nz}, faithful FP64 "
                               f"red-black sweeps ({args.sweeps} per step, "
                               "omega 1.8, sigma_lnK 2.5)",
                   "grid": [nx, ny, nz], "sweeps_per_step": args.sweeps,
                   "form": "faithful64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": "p_sweep_kernel",
                     "bytes_per_update": FAITHFUL_BYTES_PER_UPDATE,
                     "launch_us": per_launch_s * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
