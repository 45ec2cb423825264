"""Autotuner for the fused sweep kernels' launch plans.

A plan is (ns sweeps per launch, tile rows by, rows per thread ry, z-chunk
planes).  The tuner times every supported plan on the system already set up
in the context (``tse_fast_p_setup`` / ``tse_fast_c_setup`` /
``tse_fast_t_setup``) with HIP events on the library's stream, a few sweeps
each after one warm-up launch, and returns the fastest.  The FP64
candidates advance the iterate (harmless: the solve continues from any
iterate), the FP32 candidates run on the correction buffers.  Results are
cached per (system kind, grid, precision).
"""
from __future__ import annotations

import ctypes

from . import _lib

# (ns, by, ry, variant): variant 0 exchanges x neighbours through LDS,
# variant 1 across the lanes of a wavefront (two LDS plane rings, not three),
# variant 2 as 1 with register rings rotated by unrolling and buffer loads
# (fields up to 1 GiB)
PLANS64 = [(1, 8, 1, 0), (1, 16, 1, 0), (2, 16, 1, 0), (1, 24, 2, 0),
           (1, 16, 1, 1), (2, 16, 1, 1), (1, 24, 2, 1),
           (1, 16, 1, 2), (2, 16, 1, 2), (1, 24, 2, 2)]
PLANS32 = PLANS64 + [(1, 32, 2, 0), (2, 32, 2, 0), (2, 32, 4, 0), (3, 16, 1, 0),
                     (2, 24, 2, 1), (2, 32, 2, 1), (3, 16, 1, 1), (3, 24, 2, 1),
                     (2, 24, 2, 2), (2, 32, 2, 2), (3, 16, 1, 2), (3, 24, 2, 2),
                     (3, 32, 2, 2), (2, 32, 2, 3), (3, 32, 2, 3)]
# variant 5: the cell-pair sweep on the inner operator stored as bfloat16
# (faces and storage term; diagonal summed in the kernel)
PLANS32 = PLANS32 + [(2, 32, 2, 5), (3, 32, 2, 5)]
# the batched engine runs the same plans (pressure rows are always even: an
# odd nx gets one inactive column)
BATCH_PLANS32 = list(PLANS32)
# variant 7 (single systems): variant 5 with the correction's rhs stored as
# FP16 as well (18 instead of 20 bytes per cell), scaled by a power of two
# and rounded with a dither; the batched engine writes an FP32 rhs and does
# not run it
PLANS32 = PLANS32 + [(2, 32, 2, 7), (3, 32, 2, 7)]
# transport: the cell's own coefficients stay in registers, one row per thread
TR_PLANS64 = [(1, 8, 1, 0), (1, 16, 1, 0), (2, 16, 1, 0),
              (1, 8, 1, 2), (1, 16, 1, 2), (2, 16, 1, 2)]
TR_PLANS32 = TR_PLANS64 + [(3, 16, 1, 0), (3, 16, 1, 2)]
ZCHUNKS = (0, 128, 64, 32)
BUF_MAX = 1 << 30        # variants 2 and 3 read fields through buffer descriptors


def plan_fits(variant: int, shape, fp32: bool) -> bool:
    """Pressure plans on a grid of this shape (the natural layout's rows are
    nx rounded up to even): variants 2 and 3 take fields of at most 1 GiB,
    variants 5 and 7 FP32 sweeps with 8 bytes of record per cell within
    1 GiB; the library rejects the others."""
    nx, ny, nz = shape
    n = (nx + nx % 2) * ny * nz
    if variant in (5, 7):
        return fp32 and n * 8 <= BUF_MAX
    return variant < 2 or n * (4 if fp32 else 8) <= BUF_MAX

KINDS = {
    "p": ("tse_fast_p_sweeps", PLANS64, PLANS32),
    "tr": ("tse_fast_tr_sweeps", TR_PLANS64, TR_PLANS32),
}

_cache: dict = {}


def tune(hip, fp32: bool, sweeps: int = 12, sor: float = 1.9, zchunks=ZCHUNKS,
         verbose: bool = False, kind: str = "p"):
    """Return ((ns, by, ry, zchunk, variant), ms_per_sweep) of the fastest plan."""
    import torch
    key = (kind, hip.shape, bool(fp32))
    if key in _cache:
        return _cache[key]
    fn_name, p64, p32 = KINDS[kind]
    fn = getattr(hip.lib, fn_name)
    best = None
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    nz = hip.shape[2]
    for ns, by, ry, var in (p32 if fp32 else p64):
        if not plan_fits(var, hip.shape, fp32):
            continue
        for zc in zchunks:
            if zc and zc >= nz:
                continue
            n = ns * max(1, sweeps // ns)
            plan = _lib.SweepPlan(ns, by, ry, zc, var)
            _lib.check(fn(hip.ctx, ctypes.byref(plan), ns, sor, int(fp32)), "warmup")
            ev0.record(stream)
            _lib.check(fn(hip.ctx, ctypes.byref(plan), n, sor, int(fp32)), "sweeps")
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / n
            if verbose:
                print(f"  {kind} plan ns={ns} by={by} ry={ry} zchunk={zc} "
                      f"variant={var}: {ms * 1e3:.1f} us/sweep", flush=True)
            if best is None or ms < best[1]:
                best = ((ns, by, ry, zc, var), ms)
    _cache[key] = best
    return best
