"""GPU: sweep variant 5 -- FP32 cell-pair sweeps on an inner operator stored
as bfloat16 (the three faces and the storage term V R of each cell; the
diagonal is summed in the kernel from the rounded values; no validity mask).

1. One correction cycle of the mixed solve equals, to FP32 rounding, the same
   number of red/black SOR sweeps (tse.py:705-716's order: red first) run in
   FP64 by numpy on A' e = r, where A' is built here from the device's FP64
   operator with every face and V R rounded to bfloat16 (round to nearest
   even), D' = V R' + sum of the rounded faces, and r = b - A p is the
   FP32-rounded FP64 residual.  The same cycle of variant 3 equals the sweeps
   on the FP32-rounded operator (control), and the two models differ by far
   more than the tolerance, so the check sees which operator the kernel
   used.  Tolerance: 2e-5 of max|e|.  Variant 7 (the rhs stored as FP16,
   scaled by a power of two and rounded stochastically from its FP32 value
   with a dither fixed by the cell index; the scale cancels in the result,
   since it is exact and the sweeps are linear) equals the sweeps on the
   same operator with the rhs rounded that way, and differs from variant
   5's model by far more than the tolerance.
2. Solves to relative residual 1e-10 reach the oracle's converged compute_P
   fixed point within the error that residual bounds (1e-6 of the change),
   and agree with variant 3's solves (sweep counts within 10 %), with fixed
   and adaptive cycles, on grids with partial, single and unit-extent tiles.
"""
import ctypes
import types

import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu

GRIDS = [(84, 41, 23), (40, 37, 21), (160, 90, 30), (64, 33, 2), (2, 20, 30),
         (52, 1, 60)]


def _setup(f, dt):
    from transportse_amd import _lib
    from transportse_amd.device import HipContext
    hip = HipContext(*f.grid_args())
    d = {n: hip.upload(getattr(f, n)) for n in
         ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz", "BP", "WP", "c_r")}
    d["P"] = hip.upload(f.P)
    d["Pn"] = hip.upload(f.P)
    pf = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
    _lib.check(hip.lib.tse_fast_p_setup(hip.ctx, ctypes.byref(pf), dt),
               "tse_fast_p_setup")
    return hip, d


def _result(hip, d):
    from transportse_amd import _lib
    _lib.check(hip.lib.tse_fast_p_result(hip.ctx, d["P"].data_ptr()),
               "tse_fast_p_result")
    return hip.download(d["P"])


def _operator(hip, shape):
    """The FP64 operator as host arrays indexed [i, j, k]."""
    import torch
    from transportse_amd import _lib
    nx, ny, nz = shape
    bufs = [torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
            for _ in range(5)]
    _lib.check(hip.lib.tse_fast_p_operator(hip.ctx, *[b.data_ptr() for b in bufs]),
               "tse_fast_p_operator")
    torch.cuda.synchronize()
    return [b.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0).copy()
            for b in bufs]


def _bf16(w):
    """Round to bfloat16, nearest even, via the FP32 value (what the
    assembly does)."""
    u = w.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _lowbias32(x):
    """The lowbias32 integer mix (C. Wellons) on uint32 arrays."""
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def _fp16_dithered(r):
    """Variant 7's rhs: the FP32 value plus a dither of hash(t) mod 2^13 in
    its 13 low mantissa bits, truncated to FP16's 10; t is the natural-layout
    index (x fastest) of the cell.  The kernel's power-of-two scale changes
    only exponents, so it drops out (FP16 subnormals aside, whose absolute
    error is below 2^-25 of the residual's RMS).  r is indexed [i, j, k]."""
    nx, ny, nz = r.shape
    t = np.arange(nx * ny * nz, dtype=np.uint64).reshape(nz, ny, nx).transpose(2, 1, 0)
    d = _lowbias32(t.astype(np.uint32)) & np.uint32(0x1FFF)
    u = r.astype(np.float32).view(np.uint32) + d
    return (u & np.uint32(0xFFFFE000)).view(np.float32).astype(np.float64)


def _faces(Wx, Wy, Wz):
    """The six faces of every cell: x+, x-, y+, y-, z+, z-."""
    z = np.zeros_like
    xm, ym, zm = z(Wx), z(Wy), z(Wz)
    xm[1:] = Wx[:-1]
    ym[:, 1:] = Wy[:, :-1]
    zm[:, :, 1:] = Wz[:, :, :-1]
    return Wx, xm, Wy, ym, Wz, zm


def _apply_offdiag(W6, e):
    """sum_f W_f e_f over the six neighbours (zero outside the grid)."""
    out = np.zeros_like(e)
    out[:-1] += W6[0][:-1] * e[1:]
    out[1:] += W6[1][1:] * e[:-1]
    out[:, :-1] += W6[2][:, :-1] * e[:, 1:]
    out[:, 1:] += W6[3][:, 1:] * e[:, :-1]
    out[:, :, :-1] += W6[4][:, :, :-1] * e[:, :, 1:]
    out[:, :, 1:] += W6[5][:, :, 1:] * e[:, :, :-1]
    return out


def _model_sweeps(W6, invD, r, n, sor, shape):
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    colour = (i + j + k) & 1
    e = np.zeros(shape)
    for _ in range(n):
        for c in (0, 1):
            upd = (colour == c) & (invD > 0)
            new = e + sor * (invD * (r + _apply_offdiag(W6, e)) - e)
            e = np.where(upd, new, e)
    return e


@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("ns,zchunk", [(3, 0), (2, 7), (3, 9)])
def test_one_cycle_equals_sor_on_the_rounded_operator(shape, ns, zchunk):
    from transportse_amd import _lib
    f = Fields(*shape, seed=31, dirichlet_p=shape[0] > 2)
    sor, n = 1.7, 4 * ns
    fields = {}
    for var in (3, 5, 7):
        hip, d = _setup(f, 3600.0)
        Wx, Wy, Wz, D, b = _operator(hip, shape)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(ns, 32, 2, zchunk, var),
                              inner=n, max_sweeps=n, sor=sor, tol=0.0)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.sweeps == n and res.cycles == 1
        fields[var] = _result(hip, d) - f.P
    p0 = f.P
    W6 = _faces(Wx, Wy, Wz)
    active = (D > 0) & (f.BP == 0)
    r = b - D * p0 + _apply_offdiag(W6, p0)
    r32 = np.where(active, r, 0.0).astype(np.float32).astype(np.float64)
    # variant 3: FP32 faces and 1/diagonal
    W6_32 = [w.astype(np.float32).astype(np.float64) for w in W6]
    inv32 = np.where(active, 1.0 / D, 0.0).astype(np.float32).astype(np.float64)
    # variant 5: bfloat16 faces and storage term, diagonal = their sum
    W6_16 = [_bf16(w) for w in W6]
    VR = D - sum(W6)
    D5 = _bf16(np.maximum(VR, 0.0)) + sum(W6_16)
    inv5 = np.where(active & (D5 > 0), 1.0 / D5, 0.0)
    e32 = _model_sweeps(W6_32, inv32, r32, n, sor, shape)
    e16 = _model_sweeps(W6_16, inv5, r32, n, sor, shape)
    # variant 7: the rhs rounded to FP16 from its FP32 value (dithered)
    e7 = _model_sweeps(W6_16, inv5, _fp16_dithered(r32), n, sor, shape)
    scale = np.max(np.abs(e32))
    d3 = np.max(np.abs(fields[3] - e32)) / scale
    d5 = np.max(np.abs(fields[5] - e16)) / scale
    d7 = np.max(np.abs(fields[7] - e7)) / scale
    gap = np.max(np.abs(e16 - e32)) / scale
    gap7 = np.max(np.abs(e7 - e16)) / scale
    print("one cycle", shape, (ns, zchunk), "variant 3 vs FP32 model", d3,
          "variant 5 vs bf16 model", d5, "variant 7 vs fp16-rhs model", d7,
          "models apart", gap, gap7, flush=True)
    assert scale > 0
    assert d3 <= 2e-5 and d5 <= 2e-5 and d7 <= 2e-5
    if shape[0] * shape[1] * shape[2] > 1000:
        assert gap > 10 * max(d5, 1e-7)
        assert gap7 > 10 * max(d7, 1e-7)


@pytest.mark.parametrize("variant", [5, 7])
def test_solve_reaches_oracle_fixed_point(variant):
    from oracle import Oracle
    from transportse_amd import _lib
    f = Fields(40, 30, 24, seed=22)
    dt = 3600.0
    o = Oracle(*f.grid_args(), frozen_source=types.SimpleNamespace(c_r=f.c_r))
    Po = f.P.copy()
    r = o.compute_P(Po, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx,
                    f.qy, f.qz, dt, f.BP, f.WP, SOR=1.9, max_it=200000,
                    inc_it=1.0, max_dt=dt, P_conv=1e-15)
    assert r[4]
    hip, d = _setup(f, dt)
    prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(3, 32, 2, 0, variant), inner=18,
                          max_sweeps=200000, sor=1.9, tol=1e-10)
    res = _lib.FastResult()
    _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                        ctypes.byref(res)), "tse_fast_p_solve")
    assert res.converged, (res.sweeps, res.rel_residual)
    Ph = _result(hip, d)
    change = np.max(np.abs(Po - f.P))
    err = np.max(np.abs(Ph - Po))
    print(f"variant {variant} solve: sweeps", res.sweeps, "rel", res.rel_residual,
          "err/change", err / change, flush=True)
    assert err <= 1e-6 * change


@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("mode", [0, 1])
def test_solves_agree_with_the_fp32_operator(shape, mode):
    """Fixed and adaptive cycles: variants 5 and 7 converge to the same
    tolerance with about the same sweep count as variant 3, and the solutions
    agree within what the residual bounds."""
    from transportse_amd import _lib
    f = Fields(*shape, seed=32, dirichlet_p=shape[0] > 2)
    out = {}
    for var in (3, 5, 7):
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(3, 32, 2, 0, var), inner=18,
                              max_sweeps=60000, cycle_mode=mode, sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged and res.rel_residual <= 1e-10
        out[var] = (_result(hip, d), res.sweeps, res.cycles)
    a, na, ca = out[3]
    change = np.max(np.abs(a - f.P))
    for var in (5, 7):
        b, nb, cb = out[var]
        print(f"variant {var} vs 3", shape, "mode", mode, "sweeps", na, nb, "cycles",
              ca, cb, "diff/change", np.max(np.abs(a - b)) / change, flush=True)
        assert nb <= 1.1 * na + 36
        assert np.max(np.abs(a - b)) <= 1e-6 * change


@pytest.mark.parametrize("shape", [(83, 41, 23), (33, 20, 17), (1, 20, 30)])
def test_odd_nx_and_limits(shape):
    """An odd nx runs variants 5 and 7 on the padded layout (one inactive
    column) and reaches the oracle's converged fixed point like variant 3;
    FP64 sweeps have no variant 5."""
    from oracle import Oracle
    from transportse_amd import _lib
    f = Fields(*shape, seed=27, dirichlet_p=shape[0] > 2)
    dt = 3600.0
    o = Oracle(*f.grid_args(), frozen_source=types.SimpleNamespace(c_r=f.c_r))
    Po = f.P.copy()
    r = o.compute_P(Po, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx,
                    f.qy, f.qz, dt, f.BP, f.WP, SOR=1.9, max_it=400000,
                    inc_it=1.0, max_dt=dt, P_conv=1e-15)
    assert r[4]
    change = np.max(np.abs(Po - f.P))
    for var in (3, 5, 7):
        hip, d = _setup(f, dt)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(3, 32, 2, 0, var), inner=18,
                              max_sweeps=400000, sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged
        err = np.max(np.abs(_result(hip, d) - Po))
        print("odd nx", shape, "variant", var, "sweeps", res.sweeps, "err/change",
              err / change, flush=True)
        assert err <= 1e-6 * change
    prm.form = 1   # FP64 sweeps have no variant 5
    assert hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm), ctypes.byref(res)) < 0
