"""GPU: the performance forms of the pressure solve (plain FP64, mixed FP32).

1. The fused red/black z-march (NS sweeps per launch, z-chunks with halos,
   ping-pong buffers) must reproduce the same number of plain red/black
   sweeps of the oracle (reference arithmetic, tse.py:705-716). The fast
   form uses the volume-scaled symmetric operator, so rounding differs:
   tolerance 1e-10 of the change the sweeps make.
2. Solving to a relative residual target must land on the oracle's fully
   converged compute_P fixed point: plain FP64 to 1e-12 and mixed
   FP32-in-FP64 to 1e-10, each within the error that residual bounds.
"""
import ctypes
import types

import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu

# (ns, by, ry, zchunk, variant): every kernel instance, with and without
# z-chunks; variant 1 exchanges x neighbours across lanes instead of LDS,
# variant 2 is variant 1 with unrolled register rings and buffer loads
PLANS = [(1, 8, 1, 0, 0), (1, 16, 1, 5, 0), (2, 16, 1, 0, 0), (2, 16, 1, 6, 0),
         (1, 24, 2, 0, 0), (1, 24, 2, 7, 0),
         (1, 16, 1, 0, 1), (2, 16, 1, 5, 1), (1, 24, 2, 6, 1),
         (1, 16, 1, 0, 2), (2, 16, 1, 5, 2), (2, 16, 1, 0, 2), (1, 24, 2, 6, 2)]
PLANS32 = [(2, 32, 2, 6, 0), (2, 32, 4, 0, 0), (1, 32, 2, 5, 0),
           (3, 16, 1, 9, 0), (2, 24, 2, 6, 1), (2, 32, 2, 0, 1),
           (3, 16, 1, 7, 1), (3, 24, 2, 9, 1),
           (2, 24, 2, 6, 2), (2, 32, 2, 0, 2), (3, 16, 1, 7, 2), (3, 24, 2, 9, 2),
           (3, 32, 2, 0, 2), (2, 32, 2, 5, 2)]

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


@pytest.mark.parametrize("shape", [(70, 23, 19), (33, 20, 17), (9, 40, 12)])
@pytest.mark.parametrize("ns,by,ry,zchunk,variant", PLANS)
def test_fused_sweeps_match_red_black_sweeps(shape, ns, by, ry, zchunk, variant):
    from oracle import Oracle
    from transportse_amd import _lib
    f = Fields(*shape, seed=21)
    dt, sor, n = 3600.0, 1.7, 4
    o = Oracle(*f.grid_args(), frozen_source=types.SimpleNamespace(c_r=f.c_r))
    Po = f.P.copy()
    for _ in range(n):
        for colour in (0, 1):
            o.p_half_sweep(Po, f.P, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky,
                           f.Kz, f.BP, f.WP, dt, sor, colour)
    hip, d = _setup(f, dt)
    plan = _lib.SweepPlan(ns=ns, by=by, ry=ry, zchunk=zchunk, variant=variant)
    _lib.check(hip.lib.tse_fast_p_sweeps(hip.ctx, ctypes.byref(plan), n, sor, 0),
               "tse_fast_p_sweeps")
    Ph = _result(hip, d)
    change = np.max(np.abs(Po - f.P))
    err = np.max(np.abs(Ph - Po))
    print("fused", shape, ns, by, ry, zchunk, variant, "change", change, "err",
          err, flush=True)
    assert change > 0
    assert err <= 1e-10 * change


@pytest.mark.parametrize("form,tol,bound", [(1, 1e-12, 1e-8), (2, 1e-10, 1e-6)])
def test_solve_reaches_oracle_fixed_point(form, tol, bound):
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
    plan = _lib.SweepPlan(ns=2, by=16, ry=1, zchunk=0)
    prm = _lib.FastParams(form=form, plan=plan, inner=20, max_sweeps=200000,
                          sor=1.9, tol=tol)
    res = _lib.FastResult()
    _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                        ctypes.byref(res)), "tse_fast_p_solve")
    assert res.converged, (res.sweeps, res.rel_residual)
    Ph = _result(hip, d)
    change = np.max(np.abs(Po - f.P))
    err = np.max(np.abs(Ph - Po))
    print("solve form", form, "sweeps", res.sweeps, "rel", res.rel_residual,
          "err/change", err / change, "err/|P|", err / np.max(np.abs(Po)),
          flush=True)
    assert err <= bound * change


@pytest.mark.parametrize("ns,by,ry,zchunk,variant", PLANS32 + PLANS)
def test_fp32_plans_agree(ns, by, ry, zchunk, variant):
    """Every FP32 sweep plan gives the same mixed-precision solve (same
    sweep count to the same residual, fields within FP32-cycle noise)."""
    from transportse_amd import _lib
    f = Fields(70, 37, 21, seed=23)
    outs = []
    for pl in ((1, 16, 1, 0, 0), (ns, by, ry, zchunk, variant)):
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(*pl), inner=12,
                              max_sweeps=12000, sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged
        outs.append((_result(hip, d), res.sweeps))
    (a, na), (b, nb) = outs
    change = np.max(np.abs(a - f.P))
    print("fp32 plan", (ns, by, ry, zchunk, variant), "sweeps", na, nb,
          "diff/change", np.max(np.abs(a - b)) / change, flush=True)
    assert abs(na - nb) <= 24
    assert np.max(np.abs(a - b)) <= 1e-6 * change


@pytest.mark.parametrize("ns,by,ry", [(1, 16, 1), (2, 16, 1), (1, 24, 2)])
def test_lane_exchange_variant_equals_lds_variant(ns, by, ry):
    """Variant 1 reads the same operands in the same order as variant 0
    (x neighbours from the adjacent lanes instead of LDS): FP64 sweeps agree
    to rounding (the compiler may contract the two differently)."""
    from transportse_amd import _lib
    f = Fields(83, 41, 23, seed=24)
    out = []
    for var in (0, 1):
        hip, d = _setup(f, 3600.0)
        plan = _lib.SweepPlan(ns=ns, by=by, ry=ry, zchunk=9, variant=var)
        _lib.check(hip.lib.tse_fast_p_sweeps(hip.ctx, ctypes.byref(plan), 6 * ns,
                                             1.8, 0), "sweeps")
        out.append(_result(hip, d))
    change = np.max(np.abs(out[0] - f.P))
    diff = np.max(np.abs(out[0] - out[1]))
    print("variant 1 vs 0", (ns, by, ry), "identical", np.array_equal(*out),
          "diff/change", diff / change, flush=True)
    assert diff <= 1e-13 * change


@pytest.mark.parametrize("ns,by,ry,zchunk", [(1, 16, 1, 0), (2, 16, 1, 9),
                                             (2, 16, 1, 0), (1, 24, 2, 7)])
def test_unrolled_ring_variant_equals_lane_exchange_variant(ns, by, ry, zchunk):
    """Variant 2 reads the same operands in the same order as variant 1
    (rings rotated by unrolling, buffer loads that read 0 outside the grid):
    FP64 sweeps and whole mixed solves are bit-identical, including on a grid
    whose x and y extents leave partial tiles."""
    from transportse_amd import _lib
    f = Fields(83, 41, 23, seed=25)
    sw, sol = [], []
    for var in (1, 2):
        hip, d = _setup(f, 3600.0)
        plan = _lib.SweepPlan(ns=ns, by=by, ry=ry, zchunk=zchunk, variant=var)
        _lib.check(hip.lib.tse_fast_p_sweeps(hip.ctx, ctypes.byref(plan), 6 * ns,
                                             1.8, 0), "sweeps")
        sw.append(_result(hip, d))
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=2, plan=plan, inner=6 * ns, max_sweeps=12000,
                              sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged
        sol.append((_result(hip, d), res.sweeps))
    print("variant 2 vs 1", (ns, by, ry, zchunk), "sweeps identical",
          np.array_equal(*sw), "solve identical", np.array_equal(sol[0][0], sol[1][0]),
          "sweeps", sol[0][1], sol[1][1], flush=True)
    assert np.array_equal(sw[0], sw[1])
    assert sol[0][1] == sol[1][1] and np.array_equal(sol[0][0], sol[1][0])


# variant 3 (an x-pair of cells per thread): partial and single tiles in x
# and y, several middle tiles, unit y / z extents
PAIR_GRIDS = [(84, 41, 23), (40, 37, 21), (160, 90, 30), (64, 33, 2), (2, 20, 30),
              (52, 1, 60)]


@pytest.mark.parametrize("shape", PAIR_GRIDS)
@pytest.mark.parametrize("ns,zchunk", [(3, 0), (2, 7), (3, 9)])
def test_pair_variant_equals_unrolled_ring_variant(shape, ns, zchunk):
    """Variant 3 (each lane updates one cell of its x-pair in every phase)
    reads the same operands in the same order as variant 2: whole mixed
    solves (FP32 sweeps inside FP64 defect correction) are bit-identical."""
    from transportse_amd import _lib
    f = Fields(*shape, seed=26)
    sol = []
    for var in (2, 3):
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(ns=ns, by=32, ry=2,
                                                          zchunk=zchunk, variant=var),
                              inner=6 * ns, max_sweeps=12000, sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged
        sol.append((_result(hip, d), res.sweeps))
    print("variant 3 vs 2", shape, (ns, zchunk), "sweeps", sol[0][1], sol[1][1],
          "identical", np.array_equal(sol[0][0], sol[1][0]), flush=True)
    assert sol[0][1] == sol[1][1] and np.array_equal(sol[0][0], sol[1][0])


@pytest.mark.parametrize("shape", [(83, 41, 23), (33, 20, 17), (1, 20, 30), (9, 40, 12)])
def test_pair_variant_on_odd_nx(shape):
    """An odd nx: the natural layout gets one inactive column, so the cell-pair
    sweep (variant 3) runs on it, bit-identical to the unrolled-ring sweep
    (variant 2) on the same layout, single and batched; the solution matches
    the oracle's red/black sweeps (FP64 sweeps of variant 2 at 1e-10 of the
    change)."""
    from oracle import Oracle
    from transportse_amd import _lib
    from transportse_amd.batch import Batch
    from transportse_amd.pressure import FastPressure
    f = Fields(*shape, seed=27, dirichlet_p=shape[0] > 2)
    dt, sor, n = 3600.0, 1.7, 4
    o = Oracle(*f.grid_args(), frozen_source=types.SimpleNamespace(c_r=f.c_r))
    Po = f.P.copy()
    for _ in range(n):
        for colour in (0, 1):
            o.p_half_sweep(Po, f.P, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky,
                           f.Kz, f.BP, f.WP, dt, sor, colour)
    hip, d = _setup(f, dt)
    plan = _lib.SweepPlan(2, 16, 1, 0, 2)
    _lib.check(hip.lib.tse_fast_p_sweeps(hip.ctx, ctypes.byref(plan), n, sor, 0), "sweeps")
    change = np.max(np.abs(Po - f.P))
    assert np.max(np.abs(_result(hip, d) - Po)) <= 1e-10 * change + 1e-13 * np.max(np.abs(Po))
    sol = []
    for var in (2, 3):
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(3, 32, 2, 0, var), inner=18,
                              max_sweeps=12000, sor=1.9, tol=1e-10)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged
        sol.append((_result(hip, d), res.sweeps))
    assert sol[0][1] == sol[1][1] and np.array_equal(sol[0][0], sol[1][0])
    fp = FastPressure(*f.grid_args())
    fp.load(f.P, 3600.0, **{nm: getattr(f, nm) for nm in
                            ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz",
                             "BP", "WP", "c_r")})
    b = Batch(0, "mixed", (3, 32, 2, 0, 3), inner=18, tol=1e-10, sor=1.9,
              max_sweeps=12000)
    slot = b.add(fp)
    b.run(graph=False)
    ms = b.status(slot)
    b.writeback(slot)
    assert ms.converged and ms.sweeps == sol[1][1]
    assert np.array_equal(fp.result(), sol[1][0])
    b.close()
    print("odd nx", shape, "pair variant ok,", sol[1][1], "sweeps", flush=True)


DEGENERATE = [(51, 51, 1), (1, 1, 145), (200, 1, 60), (1, 20, 30), (3, 2, 1), (64, 33, 2)]


@pytest.mark.parametrize("shape", DEGENERATE)
def test_fused_variants_on_unit_dimension_grids(shape):
    """Grids with unit extents or fewer planes than the halo: every sweep
    variant matches the oracle's half-sweeps, and variant 2 (buffer loads
    that read 0 outside the grid) equals variant 1 bit for bit."""
    from oracle import Oracle
    from transportse_amd import _lib
    f = Fields(*shape, seed=26, dirichlet_p=shape[0] > 2)
    dt, sor, n = 3600.0, 1.7, 4
    o = Oracle(*f.grid_args(), frozen_source=types.SimpleNamespace(c_r=f.c_r))
    Po = f.P.copy()
    for _ in range(n):
        for colour in (0, 1):
            o.p_half_sweep(Po, f.P, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky,
                           f.Kz, f.BP, f.WP, dt, sor, colour)
    change = np.max(np.abs(Po - f.P))
    out = {}
    for pl in ((2, 16, 1, 0, 0), (2, 16, 1, 0, 1), (2, 16, 1, 0, 2), (1, 24, 2, 7, 1),
               (1, 24, 2, 7, 2)):
        hip, d = _setup(f, dt)
        plan = _lib.SweepPlan(*pl)
        _lib.check(hip.lib.tse_fast_p_sweeps(hip.ctx, ctypes.byref(plan), n, sor, 0),
                   "tse_fast_p_sweeps")
        out[pl] = _result(hip, d)
        err = np.max(np.abs(out[pl] - Po))
        print("unit-dimension", shape, pl, "change", change, "err", err, flush=True)
        assert err <= 1e-10 * change + 1e-13 * np.max(np.abs(Po))
    assert np.array_equal(out[(2, 16, 1, 0, 1)], out[(2, 16, 1, 0, 2)])
    assert np.array_equal(out[(1, 24, 2, 7, 1)], out[(1, 24, 2, 7, 2)])


@pytest.mark.parametrize("form,tol", [(2, 1e-10), (1, 1e-11)])
def test_adaptive_cycles_reach_the_same_tolerance(form, tol):
    """TSE_CYCLE_ADAPTIVE: the same stop test with cycle lengths predicted
    from the convergence rate -- converged to tol, fewer residual passes
    than fixed cycles, and the solution within tol of the fixed-cycle one."""
    from transportse_amd import _lib
    f = Fields(60, 44, 30, seed=27)
    out = {}
    for mode in (_lib.CYCLE_FIXED, _lib.CYCLE_ADAPTIVE):
        hip, d = _setup(f, 3600.0)
        prm = _lib.FastParams(form=form, plan=_lib.SweepPlan(2, 16, 1, 0, 2), inner=10,
                              max_sweeps=200000, cycle_mode=mode, sor=1.9, tol=tol)
        res = _lib.FastResult()
        _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm),
                                            ctypes.byref(res)), "solve")
        assert res.converged and res.rel_residual <= tol
        out[mode] = (_result(hip, d), res.sweeps, res.cycles)
    (a, na, ca), (b, nb, cb) = out[_lib.CYCLE_FIXED], out[_lib.CYCLE_ADAPTIVE]
    change = np.max(np.abs(a - f.P))
    print("adaptive cycles: fixed", na, "sweeps", ca, "cycles; adaptive", nb, "sweeps",
          cb, "cycles; diff/change", np.max(np.abs(a - b)) / change, flush=True)
    assert cb < ca
    assert nb <= na + 80 * 2
    assert np.max(np.abs(a - b)) <= 1e-6 * change


def test_bad_cycle_mode_rejected():
    from transportse_amd import _lib
    f = Fields(20, 12, 10, seed=28)
    hip, d = _setup(f, 3600.0)
    prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(2, 16, 1, 0, 2), inner=10,
                          max_sweeps=100, cycle_mode=7, sor=1.9, tol=1e-10)
    res = _lib.FastResult()
    assert hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm), ctypes.byref(res)) < 0
