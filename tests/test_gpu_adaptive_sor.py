"""GPU: adaptive SOR inside a fast-form pressure solve (tse_fast_params
.adapt_sor).  Started at a factor well below the optimum, a solve that may
move its factor to Young's estimate (tse_sor_model, pinned on CPU against
an operator's SOR spectrum in tests/test_omega.py) reaches the same relative
residual in fewer sweeps, ends at a larger factor, and its solution agrees
with the fixed-factor solve's within what that residual bounds (1e-6 of
the change), in the mixed and the plain FP64 forms.  Started above the
optimum (2 - omega half the estimate's), adapting costs no sweeps and
leaves the factor near its start.  A batch rejects the flag.
"""
import ctypes

import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu


def _solve(f, form, plan, sor, adapt, tol):
    from transportse_amd import _lib
    from transportse_amd.device import HipContext
    hip = HipContext(*f.grid_args())
    d = {n: hip.upload(getattr(f, n)) for n in
         ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz", "BP", "WP", "c_r")}
    d["P"] = hip.upload(f.P)
    d["Pn"] = hip.upload(f.P)
    pf = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
    _lib.check(hip.lib.tse_fast_p_setup(hip.ctx, ctypes.byref(pf), 3600.0),
               "tse_fast_p_setup")
    prm = _lib.FastParams(form=form, plan=_lib.SweepPlan(*plan), inner=30,
                          max_sweeps=200000, cycle_mode=1, adapt_sor=int(adapt),
                          sor=sor, tol=tol)
    res = _lib.FastResult()
    _lib.check(hip.lib.tse_fast_p_solve(hip.ctx, ctypes.byref(prm), ctypes.byref(res)),
               "tse_fast_p_solve")
    _lib.check(hip.lib.tse_fast_p_result(hip.ctx, d["P"].data_ptr()), "tse_fast_p_result")
    return hip.download(d["P"]), res


@pytest.mark.parametrize("form,plan,tol", [(2, (3, 32, 2, 0, 7), 1e-10),
                                           (1, (2, 16, 1, 0, 2), 1e-12)])
def test_adaptive_sor_from_below(form, plan, tol):
    f = Fields(96, 64, 48, seed=41, dirichlet_p=True)
    P0, r0 = _solve(f, form, plan, 1.8, False, tol)
    P1, r1 = _solve(f, form, plan, 1.8, True, tol)
    change = np.max(np.abs(P0 - f.P))
    print("adaptive SOR", form, "fixed 1.8:", r0.sweeps, "sweeps;", "adaptive:",
          r1.sweeps, "sweeps, ended at", r1.sor, "diff/change",
          np.max(np.abs(P1 - P0)) / change, flush=True)
    assert r0.converged and r1.converged
    assert r0.sor == 1.8 and r1.sor > 1.85
    assert r1.sweeps <= 0.7 * r0.sweeps
    assert np.max(np.abs(P1 - P0)) <= 1e-6 * change


def test_adaptive_sor_from_above_does_not_hurt():
    """Started above the optimum (2 - omega half of the estimate's, where
    the contraction is about omega - 1), the adaptive solve takes no more
    sweeps than the fixed-factor one (5 % and one cycle of slack) and stays
    within 1/4 in log2(2 - omega) of its start."""
    import math
    f = Fields(64, 48, 40, seed=42, dirichlet_p=True)
    _, r_est = _solve(f, 2, (3, 32, 2, 0, 7), 1.9, True, 1e-10)
    om = 2.0 - (2.0 - r_est.sor) / 2.0
    _, r_fix = _solve(f, 2, (3, 32, 2, 0, 7), om, False, 1e-10)
    _, r_ad = _solve(f, 2, (3, 32, 2, 0, 7), om, True, 1e-10)
    print("from above: estimate", r_est.sor, "start", om, "fixed", r_fix.sweeps,
          "adaptive", r_ad.sweeps, "ended at", r_ad.sor, "contraction",
          r_ad.contraction, flush=True)
    assert r_fix.converged and r_ad.converged
    assert r_ad.sweeps <= 1.05 * r_fix.sweeps + 240
    assert abs(math.log2(2.0 - r_ad.sor) - math.log2(2.0 - om)) <= 0.25


def test_batch_rejects_adaptive_sor():
    import torch

    from transportse_amd import _lib
    assert torch.cuda.is_available()
    lib = _lib.load()
    b = ctypes.c_void_p()
    _lib.check(lib.tse_batch_create(ctypes.byref(b), 0, None), "tse_batch_create")
    try:
        prm = _lib.FastParams(form=2, plan=_lib.SweepPlan(3, 32, 2, 0, 5), inner=30,
                              max_sweeps=1000, cycle_mode=0, adapt_sor=1, sor=1.9,
                              tol=1e-10)
        assert lib.tse_batch_begin(b, 0, ctypes.byref(prm)) < 0
        assert b"single solves" in lib.tse_last_error()
    finally:
        lib.tse_batch_destroy(b)


@pytest.mark.parametrize("form,plan", [(2, (3, 32, 2, 0, 7)), (2, (3, 32, 2, 0, 5)),
                                       (1, (2, 16, 1, 0, 2))])
def test_solve_that_runs_no_cycle(form, plan):
    """Edge cases of the adaptive and FP16-rhs paths: a solve whose initial
    iterate already meets the tolerance runs no cycle (no contraction, the
    factor unchanged, the iterate unchanged bit for bit), and so does one on
    an all-Dirichlet system, whose residual and rhs norms are zero (the FP16
    rhs scale falls back to 1)."""
    f = Fields(40, 24, 20, seed=44, dirichlet_p=True)
    P, r = _solve(f, form, plan, 1.9, True, 1e9)
    assert r.converged and r.sweeps == 0 and r.cycles == 0
    assert r.contraction == 0.0 and r.sor == 1.9
    assert np.array_equal(P, f.P)
    f.BP = np.ones_like(f.BP)
    P, r = _solve(f, form, plan, 1.9, True, 1e-10)
    assert r.converged and r.sweeps == 0 and r.rel_residual == 0.0
    assert np.array_equal(P, f.P)
