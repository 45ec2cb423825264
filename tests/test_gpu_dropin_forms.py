"""GPU: the performance forms behind the reference's own signatures.

HipSolvers(p_form / c_form / t_form = "plain64" or "mixed") runs compute_P,
compute_C and compute_T (tse.py:649-651, 785-787, 921-923) through
tse_compute_p / _c / _t with TSE_FORM_PLAIN64 / TSE_FORM_MIXED: each
attempt solves the fixed point of the reference update to the relative
residual P_conv / C_conv / T_conv, and the reference's time-step control
runs around the attempts (tse.py:724-740, 896-914, 1065-1079).

Tolerances (stated per test): against the oracle's compute_P / compute_C /
compute_T converged to an update norm of 1e-15 at the same dt,
  * plain64 at relative residual 1e-12: max error <= 1e-8 of the change the
    solve makes (pressure) or 1e-8 of max |x| (transport);
  * mixed at 1e-10: <= 1e-6 of the change (pressure), 1e-7 of max |x|
    (transport).
The control paths are exact: t, the returned dt, the reset on failure, the
sweep count of a failed attempt.
"""
import types

import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu

TOL = {"plain64": (1e-12, 1e-8, 1e-8), "mixed": (1e-10, 1e-6, 1e-7)}


def _frozen(f):
    return types.SimpleNamespace(c_r=f.c_r, l_r=f.l_r, cp_r=f.cp_r,
                                 rho_r=f.rho_r, lf_l=f.lf_l, lf_t=f.lf_t,
                                 lt_l=f.lt_l, lt_t=f.lt_t)


def _hip(f, **kw):
    from transportse_amd.solvers import HipSolvers
    return HipSolvers(*f.grid_args(), frozen_source=_frozen(f), verbose=False,
                      **kw)


def _p_args(f, dt):
    return (f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx, f.qy, f.qz,
            dt, f.BP, f.WP)


@pytest.mark.parametrize("shape", [(40, 30, 24), (33, 20, 17)])
@pytest.mark.parametrize("form", ["plain64", "mixed"])
def test_compute_p_form_reaches_oracle_fixed_point(shape, form):
    from oracle import Oracle
    tol, bound, _ = TOL[form]
    f = Fields(*shape, seed=31)
    dt = 3600.0
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f))
    Po = f.P.copy()
    ro = o.compute_P(Po, *_p_args(f, dt), SOR=1.9, max_it=200000, inc_it=1.0,
                     max_dt=dt, P_conv=1e-15)
    assert ro[4]
    h = _hip(f, p_form=form)
    Ph = f.P.copy()
    t, dt_next, P_out, Pn, conv, nit = h.compute_P(
        Ph, *_p_args(f, dt), SOR=1.9, max_it=200000, inc_it=1.25,
        max_dt=4000.0, P_conv=tol)
    assert conv and P_out is Ph
    assert t == dt and dt_next == min(dt * 1.25, 4000.0)  # tse.py:724-729
    assert np.array_equal(Pn, f.P)                        # Pn = P at entry
    change = np.max(np.abs(Po - f.P))
    err = np.max(np.abs(Ph - Po))
    print("compute_P", form, shape, "sweeps", nit, "err/change", err / change,
          flush=True)
    assert err <= bound * change


@pytest.mark.parametrize("form", ["plain64", "mixed"])
def test_compute_p_form_halves_to_failure(form):
    """Every attempt exhausts max_it sweeps: dt halves until <= min_dt, the
    call returns converged=False with P reset to its entry value and
    max_it sweeps per attempt (tse.py:730-740)."""
    f = Fields(20, 14, 10, seed=32)
    h = _hip(f, p_form=form)
    P = f.P.copy()
    dt0, min_dt, max_it = 3600.0, 400.0, 6
    t, dt, P_out, Pn, conv, nit = h.compute_P(
        P, *_p_args(f, dt0), SOR=1.9, max_it=max_it, inc_it=1.25,
        max_dt=1e6, min_dt=min_dt, P_conv=1e-14)
    halvings = 0
    d = dt0
    while True:          # the reference's sequence of dt values
        d /= 2
        halvings += 1
        if d <= min_dt:
            break
    assert not conv and t == 0 and dt == d
    assert h.last.num_halvings == halvings and nit == halvings * max_it
    assert np.array_equal(P, f.P) and np.array_equal(Pn, f.P)


@pytest.mark.parametrize("form", ["plain64", "mixed"])
def test_compute_p_form_one_halving_then_accept(form):
    """A max_it that the attempt at dt exceeds and the attempt at dt/2 does
    not: the call halves once, accepts at dt/2 with the dt/2 attempt's
    result, and counts both attempts' sweeps."""
    from oracle import Oracle
    tol, bound, _ = TOL[form]
    f = Fields(24, 18, 12, seed=33)
    # a storage term comparable to the face terms, so that halving dt makes
    # the (Gauss-Seidel, SOR 1) attempt converge in markedly fewer sweeps
    f.c_f = np.full(f.shape, 1e-5)
    dt = 86400.0
    args = lambda d: _p_args(f, d)
    chosen = None
    h = _hip(f, p_form=form)
    for max_it in range(6, 1200, 6):
        P1 = f.P.copy()
        r1 = h.compute_P(P1, *args(dt), SOR=1.0, max_it=max_it, inc_it=1.0,
                         max_dt=dt, min_dt=0.75 * dt, P_conv=tol)
        P2 = f.P.copy()
        r2 = h.compute_P(P2, *args(dt / 2), SOR=1.0, max_it=max_it, inc_it=1.0,
                         max_dt=dt, min_dt=1e-8, P_conv=tol)
        if not r1[4] and r2[4] and h.last.num_halvings == 0:
            chosen = (max_it, P2, r2)
            break
    assert chosen is not None, "no max_it separates dt from dt/2"
    max_it, P2, r2 = chosen
    P = f.P.copy()
    t, dt_next, _, _, conv, nit = h.compute_P(
        P, *args(dt), SOR=1.0, max_it=max_it, inc_it=1.0, max_dt=dt,
        min_dt=1e-8, P_conv=tol)
    assert conv and h.last.num_halvings == 1 and t == dt / 2
    assert nit == max_it + r2[5]
    assert np.array_equal(P, P2)
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f))
    Po = f.P.copy()
    ro = o.compute_P(Po, *args(dt / 2), SOR=1.0, max_it=200000, inc_it=1.0,
                     max_dt=dt, P_conv=1e-15)
    assert ro[4]
    change = np.max(np.abs(Po - f.P))
    err = np.max(np.abs(P - Po))
    print("one halving", form, "max_it", max_it, "err/change", err / change,
          flush=True)
    assert err <= bound * change


def _c_args(f, dt):
    return (f.D_fx[..., 0], f.D_fy[..., 0], f.D_fz[..., 0], f.phi, f.rho_f,
            f.qx, f.qy, f.qz, dt, f.BC, f.WC[..., 0], f.R, f.e)


def _t_args(f, dt):
    return (f.l_f, f.cp_f, f.rho_f, f.phi, f.qx, f.qy, f.qz, dt, f.BT, f.WT,
            f.e)


@pytest.mark.parametrize("form", ["plain64", "mixed"])
@pytest.mark.parametrize("omega", [1.0, 0.6, 0.0])
def test_transport_forms_reach_oracle_fixed_point(form, omega):
    from oracle import Oracle
    tol, _, bound = TOL[form]
    f = Fields(33, 20, 17, seed=34)
    dt = 3600.0
    kw = dict(c_imblflo=True, t_imblflo=True, dt_max=1e6)
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f), **kw)
    h = _hip(f, c_form=form, t_form=form, **kw)
    conv_o = 1e-15 if omega else 1.0
    Co, Ch = f.C.copy(), f.C.copy()
    ro = o.compute_C(Co[..., 0], *_c_args(f, dt), omega=omega, SOR=1.0,
                     max_it=100000, C_conv=conv_o)
    rh = h.compute_C(Ch[..., 0], *_c_args(f, dt), omega=omega, SOR=1.0,
                     max_it=100000, C_conv=tol)
    assert ro[3] and rh[3] and rh[0] == dt and rh[1] == ro[1]
    errc = np.max(np.abs(Ch - Co)) / np.max(np.abs(Co))
    assert np.array_equal(Ch[..., 1:], f.C[..., 1:])   # other species untouched
    To, Th = f.T.copy(), f.T.copy()
    ro = o.compute_T(To, *_t_args(f, dt), omega=omega, SOR=1.0, max_it=100000,
                     T_conv=conv_o)
    rh = h.compute_T(Th, *_t_args(f, dt), omega=omega, SOR=1.0, max_it=100000,
                     T_conv=tol)
    assert ro[3] and rh[3] and rh[0] == dt and rh[1] == ro[1]
    errt = np.max(np.abs(Th - To)) / np.max(np.abs(To))
    print("transport", form, "omega", omega, "C err", errc, "T err", errt,
          flush=True)
    if omega == 0.0 and form == "plain64":
        # one explicit sweep in FP64: the same arithmetic up to rounding
        assert errc <= 1e-12 and errt <= 1e-12
    else:
        assert errc <= bound and errt <= bound


def test_driver_with_mixed_forms_follows_faithful_run(tmp_path):
    """The 64x32x32 salt config's first 10 steps through the SIA driver with
    every solver in the mixed form (CONV_P / CONV_C of the config read as
    relative residuals, here 1e-8).  The CFL clamp (tse.py:1289-1292) sets
    dt from the fluxes of each form's own solution, so the dt sequences agree
    to 1e-5 relative, and P and C stay within 1e-5 relative L2 of the
    faithful run (update norm 1e-8) at every step."""
    import functools
    import os

    from helpers import ROOT
    from transportse_amd.driver import Simulation, make_hip_backend
    cfg = os.path.join(ROOT, "transportse_amd", "configs", "salt_64x32x32.py")
    runs = {}
    for form in ("faithful64", "mixed"):
        be = functools.partial(make_hip_backend, p_form=form, c_form=form,
                               t_form=form)
        sim = Simulation(cfg, be, verbose=False, record_fields=True,
                         out_dir=str(tmp_path), max_steps=10)
        runs[form] = sim.run()
    worst = 0.0
    for s, (a, b) in enumerate(zip(runs["faithful64"], runs["mixed"])):
        assert abs(a["dt"] - b["dt"]) <= 1e-5 * a["dt"], (s, a["dt"], b["dt"])
        assert abs(a["t_sim"] - b["t_sim"]) <= 1e-5 * a["t_sim"]
        for k in ("P", "C"):
            rel = np.linalg.norm(b[k] - a[k]) / np.linalg.norm(a[k])
            worst = max(worst, rel)
            assert rel <= 1e-5, (s, k, rel)
    print("driver mixed vs faithful: 10 steps, worst rel L2", worst, flush=True)
