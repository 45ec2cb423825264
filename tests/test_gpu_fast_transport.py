"""GPU: the performance forms of the transport solves (compute_C, compute_T).

The fast form solves the fixed point of the reference update (tse.py:870-875,
1038-1043) with the reference's coefficients, by fused red/black z-marches,
to a relative residual.  It must land on the oracle's fully converged
compute_C / compute_T result (update norm 1e-15, far below what the fast
tolerances leave):
  * plain FP64 to rel. residual 1e-13: within 1e-10 of max|x|,
  * FP32 sweeps in FP64 defect correction to 1e-11: within 1e-8 of max|x|,
for every kernel variant, with and without z-chunks, weak and strong
advection, omega < 1, both imbalance-flow settings, Dirichlet cells.
The reported residual must agree with a fresh residual evaluation.
"""
import types

import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu

# (ns, by, ry, zchunk, variant): every transport kernel variant, with and
# without z-chunks; variant 2 rotates its rings by unrolling, reads through
# buffer loads and takes x neighbours from the adjacent lanes
PLANS64 = [(1, 8, 1, 0, 0), (1, 16, 1, 5, 0), (2, 16, 1, 0, 0), (2, 16, 1, 6, 0),
           (1, 8, 1, 0, 2), (2, 16, 1, 6, 2)]
PLANS32 = [(1, 16, 1, 0, 0), (2, 16, 1, 7, 0), (3, 16, 1, 0, 0), (3, 16, 1, 9, 0),
           (1, 16, 1, 5, 2), (2, 16, 1, 0, 2), (3, 16, 1, 9, 2)]
SHAPES = [(70, 23, 19), (33, 20, 17), (9, 40, 12)]


def _frozen(f):
    return types.SimpleNamespace(c_r=f.c_r, l_r=f.l_r, cp_r=f.cp_r,
                                 rho_r=f.rho_r, lf_l=f.lf_l, lf_t=f.lf_t,
                                 lt_l=f.lt_l, lt_t=f.lt_t)


def _oracle_c(f, dt, omega, imbl, n=0, sor=1.1):
    from oracle import Oracle
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f), maxit_c=20000,
               dt_max=1e12, c_imblflo=imbl)
    C = f.C[..., n].copy()
    r = o.compute_C(C, f.D_fx[..., n], f.D_fy[..., n], f.D_fz[..., n], f.phi,
                    f.rho_f, f.qx, f.qy, f.qz, dt, f.BC, f.WC[..., n], f.R, f.e,
                    omega=omega, SOR=sor, max_it=20000, C_conv=1e-15)
    assert r[3] and o.last.num_halvings == 0
    return C


def _oracle_t(f, dt, omega, imbl):
    from oracle import Oracle
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f), maxit_t=20000,
               dt_max=1e12, t_imblflo=imbl)
    T = f.T.copy()
    r = o.compute_T(T, f.l_f, f.cp_f, f.rho_f, f.phi, f.qx, f.qy, f.qz, dt,
                    f.BT, f.WT, f.e, omega=omega, SOR=1.1, max_it=20000,
                    T_conv=1e-15)
    assert r[3] and o.last.num_halvings == 0
    return T


def _fast_c(f, dt, omega, imbl, n=0):
    from transportse_amd.transport import FastTransport
    ft = FastTransport(*f.grid_args())
    ft.load_c(f.C[..., n], f.D_fx[..., n], f.D_fy[..., n], f.D_fz[..., n],
              f.phi, f.rho_f, f.qx, f.qy, f.qz, dt, f.BC, f.WC[..., n], f.R,
              f.e, omega, lf_l=f.lf_l, lf_t=f.lf_t, imblflo=imbl)
    return ft


def _fast_t(f, dt, omega, imbl):
    from transportse_amd.transport import FastTransport
    ft = FastTransport(*f.grid_args())
    ft.load_t(f.T, f.l_f, f.cp_f, f.rho_f, f.phi, f.qx, f.qy, f.qz, dt, f.BT,
              f.WT, f.e, omega, l_r=f.l_r, cp_r=f.cp_r, rho_r=f.rho_r,
              lt_l=f.lt_l, lt_t=f.lt_t, imblflo=imbl)
    return ft


def _check(ft, ref, form, tol, bound, plan, what):
    st = ft.solve(form, tol=tol, sor=1.3, max_sweeps=4000, inner=8, plan=plan)
    x = ft.result()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    again = ft.residual()
    print(what, form, plan, "sweeps", st.sweeps, "rel", st.rel_residual,
          "err", err, flush=True)
    assert st.converged and st.rel_residual <= tol
    assert abs(again - st.rel_residual) <= 1e-3 * st.rel_residual + 1e-300
    assert err <= bound
    return x


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("plan", PLANS64)
def test_c_plain64_reaches_oracle_fixed_point(shape, plan):
    f = Fields(*shape, seed=41)
    dt = 86400.0 * 30
    ref = _oracle_c(f, dt, 1.0, True)
    _check(_fast_c(f, dt, 1.0, True), ref, "plain64", 1e-13, 1e-10, plan, "C")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("plan", PLANS32)
def test_t_mixed_reaches_oracle_fixed_point(shape, plan):
    f = Fields(*shape, seed=42)
    dt = 86400.0 * 30
    ref = _oracle_t(f, dt, 1.0, False)
    _check(_fast_t(f, dt, 1.0, False), ref, "mixed", 1e-11, 1e-8, plan, "T")


@pytest.mark.parametrize("which", ["C", "T"])
@pytest.mark.parametrize("omega,imbl,qscale", [(1.0, True, 100.0),
                                               (0.6, False, 1.0),
                                               (0.7, True, 30.0)])
def test_transport_cases(which, omega, imbl, qscale):
    f = Fields(48, 26, 21, seed=43)
    f.qx *= qscale
    f.qy *= qscale
    f.qz *= qscale
    dt = 86400.0 * 10
    if which == "C":
        ref = _oracle_c(f, dt, omega, imbl, n=1)
        ft = _fast_c(f, dt, omega, imbl, n=1)
    else:
        ref = _oracle_t(f, dt, omega, imbl)
        ft = _fast_t(f, dt, omega, imbl)
    _check(ft, ref, "plain64", 1e-13, 1e-10, (2, 16, 1, 0), which)
    ft.assemble()
    _check(ft, ref, "mixed", 1e-11, 1e-8, (3, 16, 1, 8), which)


def test_explicit_omega_zero():
    """omega = 0: the update has no neighbour term, so one unrelaxed sweep is
    the reference's accepted result (tse.py:897 accepts at once)."""
    f = Fields(33, 20, 17, seed=44)
    dt = 3600.0
    ref = _oracle_c(f, dt, 0.0, True, sor=1.0)
    ft = _fast_c(f, dt, 0.0, True)
    st = ft.solve("plain64", tol=1e-14, sor=1.0, max_sweeps=8, inner=1,
                  plan=(1, 16, 1, 0))
    x = ft.result()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print("omega=0 sweeps", st.sweeps, "err", err, flush=True)
    assert st.converged and err <= 1e-14


@pytest.mark.parametrize("shape", [(51, 51, 1), (200, 1, 60), (1, 1, 145),
                                   (1, 20, 30)])
def test_unit_dimension_grids_fast_forms(shape):
    """Pressure and transport fast forms on grids with a unit dimension reach
    the oracle's fixed points (plain FP64 and mixed)."""
    from oracle import Oracle
    from transportse_amd.pressure import FastPressure
    f = Fields(*shape, seed=45, dirichlet_p=shape[0] > 2)
    dt = 86400.0
    o = Oracle(*f.grid_args(), frozen_source=_frozen(f), dt_max=dt)
    Po = f.P.copy()
    r = o.compute_P(Po, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx,
                    f.qy, f.qz, dt, f.BP, f.WP, SOR=1.8, max_it=10 ** 6,
                    inc_it=1.0, max_dt=dt, P_conv=1e-15)
    assert r[4]
    fp = FastPressure(*f.grid_args())
    fp.load(f.P, dt, **{n: getattr(f, n) for n in ("c_f", "nue", "rho_f", "phi",
                                                    "Kx", "Ky", "Kz", "BP", "WP",
                                                    "c_r")})
    change = np.max(np.abs(Po - f.P))
    for form, tol, bound, plan in (("plain64", 1e-13, 1e-8, (2, 16, 1, 0, 1)),
                                   ("mixed", 1e-11, 1e-5, (2, 16, 1, 0, 1))):
        fp.assemble()
        st = fp.solve(form, tol=tol, sor=1.8, inner=20, plan=plan,
                      max_sweeps=200000)
        err = np.max(np.abs(fp.result() - Po))
        print("P", shape, form, "sweeps", st.sweeps, "err/change", err / change,
              flush=True)
        assert st.converged and err <= bound * change
    ref = _oracle_c(f, dt, 1.0, True)
    _check(_fast_c(f, dt, 1.0, True), ref, "plain64", 1e-13, 1e-10,
           (2, 16, 1, 0), "C")


@pytest.mark.parametrize("form,plan", [("plain64", (1, 8, 1, 0)), ("plain64", (2, 16, 1, 6)),
                                       ("mixed", (2, 16, 1, 7)), ("mixed", (3, 16, 1, 0))])
def test_transport_variant2_equals_variant0(form, plan):
    """The transport variant 2 reads the same operands in the same order as
    variant 0 (x neighbours from the adjacent lanes instead of LDS): whole
    solves are bit-identical, with the same sweep count."""
    f = Fields(67, 37, 19, seed=31)
    out = []
    for var in (0, 2):
        ft = _fast_c(f, 43200.0, 1.0, True)
        st = ft.solve(form, tol=1e-12 if form == "plain64" else 1e-11, sor=1.2,
                      max_sweeps=4000, inner=6 * plan[0], plan=plan + (var,))
        assert st.converged
        out.append((ft.result(), st.sweeps))
    print("transport variant 2 vs 0", form, plan, "identical",
          np.array_equal(out[0][0], out[1][0]), "sweeps", out[0][1], out[1][1],
          flush=True)
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("shape", [(51, 51, 1), (1, 20, 30), (64, 33, 2), (3, 2, 1),
                                   (200, 1, 9)])
def test_transport_variant2_on_unit_dimension_grids(shape):
    """Tiles whose edges lie on the grid boundary (variant 2 drops their
    halo there): the same solve as variant 0, bit for bit."""
    f = Fields(*shape, seed=32)
    out = []
    for var in (0, 2):
        ft = _fast_c(f, 43200.0, 1.0, True)
        st = ft.solve("plain64", tol=1e-12, sor=1.1, max_sweeps=4000, inner=4,
                      plan=(2, 16, 1, 0, var))
        out.append((ft.result(), st.sweeps, st.converged))
    print("transport unit-dimension", shape, "sweeps", out[0][1], out[1][1],
          "identical", np.array_equal(out[0][0], out[1][0]), flush=True)
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0])
