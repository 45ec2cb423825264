"""GPU parity: the HIP faithful FP64 path against the oracle, bit for bit.

Every hot-path function is called through the C-ABI (HipSolvers -> ctypes ->
libtse_hip.so) on seeded inputs and compared with the oracle on the same
inputs. The faithful form keeps the reference's operand order and
colour order, so fields must be identical (array_equal), iteration counts,
t and dt equal. The change norm is a parallel reduction on the GPU; its
rounding differs from the sequential sum by O(1e-16) relative, which can
only flip an accept decision if the norm lies within that margin of the
tolerance -- none of these cases does, and the iteration counts prove it.
"""
import numpy as np
import pytest

from helpers import CASES, GOLDEN, Fields, oracle_backend

pytestmark = pytest.mark.gpu

GRIDS = [(7, 6, 5), (9, 7, 5), (16, 9, 12), (33, 20, 17), (64, 32, 32)]


def _pair(f, **kw):
    from oracle import Oracle
    from transportse_amd.solvers import HipSolvers
    import types
    frozen = types.SimpleNamespace(c_r=f.c_r, l_r=f.l_r, cp_r=f.cp_r,
                                   rho_r=f.rho_r, lf_l=f.lf_l, lf_t=f.lf_t,
                                   lt_l=f.lt_l, lt_t=f.lt_t)
    o = Oracle(*f.grid_args(), frozen_source=frozen, **kw)
    h = HipSolvers(*f.grid_args(), frozen_source=frozen, verbose=False, **kw)
    return o, h


@pytest.mark.parametrize("shape", GRIDS)
def test_layout_roundtrip(shape):
    from transportse_amd.device import HipContext
    f = Fields(*shape, seed=1)
    ctx = HipContext(*f.grid_args())
    for a in (f.P, f.C[:, :, :, 1], f.Kx):
        d = ctx.upload(a)
        back = ctx.download(d)
        assert np.array_equal(back, a)
    print("layout ok", shape, flush=True)


@pytest.mark.parametrize("shape", GRIDS)
def test_compute_q_and_flow_error(shape):
    f = Fields(*shape, seed=2)
    o, h = _pair(f)
    qo = [f.qx.copy(), f.qy.copy(), f.qz.copy()]
    qh = [f.qx.copy(), f.qy.copy(), f.qz.copy()]
    o.compute_q(f.P, f.BP, f.Kx, f.Ky, f.Kz, *qo, f.nue, f.rho_f)
    h.compute_q(f.P, f.BP, f.Kx, f.Ky, f.Kz, *qh, f.nue, f.rho_f)
    for a, b in zip(qo, qh):
        assert np.array_equal(a, b)
    Pn = f.P + 10.0 * np.random.default_rng(3).standard_normal(f.shape)
    eo = o.flow_error(f.P, Pn, *qo, f.WP, f.phi, f.c_f, f.c_r, f.rho_f, 3600.0)
    eh = h.flow_error(f.P, Pn, *qh, f.WP, f.phi, f.c_f, f.c_r, f.rho_f, 3600.0)
    assert np.array_equal(eo, eh)
    print("q/e ok", shape, flush=True)


@pytest.mark.parametrize("shape", GRIDS)
def test_p_sweeps_bitwise(shape):
    """Fixed-dt red/black sweeps (no control) against the oracle half-sweeps."""
    import torch
    from transportse_amd import _lib
    import ctypes
    f = Fields(*shape, seed=4)
    o, h = _pair(f)
    Po = f.P.copy()
    Pn = f.P.copy()
    for _ in range(3):
        for colour in (0, 1):
            o.p_half_sweep(Po, Pn, f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky,
                           f.Kz, f.BP, f.WP, 3600.0, 1.8, colour)
    hx = h.hip
    d = {n: hx.upload(getattr(f, n)) for n in
         ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz", "BP", "WP", "c_r")}
    d["P"] = hx.upload(f.P)
    d["Pn"] = hx.upload(Pn)
    pf = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
    _lib.check(hx.lib.tse_p_sweeps(hx.ctx, ctypes.byref(pf), 3600.0, 1.8, 3, 1),
               "tse_p_sweeps")
    Ph = hx.download(d["P"])
    torch.cuda.synchronize()
    assert np.array_equal(Po, Ph)
    print("p sweeps ok", shape, flush=True)


@pytest.mark.parametrize("shape,sor,conv",
                         [(s, 1.8, 1e-8) for s in GRIDS]
                         + [(s, 1.0, 1e-10) for s in GRIDS[:3]])
def test_compute_p(shape, sor, conv):
    f = Fields(*shape, seed=5)
    o, h = _pair(f)
    Po, Ph = f.P.copy(), f.P.copy()
    args = (f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx, f.qy, f.qz,
            86400.0, f.BP, f.WP)
    kw = dict(SOR=sor, max_it=20000, inc_it=1.25, max_dt=86400.0 * 2,
              P_conv=conv)
    ro = o.compute_P(Po, *args, **kw)
    rh = h.compute_P(Ph, *args, **kw)
    assert ro[4] and rh[4]
    assert (ro[0], ro[1], ro[5]) == (rh[0], rh[1], rh[5])
    assert np.array_equal(ro[2], rh[2])
    assert np.array_equal(ro[3], rh[3])
    print("compute_P ok", shape, sor, "iters", ro[5], flush=True)


def test_compute_p_halving():
    f = Fields(6, 5, 7, seed=6, stretch=1.0)
    f.c_f[:] = 4.5e-8
    o, h = _pair(f)
    Po, Ph = f.P.copy(), f.P.copy()
    args = (f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx, f.qy, f.qz,
            2000.0, f.BP, f.WP)
    kw = dict(SOR=1.8, max_it=20, inc_it=1.25, max_dt=2000.0, P_conv=1e-8)
    ro = o.compute_P(Po, *args, **kw)
    rh = h.compute_P(Ph, *args, **kw)
    assert o.last.num_halvings > 0
    assert h.last.num_halvings == o.last.num_halvings
    assert (ro[0], ro[1], ro[4], ro[5]) == (rh[0], rh[1], rh[4], rh[5])
    assert np.array_equal(ro[2], rh[2])
    print("halving ok", o.last.num_halvings, flush=True)


# 65 x 65 x 21 = 88,725 slots per colour (odd): 174 sweep blocks, more than
# the 128 up to which the control step runs inside the next red sweep, so
# these solves use a control launch after every iteration
LARGE = (129, 65, 21)


def test_compute_p_control_launches():
    test_compute_p(LARGE, 1.8, 1e-8)


def test_compute_c_control_launches():
    test_compute_c(LARGE, 0.6, 1.2, False)


def test_compute_t_control_launches():
    test_compute_t(LARGE, 1, 1.0, True)


@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("omega,sor,imbl", [(1, 1.0, True), (0.6, 1.2, False),
                                            (0, 1.0, True)])
def test_compute_c(shape, omega, sor, imbl):
    f = Fields(*shape, seed=7)
    o, h = _pair(f, c_imblflo=imbl, dt_max=1e6)
    for n in range(f.S):
        Co, Ch = f.C.copy(), f.C.copy()
        args = (f.D_fx[..., n], f.D_fy[..., n], f.D_fz[..., n], f.phi, f.rho_f,
                f.qx, f.qy, f.qz, 3600.0, f.BC, f.WC[..., n], f.R, f.e)
        kw = dict(omega=omega, SOR=sor, max_it=100, C_conv=1e-10)
        ro = o.compute_C(Co[..., n], *args, **kw)
        rh = h.compute_C(Ch[..., n], *args, **kw)
        assert (ro[0], ro[1], ro[3], ro[4]) == (rh[0], rh[1], rh[3], rh[4])
        assert np.array_equal(Co, Ch)
    print("compute_C ok", shape, omega, flush=True)


@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("omega,sor,imbl", [(1, 1.0, True), (0.7, 1.3, False)])
def test_compute_t(shape, omega, sor, imbl):
    f = Fields(*shape, seed=8)
    o, h = _pair(f, t_imblflo=imbl, dt_max=1e6)
    To, Th = f.T.copy(), f.T.copy()
    args = (f.l_f, f.cp_f, f.rho_f, f.phi, f.qx, f.qy, f.qz, 3600.0, f.BT,
            f.WT, f.e)
    kw = dict(omega=omega, SOR=sor, max_it=100, T_conv=1e-10)
    ro = o.compute_T(To, *args, **kw)
    rh = h.compute_T(Th, *args, **kw)
    assert (ro[0], ro[1], ro[3], ro[4]) == (rh[0], rh[1], rh[3], rh[4])
    assert np.array_equal(ro[2], rh[2])
    print("compute_T ok", shape, omega, "iters", ro[4], flush=True)


@pytest.mark.parametrize("shape", GRIDS)
def test_smolarkiewicz(shape):
    """Anti-diffusive velocities from a species view with C <= 0 cells; the
    last index of each axis keeps the previous vx / vy / vz."""
    f = Fields(*shape, seed=9)
    o, h = _pair(f)
    rng = np.random.default_rng(10)
    C4 = f.C.copy()
    C4[..., 1] -= 0.2          # some cells <= 0
    q = [1e-6 * rng.standard_normal(f.shape) for _ in range(3)]
    v0 = [rng.standard_normal(f.shape) for _ in range(3)]
    vo = [a.copy() for a in v0]
    vh = [a.copy() for a in v0]
    o.smolarkiewicz(C4[..., 1], *q, *vo, 3600.0)
    h.smolarkiewicz(C4[..., 1], *q, *vh, 3600.0)
    for a, b in zip(vo, vh):
        assert np.array_equal(a, b)
    keep = (C4[-1, :, :, 1] > 0)
    assert np.array_equal(vo[0][-1][keep], v0[0][-1][keep])
    assert np.all(vo[0][C4[..., 1] <= 0] == 0)
    print("smolarkiewicz ok", shape, flush=True)


# unit dimensions, as in the reference's course examples (51x51x1 planes,
# 200x1x60 sections, 1x1x145 columns)
DEGENERATE = [(51, 51, 1), (1, 1, 145), (200, 1, 60), (1, 20, 30), (3, 2, 1)]


@pytest.mark.parametrize("shape", DEGENERATE)
def test_unit_dimension_grids(shape):
    f = Fields(*shape, seed=11, dirichlet_p=shape[0] > 2)
    o, h = _pair(f, dt_max=1e6)
    Po, Ph = f.P.copy(), f.P.copy()
    args = (f.c_f, f.nue, f.rho_f, f.phi, f.Kx, f.Ky, f.Kz, f.qx, f.qy, f.qz,
            3600.0, f.BP, f.WP)
    kw = dict(SOR=1.5, max_it=20000, inc_it=1.25, max_dt=7200.0, P_conv=1e-10)
    ro = o.compute_P(Po, *args, **kw)
    rh = h.compute_P(Ph, *args, **kw)
    assert (ro[0], ro[1], ro[4], ro[5]) == (rh[0], rh[1], rh[4], rh[5])
    assert np.array_equal(ro[2], rh[2])
    qo = [f.qx.copy(), f.qy.copy(), f.qz.copy()]
    qh = [f.qx.copy(), f.qy.copy(), f.qz.copy()]
    o.compute_q(Po, f.BP, f.Kx, f.Ky, f.Kz, *qo, f.nue, f.rho_f)
    h.compute_q(Ph, f.BP, f.Kx, f.Ky, f.Kz, *qh, f.nue, f.rho_f)
    for a, b in zip(qo, qh):
        assert np.array_equal(a, b)
    eo = o.flow_error(Po, ro[3], *qo, f.WP, f.phi, f.c_f, f.c_r, f.rho_f, 3600.0)
    eh = h.flow_error(Ph, rh[3], *qh, f.WP, f.phi, f.c_f, f.c_r, f.rho_f, 3600.0)
    assert np.array_equal(eo, eh)
    Co, Ch = f.C[..., 0].copy(), f.C[..., 0].copy()
    cargs = (f.D_fx[..., 0], f.D_fy[..., 0], f.D_fz[..., 0], f.phi, f.rho_f,
             *qo, 3600.0, f.BC, f.WC[..., 0], f.R, eo)
    ckw = dict(omega=1.0, SOR=1.0, max_it=100, C_conv=1e-10)
    rco, rch = o.compute_C(Co, *cargs, **ckw), h.compute_C(Ch, *cargs, **ckw)
    assert (rco[0], rco[1], rco[3], rco[4]) == (rch[0], rch[1], rch[3], rch[4])
    assert np.array_equal(Co, Ch)
    To, Th = f.T.copy(), f.T.copy()
    targs = (f.l_f, f.cp_f, f.rho_f, f.phi, *qo, 3600.0, f.BT, f.WT, eo)
    tkw = dict(omega=0.8, SOR=1.0, max_it=100, T_conv=1e-10)
    rto, rth = o.compute_T(To, *targs, **tkw), h.compute_T(Th, *targs, **tkw)
    assert (rto[0], rto[1], rto[3], rto[4]) == (rth[0], rth[1], rth[3], rth[4])
    assert np.array_equal(rto[2], rth[2])
    print("unit-dimension grid ok", shape, "P iterations", ro[5], flush=True)


@pytest.mark.parametrize("case", CASES)
def test_driver_matches_reference_trajectory(case, tmp_path):
    """Full SIA driver on the GPU against the reference's golden vectors."""
    from transportse_amd.driver import Simulation, make_hip_backend
    g = np.load(f"{GOLDEN}/{case}.npz")
    sim = Simulation(f"{GOLDEN}/configs/{case}.py", make_hip_backend,
                     verbose=False, record_fields=True, out_dir=str(tmp_path))
    traj = sim.run()
    assert len(traj) == g["P"].shape[0]
    for s, r in enumerate(traj):
        for k in ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f"):
            assert np.array_equal(r[k], g[k][s]), (case, s, k)
        for k in ("numitP", "numitC", "numitT"):
            assert r[k] == int(g[k][s]), (case, s, k)
        assert r["t_sim"] == g["t_end"][s]
    print("driver ok", case, flush=True)


def test_driver_output_tuples_match_reference(tmp_path):
    """PLOTSAVE on the GPU backend (golden g9): the output clamp's dt
    sequence and every element of the 34-tuples written at PLOTTIME equal
    the reference's (tse.py:312-329)."""
    from helpers import check_plot_files
    from transportse_amd.driver import Simulation, make_hip_backend
    sim = Simulation(f"{GOLDEN}/configs/g9_plot.py", make_hip_backend,
                     verbose=False, record_fields=True, out_dir=str(tmp_path))
    traj = sim.run()
    g = np.load(f"{GOLDEN}/g9_plot.npz")
    assert [r["dt"] for r in traj] == g["dt"].tolist()
    assert check_plot_files(str(tmp_path), "g9_plot") == 3
    print("output tuples ok: 3 files, dt", [r["dt"] for r in traj], flush=True)
