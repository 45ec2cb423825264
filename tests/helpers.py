"""Shared test helpers: seeded field sets and backend factories."""
import math
import types
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ("g1_salt", "g2_sia_heat", "g3_halving", "g4_smolar", "g5_nlte",
         "g6_xz", "g7_column", "g8_xy", "g9_plot")


def oracle_backend(sim):
    from oracle import Oracle
    st, c = sim.st, sim.config
    return Oracle(st.nx, st.ny, st.nz, st.dx, st.dy, st.dz, st.g, st.beta_deg,
                  frozen_source=st, maxit_p=sim.MAXIT_P, maxit_c=sim.MAXIT_C,
                  maxit_t=sim.MAXIT_T, dt_max=c.dt_max, c_imblflo=c.C_IMBLFLO,
                  t_imblflo=c.T_IMBLFLO)


class Fields:
    """A seeded, physically plausible set of every hot-path input."""

    def __init__(self, nx, ny, nz, seed=0, beta=5.0, stretch=1.05, S=2,
                 dirichlet_p=True, bc_frac=0.05, disp=(1.0, 0.1),
                 zero_k_frac=0.0):
        rng = np.random.default_rng(seed)
        self.nx, self.ny, self.nz, self.S = nx, ny, nz, S
        shp = (nx, ny, nz)
        self.shape = shp
        self.dx = 10.0 * (1.0 + 0.2 * rng.random(nx))
        self.dy = np.full(ny, 8.0)
        self.dz = 2.0 * stretch ** np.arange(nz)[::-1]
        self.g = 9.80665
        self.beta = beta
        K = 1e-12 * np.exp(rng.standard_normal(shp))
        if zero_k_frac:
            K[rng.random(shp) < zero_k_frac] = 0.0
        self.Kx, self.Ky, self.Kz = K.copy(), 0.5 * K, 0.1 * K
        self.phi = 0.15 + 0.1 * rng.random(shp)
        self.c_f = np.full(shp, 4.5e-10)
        self.c_r = np.full(shp, 1e-10)
        self.nue = 1e-3 * (1.0 + 0.1 * rng.random(shp))
        C0 = rng.random(shp + (S,)) * 0.5
        self.C = C0
        self.rho_f = 1000.0 + 700.0 * C0[..., 0]
        self.cp_f = np.full(shp, 4150.0)
        self.l_f = 0.6 + 0.1 * rng.random(shp)
        self.cp_r = np.full(shp, 1000.0)
        self.l_r = 2.0 + rng.random(shp)
        self.rho_r = np.full(shp, 2500.0)
        self.lf_l = np.full(shp, disp[0])
        self.lf_t = np.full(shp, disp[1])
        self.lt_l = np.full(shp, disp[0])
        self.lt_t = np.full(shp, disp[1])
        self.R = 1.0 + 0.2 * rng.random(shp)
        Df = 1e-9 * (1 + rng.random(shp + (S,)))
        self.D_fx, self.D_fy, self.D_fz = Df.copy(), Df.copy(), 0.5 * Df
        self.BP = np.zeros(shp)
        if dirichlet_p:
            self.BP[0] = 1
            self.BP[-1] = 1
        self.BC = (rng.random(shp) < bc_frac).astype(float)
        self.BT = np.zeros(shp)
        self.BT[:, :, 0] = 1
        self.WP = np.zeros(shp)
        self.WP[nx // 2, ny // 2, nz // 2] = 1e-6
        self.WC = np.zeros(shp + (S,))
        self.WC[nx // 2, ny // 2, nz // 2, 0] = 1e-8
        self.WT = np.zeros(shp)
        self.WT[nx // 3, ny // 3, nz // 3] = 3.0
        b = math.radians(beta)
        z = np.cumsum(self.dz[::-1])[::-1] - 0.5 * self.dz
        x = np.cumsum(self.dx) - 0.5 * self.dx
        self.P = (1e5 + 1000 * self.g * (math.cos(b) * z[None, None, :]
                                         - math.sin(b) * x[:, None, None])
                  + 100.0 * rng.standard_normal(shp))
        self.P = np.broadcast_to(self.P, shp).copy()
        self.T = 10.0 + 50.0 * rng.random(shp)
        self.qx = 1e-7 * rng.standard_normal(shp)
        self.qy = 1e-7 * rng.standard_normal(shp)
        self.qz = 1e-7 * rng.standard_normal(shp)
        self.e = 1e-6 * rng.standard_normal(shp)

    def grid_args(self):
        return (self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, self.g,
                self.beta)


def oracle_sia_steps(spec, nsteps, dts=None, eos=None):
    """The SIA steps of a ScenarioSpec (transportse_amd.scenario) run with the
    oracle's reference functions in tse.py order (compute_P, compute_q,
    flow_error, compute_C per species, compute_T, linear EOS, density norm,
    reset on a rejected Picard iteration), each solve converged to an update
    norm of 1e-15 at the step's dt (dts[s], default the spec's dt), i.e. the
    FP64 fixed point of every solve.  eos: a fluid_properties callable
    (tse.py:178-229 contract, whole fields, species axis first) instead of
    the linear EOS; it may change rho_f, l_f, nue, c_f and cp_f."""
    from oracle import Oracle
    from transportse_amd.scenario import EOS, UNIFORM, host_fields
    hf = host_fields(spec)
    shp = (spec.nx, spec.ny, spec.nz)
    U = {k: np.full(shp, v) for k, v in UNIFORM.items()}
    disp_l, disp_t = np.full(shp, spec.disp[0]), np.full(shp, spec.disp[1])
    frozen = types.SimpleNamespace(c_r=U["c_r"], l_r=U["l_r"], cp_r=U["cp_r"],
                                   rho_r=U["rho_r"], lf_l=disp_l, lf_t=disp_t,
                                   lt_l=disp_l, lt_t=disp_t)
    d = np.full(spec.nx, spec.spacing)
    o = Oracle(spec.nx, spec.ny, spec.nz, d, np.full(spec.ny, spec.spacing),
               np.full(spec.nz, spec.spacing), 9.80665, spec.beta,
               frozen_source=frozen, dt_max=spec.dt, c_imblflo=spec.imblflo,
               t_imblflo=spec.imblflo)
    K, P, T = hf["K"], hf["P"].copy(), hf["T"].copy()
    C = [c.copy() for c in hf["C"]]
    rho, l_f = hf["rho_f"].copy(), hf["l_f"].copy()
    rho_prev = np.zeros(shp)
    qx, qy, qz = np.zeros(shp), np.zeros(shp), np.zeros(shp)
    for step in range(nsteps):
        dt = spec.dt if dts is None else dts[step]
        Pn0, Cn0, Tn0 = P.copy(), [c.copy() for c in C], T.copy()
        for it in range(spec.sia_maxit):
            r = o.compute_P(P, U["c_f"], U["nue"], rho, U["phi"], K, K, K, qx, qy,
                            qz, dt, hf["BP"], U["zero"], SOR=1.8, max_it=10 ** 6,
                            inc_it=1.25, max_dt=dt, P_conv=1e-15)
            assert r[4]
            P, Pn = r[2], r[3]
            o.compute_q(P, hf["BP"], K, K, K, qx, qy, qz, U["nue"], rho)
            e = o.flow_error(P, Pn, qx, qy, qz, U["zero"], U["phi"], U["c_f"],
                             U["c_r"], rho, dt)
            for s in range(spec.species):
                r = o.compute_C(C[s], U["D_f"], U["D_f"], U["D_f"], U["phi"], rho,
                                qx, qy, qz, dt, hf["BC"], U["zero"], U["R"], e,
                                omega=spec.omega, SOR=1.0, max_it=10 ** 5,
                                max_dt=dt, C_conv=1e-15)
                assert r[3]
                C[s] = r[2]
            r = o.compute_T(T, l_f, U["cp_f"], rho, U["phi"], qx, qy, qz, dt,
                            hf["BT"], U["zero"], e, omega=spec.omega, SOR=1.0,
                            max_it=10 ** 5, max_dt=dt, T_conv=1e-15)
            assert r[3]
            T = r[2]
            if eos is None:
                rho = EOS["rho0"] + EOS["drho_dc"] * C[0]
                l_f = EOS["lf0"] * (1.0 + EOS["dlf_dt"] * (T - EOS["t_ref"]))
            else:
                D = np.broadcast_to(U["D_f"], (spec.species,) + shp)
                out = eos(np.stack(C), T, P, U["nue"], rho, U["c_f"], U["cp_f"],
                          l_f, U["cp_r"], U["l_r"], U["phi"], K, K, K, D, D, D,
                          U["rho_r"])
                U["nue"], rho, U["c_f"], U["cp_f"], l_f = (
                    np.broadcast_to(np.asarray(out[n], np.float64), shp).copy()
                    for n in (3, 4, 5, 6, 7))
            diff = o.l2_norm(rho, rho_prev)
            if diff < spec.sia_conv or it == spec.sia_maxit - 1:
                break
            rho_prev = rho.copy()
            P, T, C = Pn0.copy(), Tn0.copy(), [c.copy() for c in Cn0]
    return {"P": P, "T": T, "C": C, "rho_f": rho, "l_f": l_f}


def check_plot_files(out_dir, case):
    """The 34-tuple files a driver wrote at PLOTTIME (tse.py:312-329) against
    the reference's own tuples in tests/golden/<case>.npz: the same number
    of files, names <case>-<index>.p, and every element equal (arrays bit
    for bit).  The files are this build's own output, read back with
    joblib."""
    from joblib import load
    g = np.load(f"{GOLDEN}/{case}.npz")
    count = int(g["plot_count"])
    files = sorted(f for f in os.listdir(out_dir) if f.endswith(".p"))
    assert files == [f"{case}-{i:05d}.p" for i in range(count)], files
    for n, fn in enumerate(files):
        tup = load(os.path.join(out_dir, fn))
        assert len(tup) == 34
        for e, v in enumerate(tup):
            ref = g[f"plot{n}_e{e:02d}"]
            got = np.asarray(v)
            assert got.shape == ref.shape and np.array_equal(got, ref), (fn, e)
    return count

