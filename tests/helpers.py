"""Shared test helpers: seeded field sets and backend factories."""
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ("g1_salt", "g2_sia_heat", "g3_halving", "g4_smolar", "g5_nlte",
         "g6_xz", "g7_column", "g8_xy")


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
