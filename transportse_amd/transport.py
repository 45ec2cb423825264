"""Performance forms of the transport solves (compute_C, compute_T).

``FastTransport`` drives ``tse_fast_c_setup`` / ``tse_fast_t_setup`` and the
``tse_fast_tr_*`` entry points.  The coefficients are the reference's own
(evaluated by the same kernels as the bit-faithful solvers, with the
reference arithmetic); the system solved is the fixed point of the update of
tse.py:870-875 (compute_C) and tse.py:1038-1043 (compute_T):

    (R - omega M) x - omega sum_f a_f x_f = X + R Xn + W - E

with Xn the unknown at entry.  It is non-symmetric (upwind advection), so
each cell keeps its own six neighbour coefficients.  The solve stops on the
relative residual ||b - A x|| / ||b|| instead of the reference's update
norm, in plain FP64 or with FP32 sweeps inside FP64 defect correction.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .pressure import FastSystem

C_INPUTS = ("D_fx", "D_fy", "D_fz", "phi", "rho_f", "qx", "qy", "qz", "BC",
            "WC", "RC", "e", "lf_l", "lf_t")
T_INPUTS = ("l_f", "cp_f", "rho_f", "phi", "qx", "qy", "qz", "BT", "WT", "e",
            "l_r", "cp_r", "rho_r", "lt_l", "lt_t")


class FastTransport(FastSystem):
    """Fast-form solver for one species (compute_C) or heat (compute_T)."""

    KIND = "tr"
    DEFAULT_PLAN = (2, 16, 1, 64)

    def _upload(self, names, fields):
        for n in names:
            self.d[n] = self.hip.upload(fields[n], out=self.d.get(n))

    def load_c(self, C, D_fx, D_fy, D_fz, phi, rho_f, qx, qy, qz, dt, BC, WC,
               R, e, omega=1.0, *, lf_l, lf_t, imblflo=False):
        """compute_C's arguments (tse.py:783-785) plus its module-level
        dispersion lengths lf_l, lf_t and the C_IMBLFLO switch."""
        self._upload(C_INPUTS, dict(D_fx=D_fx, D_fy=D_fy, D_fz=D_fz, phi=phi,
                                    rho_f=rho_f, qx=qx, qy=qy, qz=qz, BC=BC,
                                    WC=WC, RC=R, e=e, lf_l=lf_l, lf_t=lf_t))
        self.d["X"] = self.hip.upload(C, out=self.d.get("X"))
        if "Xn" not in self.d:
            self.d["Xn"] = self.hip.empty()
        self.which, self.dt, self.omega = "C", float(dt), float(omega)
        self.imblflo = bool(imblflo)
        self.assemble()

    def load_t(self, T, l_f, cp_f, rho_f, phi, qx, qy, qz, dt, BT, WT, e,
               omega=1.0, *, l_r, cp_r, rho_r, lt_l, lt_t, imblflo=False):
        """compute_T's arguments (tse.py:919-921) plus its module-level rock
        properties l_r, cp_r, rho_r, dispersion lengths lt_l, lt_t and the
        T_IMBLFLO switch."""
        self._upload(T_INPUTS, dict(l_f=l_f, cp_f=cp_f, rho_f=rho_f, phi=phi,
                                    qx=qx, qy=qy, qz=qz, BT=BT, WT=WT, e=e,
                                    l_r=l_r, cp_r=cp_r, rho_r=rho_r,
                                    lt_l=lt_l, lt_t=lt_t))
        self.d["X"] = self.hip.upload(T, out=self.d.get("X"))
        if "Xn" not in self.d:
            self.d["Xn"] = self.hip.empty()
        self.which, self.dt, self.omega = "T", float(dt), float(omega)
        self.imblflo = bool(imblflo)
        self.assemble()

    def assemble(self):
        """(Re)assemble from the loaded fields and load the unknown."""
        ptr = {n: v.data_ptr() for n, v in self.d.items()}
        if self.which == "C":
            f = _lib.CFields(C=ptr["X"], Cn=ptr["Xn"], imblflo=int(self.imblflo),
                             **{n: ptr[n] for n in C_INPUTS})
            rc = self.lib.tse_fast_c_setup(self.hip.ctx, ctypes.byref(f),
                                           self.dt, self.omega)
            _lib.check(rc, "tse_fast_c_setup")
        else:
            f = _lib.TFields(T=ptr["X"], Tn=ptr["Xn"], imblflo=int(self.imblflo),
                             **{n: ptr[n] for n in T_INPUTS})
            rc = self.lib.tse_fast_t_setup(self.hip.ctx, ctypes.byref(f),
                                           self.dt, self.omega)
            _lib.check(rc, "tse_fast_t_setup")
        self._f = f
