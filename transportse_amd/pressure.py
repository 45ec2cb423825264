"""Performance forms of the pressure solve (plain FP64, mixed FP32-in-FP64).

``FastPressure`` drives the C-ABI ``tse_fast_p_*`` entry points for one grid:
it uploads compute_P's inputs (reference layout, as tse.py passes them),
assembles the volume-scaled symmetric operator, solves to a relative
residual ||b - A p|| / ||b|| with fused red/black sweeps, and returns P.

Forms:
  * ``"plain64"`` -- FP64 sweeps on P.
  * ``"mixed"``   -- FP32 sweeps on the correction A e = r, with r = b - A p,
    the update p += e and the convergence norm in FP64 (defect correction).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import HipContext

FORMS = {"plain64": _lib.FORM_PLAIN64, "mixed": _lib.FORM_MIXED}
P_INPUTS = ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz", "BP", "WP", "c_r")


@dataclass
class SolveStats:
    form: str
    sweeps: int
    cycles: int
    converged: bool
    initial_rel: float
    rel_residual: float
    seconds: float
    plan: tuple
    contraction: float = 0.0   # per sweep, last cycle after the first (0: none)
    sor: float = 0.0           # over-relaxation factor of the last cycle


class FastSystem:
    """Shared driver of one performance-form system (``tse_fast_<kind>_*``)."""

    KIND = ""
    DEFAULT_PLAN = (2, 16, 1, 64)

    def __init__(self, nx, ny, nz, dx, dy, dz, g, beta_deg, device=0,
                 hip: HipContext | None = None):
        self.hip = hip or HipContext(nx, ny, nz, dx, dy, dz, g, beta_deg, device)
        self.lib = self.hip.lib
        self.N = self.hip.nx * self.hip.ny * self.hip.nz
        self.d = {}
        self.plans = {}

    def _fn(self, what):
        return getattr(self.lib, f"tse_fast_{self.KIND}_{what}")

    def assemble(self):  # pragma: no cover - per system
        raise NotImplementedError

    def plan(self, form: str, tune: bool = True):
        """Launch plan for the form's sweep precision (autotuned once)."""
        if form not in self.plans:
            if tune:
                from .autotune import tune as _tune
                if form == "mixed":
                    # the FP32 timing path needs a residual in place
                    self.solve("mixed", max_sweeps=0, plan=(1, 16, 1, 0))
                plan, _ = _tune(self.hip, form == "mixed", kind=self.KIND)
                self.plans[form] = tuple(plan)
                self.assemble()  # restore the iterate the tuning advanced
            else:
                self.plans[form] = self.DEFAULT_PLAN
        return self.plans[form]

    def solve(self, form="mixed", tol=1e-10, sor=1.996, max_sweeps=200000,
              inner=None, plan=None, adaptive=False, adapt_sor=False) -> SolveStats:
        """Solve to relative residual `tol`.  adaptive: cycle lengths from
        the observed convergence rate (first cycle `inner`, then up to
        8*inner), fewer FP64 residual passes; else every cycle is `inner`.
        adapt_sor: `sor` is the starting factor; a cycle whose contraction
        shows it below the optimum moves it to Young's estimate (at most
        twice; ``SolveStats.sor`` is the last cycle's)."""
        pl = plan or self.plan(form)
        ns = pl[0]
        inner = inner or 100
        inner = max(ns, (inner // ns) * ns)
        prm = _lib.FastParams(form=FORMS[form], plan=_lib.SweepPlan(*pl),
                              inner=inner, max_sweeps=int(max_sweeps),
                              cycle_mode=_lib.CYCLE_ADAPTIVE if adaptive
                              else _lib.CYCLE_FIXED, adapt_sor=int(bool(adapt_sor)),
                              sor=float(sor), tol=float(tol))
        res = _lib.FastResult()
        t0 = time.perf_counter()
        _lib.check(self._fn("solve")(self.hip.ctx, ctypes.byref(prm),
                                     ctypes.byref(res)), f"tse_fast_{self.KIND}_solve")
        sec = time.perf_counter() - t0
        return SolveStats(form, res.sweeps, res.cycles, bool(res.converged),
                          res.initial_rel, res.rel_residual, sec, tuple(pl),
                          res.contraction, res.sor)

    def result_device(self):
        """Write the current iterate into a split-layout device field (kept
        apart from the input field, which ``assemble`` reloads)."""
        if "res" not in self.d:
            self.d["res"] = self.hip.empty()
        _lib.check(self._fn("result")(self.hip.ctx, self.d["res"].data_ptr()),
                   f"tse_fast_{self.KIND}_result")
        return self.d["res"]

    def result(self, out: np.ndarray | None = None) -> np.ndarray:
        return self.hip.download(self.result_device(), out=out)

    def residual(self) -> float:
        rel = ctypes.c_double()
        _lib.check(self._fn("residual")(self.hip.ctx, ctypes.byref(rel)),
                   f"tse_fast_{self.KIND}_residual")
        return rel.value

    def sweeps(self, n, plan, sor=1.996, fp32=True):
        """Timing hook: n fused sweeps without residual checks."""
        _lib.check(self._fn("sweeps")(self.hip.ctx,
                                      ctypes.byref(_lib.SweepPlan(*plan)),
                                      int(n), float(sor), int(fp32)),
                   f"tse_fast_{self.KIND}_sweeps")


class FastPressure(FastSystem):
    """Fast-form pressure solver bound to one grid on one device."""

    KIND = "p"

    def load(self, P, dt, Pn=None, **fields):
        """Upload P, Pn and the ten compute_P inputs and assemble for dt."""
        for n in P_INPUTS:
            self.d[n] = self.hip.upload(fields[n], out=self.d.get(n))
        self.d["P"] = self.hip.upload(P, out=self.d.get("P"))
        self.d["Pn"] = self.hip.upload(P if Pn is None else Pn, out=self.d.get("Pn"))
        self.dt = float(dt)
        self.assemble()

    def assemble(self):
        """(Re)assemble the operator and load P as the iterate."""
        self._pf = _lib.PFields(**{n: self.d[n].data_ptr() for n in _lib.P_FIELDS})
        _lib.check(self.lib.tse_fast_p_setup(self.hip.ctx, ctypes.byref(self._pf),
                                             self.dt), "tse_fast_p_setup")
