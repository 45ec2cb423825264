"""Batched performance-form solves over many grids (``tse_batch_*``).

A ``Batch`` holds members: ``FastPressure`` or ``FastTransport`` systems
(each on its own ``HipContext``) that are set up for a solve.  Every cycle
runs, for each active member, the same sweeps, residual pass and stop test
as the member's own ``solve`` would run per cycle, with the sweeps of all
members in one persistent tile-queue launch per sweep group.  A member's
result is therefore the same whether it is solved alone or in a batch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib
from .pressure import FORMS

KIND_P = 0
KIND_TR = 1


@dataclass
class MemberStatus:
    sweeps: int
    cycles: int
    converged: bool
    active: bool
    initial_rel: float
    rel_residual: float
    contraction: float = 0.0   # per sweep, last cycle after the first (0: none)


class Batch:
    """One batch of fast-form solves of one kind, form and launch plan."""

    def __init__(self, kind: int, form: str = "mixed", plan=(2, 16, 1, 0),
                 inner: int = 100, tol: float = 1e-10, sor: float = 1.996,
                 max_sweeps: int = 200000, device: int = 0, stream=None):
        import torch
        self.lib = _lib.load()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        h = ctypes.c_void_p()
        _lib.check(self.lib.tse_batch_create(ctypes.byref(h), device,
                                             ctypes.c_void_p(stream.cuda_stream)),
                   "tse_batch_create")
        self.h = h
        self.kind = kind
        self.members = {}
        self.configure(form, plan, inner, tol, sor, max_sweeps)

    def configure(self, form, plan, inner, tol, sor, max_sweeps):
        ns = plan[0]
        inner = max(ns, (int(inner) // ns) * ns)
        self.form, self.plan, self.inner = form, tuple(plan), inner
        self.max_sweeps = int(max_sweeps)
        prm = _lib.FastParams(form=FORMS[form], plan=_lib.SweepPlan(*plan),
                              inner=inner, max_sweeps=int(max_sweeps),
                              sor=float(sor), tol=float(tol))
        _lib.check(self.lib.tse_batch_begin(self.h, self.kind, ctypes.byref(prm)),
                   "tse_batch_begin")

    def add(self, system) -> int:
        """Add a set-up FastPressure / FastTransport; returns its slot."""
        rc = self.lib.tse_batch_add(self.h, system.hip.ctx)
        if rc < 0:
            _lib.check(rc, "tse_batch_add")
        self.members[rc] = system
        return rc

    def rearm(self, slot: int):
        _lib.check(self.lib.tse_batch_rearm(self.h, slot), "tse_batch_rearm")

    def rearm_many(self, slots):
        if not slots:
            return
        arr = (ctypes.c_int32 * len(slots))(*slots)
        _lib.check(self.lib.tse_batch_rearm_many(self.h, arr, len(slots)),
                   "tse_batch_rearm_many")

    def writeback(self, slot: int):
        _lib.check(self.lib.tse_batch_writeback(self.h, slot), "tse_batch_writeback")

    def set_sor(self, slot: int, sor: float):
        _lib.check(self.lib.tse_batch_set_sor(self.h, slot, float(sor)),
                   "tse_batch_set_sor")

    def remove(self, slot: int):
        _lib.check(self.lib.tse_batch_remove(self.h, slot), "tse_batch_remove")
        self.members.pop(slot, None)

    def cycle(self, graph: bool = False) -> int:
        n = ctypes.c_int32()
        _lib.check(self.lib.tse_batch_cycle(self.h, int(graph), ctypes.byref(n)),
                   "tse_batch_cycle")
        return n.value

    def run(self, graph: bool = False) -> int:
        n = ctypes.c_int32()
        _lib.check(self.lib.tse_batch_run(self.h, int(graph), ctypes.byref(n)),
                   "tse_batch_run")
        return n.value

    def status(self, slot: int) -> MemberStatus:
        r = _lib.FastResult()
        _lib.check(self.lib.tse_batch_status(self.h, slot, ctypes.byref(r)),
                   "tse_batch_status")
        return MemberStatus(r.sweeps, r.cycles, bool(r.converged), bool(r.pad),
                            r.initial_rel, r.rel_residual, r.contraction)

    def close(self):
        """Hand every member's iterate back to its context, then free the
        batch."""
        if self.h:
            for s in list(self.members):
                self.remove(s)
            self.lib.tse_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        # only the batch's own buffers: a finalizer may run after a member's
        # context was destroyed (garbage cycles are finalized in any order),
        # so no write-back into the members here
        try:
            if self.h:
                self.lib.tse_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
