"""Serving loop: scenario requests with continuous batching on one GPU.

Requests arrive as a Poisson process.  Each request is a scenario (grid,
permeability, species) that runs a number of SIA steps to a relative
residual in mixed precision.  The server

* admits waiting requests in arrival order while their HBM footprint fits
  the budget (``scenario.field_bytes``, plus ``scenario.uniform_bytes`` for
  the first live request of a grid shape, whose shared uniform fields stay
  until the last request of that shape finishes) -- fields are uploaded
  once and stay device-resident across the request's time steps;
* every tick runs one cycle (``inner`` sweeps + residual + stop test) of the
  pressure batch and one of the transport batch over all requests that are
  in a solve right now, whichever step or Picard iteration each is in;
* between ticks, moves every request whose solve finished on to its next
  solve (device work of the SIA step in between) and adds it to the batch
  of that solve's kind; requests that arrive meanwhile join at the next
  tick.

A request's arithmetic does not depend on what else is in the batches (each
member's sweeps, residuals and stop test are its own), so its results are
the same alone and in the mix.  Latency is arrival to completion (queueing
included); throughput is completed requests per second of wall time.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .batch import KIND_P, KIND_TR, Batch
from .scenario import (DeviceScenario, ScenarioSpec, field_bytes, resume_all,
                       uniform_bytes)

GRIDS = ((64, 64, 64), (128, 128, 128), (256, 256, 128))


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    spec: ScenarioSpec
    steps: int
    admitted: float = -1.0
    done: float = -1.0
    solves: int = 0
    sweeps: int = 0
    fields: dict | None = None   # request payload: initial host fields
    result: dict | None = None


def prepare(requests, cache: dict | None = None) -> None:
    """Build each request's input fields (the payload a client sends).  With
    `cache`, payloads are kept by scenario (grid, seeds, physics) and reused
    by requests of the same scenario; the server only reads them."""
    from .scenario import host_fields
    for r in requests:
        if r.fields is not None:
            continue
        if cache is None:
            r.fields = host_fields(r.spec)
            continue
        key = repr(r.spec)
        if key not in cache:
            cache[key] = host_fields(r.spec)
        r.fields = cache[key]


def make_requests(n: int, rate: float, steps: int = 10, species: int = 3,
                  arrival_seed: int = 5, grid_seed: int = 6,
                  grids=GRIDS, dt: float = 3600.0) -> list[Request]:
    """n requests, Poisson arrivals at `rate` per second (seed 5), grids drawn
    from `grids` (seed 6), sigma_lnK ~ U[0.5, 2.5], heat + `species`, time
    step dt (1 h keeps the transport update a contraction through the
    initial density-driven transient, as for the ensemble)."""
    ra = np.random.default_rng(arrival_seed)
    rg = np.random.default_rng(grid_seed)
    t = np.cumsum(ra.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        g = grids[int(rg.integers(0, len(grids)))]
        spec = ScenarioSpec(*g, spacing=5.0, beta=0.0,
                            sigma=float(rg.uniform(0.5, 2.5)), corr=5.0,
                            seed=int(rg.integers(0, 2 ** 31 - 1)),
                            species=species, heat=True, salt_patch=4, dt=dt)
        out.append(Request(i, float(t[i]), spec, steps))
    return out


def admissible(need: int, used: int, budget: float, n_live: int) -> bool:
    """Admission rule: a request is admitted when its footprint fits the
    remaining budget, or when nothing is running (a request larger than the
    whole budget still runs, alone)."""
    return used + need <= budget or n_live == 0


@dataclass
class ServeStats:
    rate: float
    completed: int
    seconds: float
    p50: float
    p99: float
    mean: float
    solves_per_s: float       # completed requests per second
    peak_bytes: int
    max_concurrent: int
    latencies: list = field(default_factory=list)
    # latency = queueing (arrival -> admission) + service (admission -> done)
    queue_p50: float = float("nan")
    queue_p99: float = float("nan")
    service_p50: float = float("nan")
    service_p99: float = float("nan")
    waited: int = 0           # requests that waited for admission (> 1 ms)


class Server:
    def __init__(self, budget_bytes=200e9, form="mixed", tol=1e-10,
                 plan_p=None, plan_tr=(2, 16, 1, 64, 2), inner_p=51,
                 inner_tr=6, sor_p=1.95, sor_tr=1.0, device=0,
                 max_sweeps=400000, keep_results=False):
        """plan_p None: the cell-pair sweep on the bfloat16 inner operator
        (variant 5)."""
        import torch
        if plan_p is None:
            plan_p = (3, 32, 2, 64, 5)
        self.budget = float(budget_bytes)
        self.device = device
        self.stream = torch.cuda.Stream(device)   # default priority
        self.keep = keep_results
        mk = lambda kind, plan, inner, sor: Batch(
            kind, form, plan, inner=inner, tol=tol, sor=sor,
            max_sweeps=max_sweeps, device=device, stream=self.stream)
        self.batches = {KIND_P: mk(KIND_P, plan_p, inner_p, sor_p),
                        KIND_TR: mk(KIND_TR, plan_tr, inner_tr, sor_tr)}

    def run(self, requests: list[Request], time_scale: float = 1.0) -> ServeStats:
        """Serve `requests` (arrival times in seconds, scaled by
        time_scale); returns latency / throughput statistics."""
        import torch
        reqs = sorted(requests, key=lambda r: r.arrival)
        waiting: deque = deque()
        live = {}            # rid -> [request, scenario, generator, kind, slots, need]
        shapes = {}          # grid shape -> live requests of that shape
        used = 0
        peak = 0
        max_conc = 0
        nxt = 0
        done = []
        with torch.cuda.stream(self.stream):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            now = lambda: time.perf_counter() - t0
            while nxt < len(reqs) or waiting or live:
                t = now()
                while nxt < len(reqs) and reqs[nxt].arrival * time_scale <= t:
                    waiting.append(reqs[nxt])
                    nxt += 1
                # admission by HBM footprint, in arrival order
                while waiting:
                    r = waiting[0]
                    shp = (r.spec.nx, r.spec.ny, r.spec.nz)
                    # the first live request of a shape also brings the
                    # shape's shared uniform fields
                    need = field_bytes(r.spec) + (0 if shapes.get(shp) else
                                                  uniform_bytes(r.spec))
                    if not admissible(need, used, self.budget, len(live)):
                        break
                    waiting.popleft()
                    shapes[shp] = shapes.get(shp, 0) + 1
                    used += need
                    peak = max(peak, used)
                    r.admitted = now()
                    scn = DeviceScenario(r.spec, self.device, fields=r.fields)
                    r.fields = None
                    gen = scn.steps(r.steps)
                    kind, sysm = resume_all({0: gen}, [0])[0]
                    slots = {KIND_P: None, KIND_TR: None}
                    self._enter(scn, kind, sysm, slots)
                    live[r.rid] = [r, scn, gen, kind, slots, need]
                max_conc = max(max_conc, len(live))
                if not live:
                    # idle until the next arrival
                    if nxt < len(reqs):
                        dt = reqs[nxt].arrival * time_scale - now()
                        if dt > 0:
                            time.sleep(min(dt, 0.05))
                    continue
                # one cycle of each batch over every request in a solve
                for b in self.batches.values():
                    b.cycle(graph=False)
                # advance requests whose solve finished; their re-arms go to
                # each batch at once after the pass
                rearm = {KIND_P: [], KIND_TR: []}
                fin = []
                for rid in list(live):
                    r, scn, gen, kind, slots, need = live[rid]
                    b = self.batches[kind]
                    st = b.status(slots[kind])
                    if st.active:
                        continue
                    # a solve that did not converge halves the request's dt
                    # and comes back as the same stage
                    scn.solved(st.converged, st.sweeps, b.max_sweeps)
                    r.sweeps += st.sweeps
                    fin.append(rid)
                    if not st.converged:
                        continue
                    b.writeback(slots[kind])
                    if kind == KIND_P:
                        scn.omega_p.report(st.sweeps, st.contraction)
                    r.solves += 1
                nxt_items = resume_all({rid: live[rid][2] for rid in fin}, fin)
                for rid in fin:
                    r, scn, gen, kind, slots, need = live[rid]
                    it = nxt_items[rid]
                    if it is None:
                        for k, sl in slots.items():
                            if sl is not None:
                                self.batches[k].remove(sl)
                        torch.cuda.synchronize()
                        r.done = now()
                        if self.keep:
                            r.result = scn.download()
                        scn.close()
                        used -= field_bytes(r.spec)
                        shp = (r.spec.nx, r.spec.ny, r.spec.nz)
                        shapes[shp] -= 1
                        if shapes[shp] == 0:     # its shared fields go too
                            del shapes[shp]
                            used -= uniform_bytes(r.spec)
                        del live[rid]
                        done.append(r)
                        continue
                    kind, sysm = it
                    self._enter(scn, kind, sysm, slots, rearm)
                    live[rid][3] = kind
                for k, lst in rearm.items():
                    self.batches[k].rearm_many(lst)
            sec = now()
        lat = np.array([r.done - r.arrival * time_scale for r in done])
        wait = np.array([r.admitted - r.arrival * time_scale for r in done])
        serv = np.array([r.done - r.admitted for r in done])
        pct = lambda a, q: float(np.percentile(a, q)) if len(a) else float("nan")
        return ServeStats(
            rate=0.0, completed=len(done), seconds=sec,
            p50=pct(lat, 50), p99=pct(lat, 99),
            mean=float(lat.mean()) if len(lat) else float("nan"),
            solves_per_s=len(done) / sec if sec > 0 else 0.0, peak_bytes=peak,
            max_concurrent=max_conc, latencies=lat.tolist(),
            queue_p50=pct(wait, 50), queue_p99=pct(wait, 99),
            service_p50=pct(serv, 50), service_p99=pct(serv, 99),
            waited=int(np.sum(wait > 1e-3)))

    def _enter(self, scn, kind, sysm, slots, rearm=None):
        """Put a scenario's set-up system into the batch of its kind (a slot
        per kind for the request's lifetime; re-armed for later solves, at
        once if `rearm` is None, else queued in rearm[kind])."""
        b = self.batches[kind]
        if slots[kind] is None:
            slots[kind] = b.add(sysm)
        elif rearm is None:
            b.rearm(slots[kind])
        else:
            rearm[kind].append(slots[kind])
        if kind == KIND_P:
            b.set_sor(slots[kind], scn.omega_p.omega)

    def close(self):
        for b in self.batches.values():
            b.close()
