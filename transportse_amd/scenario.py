"""Device-resident scenarios: the SIA time step on the performance forms.

A scenario is one TRANSPORTSE run (grid, permeability realization, boundary
values) whose fields stay in HBM across time steps.  ``DeviceScenario.steps``
is the SIA step of tse.py:1181-1494 as a generator: between solves it runs
the device pieces (compute_q, flow_error, the linear EOS, the SIA density
norm, the resets of P/C/T on a rejected Picard iteration, mass sums) and at
each solve it yields ``(kind, system)`` -- ``kind`` 0 for the pressure
system, 1 for a transport system -- to whoever drives it: a plain loop that
calls ``system.solve``, a lockstep ensemble that batches the same stage of
many scenarios, or a serving loop that batches whatever stage each request
is in.

Solver form: the pressure and transport systems are solved to a relative
residual (FP32 sweeps in FP64 defect correction, or plain FP64) instead of
the reference's update norm.  The time-step control is the reference's: the
driver reports each solve's outcome (``solved``); a solve that does not
converge within its sweep budget halves dt and runs again from its entry
state, and the rest of the step uses the halved dt (tse.py:730-740, the
returned t_X feeding the next call); the next step's dt is the smallest of
the solvers' next dt, each grown by its inc_it (1.25 pressure, 1.05
transport) when its solve took fewer than half its sweeps and capped at the
scenario's dt (tse.py:724-729, 1486-1492).  When every solve converges dt
stays at the scenario's dt, as in the reference with SIA on
(tse.py:1262-1271).  Everything else follows tse.py: the solver call order,
the time level of each input, rhof_prev starting at zero and persisting
across steps (tse.py:1179, 1430), accept on SIA_conv or the last iteration
(tse.py:1429).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import HipContext
from .omega import OmegaTuner
from .pressure import FastPressure
from .transport import FastTransport
from .workloads import hydrostatic_P, lognormal_K

EOS = dict(rho0=1000.0, drho_dc=700.0, lf0=2.5, dlf_dt=0.002, t_ref=10.0)

# DeviceScenario.steps() yields (KIND_WAIT, None) where it next reads a
# device reduction; the solve kinds 0 (pressure) and 1 (transport) are
# batch.KIND_P and batch.KIND_TR
KIND_WAIT = 2


def resume_all(gens: dict, ids) -> dict:
    """next() on the step generators gens[i] for i in ids, resuming those
    that stop at a reduction read (KIND_WAIT) until each yields a solve or
    ends.  Every member's reduction is enqueued before the first read waits,
    so the members share one wait.  Returns {i: (kind, system) or None}."""
    out = {}
    todo = list(ids)
    while todo:
        again = []
        for i in todo:
            try:
                item = next(gens[i])
            except StopIteration:
                out[i] = None
                continue
            if item[0] == KIND_WAIT:
                again.append(i)
            else:
                out[i] = item
        todo = again
    return out


@dataclass
class ScenarioSpec:
    """Inputs of one scenario (sizes in cells, lengths in m, dt in s)."""
    nx: int
    ny: int
    nz: int
    spacing: float = 5.0
    beta: float = 0.0
    sigma: float = 1.0
    corr: float = 5.0
    seed: int = 1
    species: int = 1          # species 0 is salt (linear EOS), others tracers
    heat: bool = True
    salt_patch: int = 4       # top layers with salt C = 1 over the central x half
    c_top: float | None = None  # salt Dirichlet value on the top face
    t0: float = 10.0
    t_bottom: float = 60.0    # Dirichlet temperature of the bottom face
    disp: tuple = (1.0, 0.1)  # longitudinal / transversal dispersivity
    dt: float = 86400.0
    sia_maxit: int = 2
    sia_conv: float = 1e-6
    omega: float = 1.0        # OMEGA_C = OMEGA_T
    imblflo: bool = True

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def host_fields(spec: ScenarioSpec) -> dict:
    """Initial state of the variable fields (reference layout, float64)."""
    shp = (spec.nx, spec.ny, spec.nz)
    nx, nz = spec.nx, spec.nz
    dx = np.full(spec.nx, float(spec.spacing))
    dz = np.full(spec.nz, float(spec.spacing))
    f = {"K": lognormal_K(shp, spec.sigma, spec.corr, spec.seed)}
    f["P"] = hydrostatic_P(dx, dz, spec.beta, ny=spec.ny)
    f["BP"] = np.zeros(shp)
    f["BP"][0] = 1.0
    f["BP"][-1] = 1.0
    f["BC"] = np.zeros(shp)
    C = [np.zeros(shp) for _ in range(spec.species)]
    if spec.salt_patch:
        C[0][nx // 4: nx - nx // 4, :, nz - spec.salt_patch:] = 1.0
    if spec.c_top is not None:
        C[0][:, :, nz - 1] = spec.c_top
        f["BC"][:, :, nz - 1] = 1.0
    for s in range(1, spec.species):  # tracer blobs, one per species
        cx = (s * nx) // (spec.species + 1)
        w = max(1, min(nx, spec.ny, nz) // 8)
        C[s][max(0, cx - w): cx + w, spec.ny // 2 - w: spec.ny // 2 + w,
             nz // 2 - w: nz // 2 + w] = 1.0
    f["C"] = C
    T = np.full(shp, spec.t0)
    f["BT"] = np.zeros(shp)
    if spec.heat:
        T[:, :, 0] = spec.t_bottom
        f["BT"][:, :, 0] = 1.0
    f["T"] = T
    f["rho_f"] = EOS["rho0"] + EOS["drho_dc"] * C[0]
    f["l_f"] = EOS["lf0"] * (1.0 + EOS["dlf_dt"] * (T - EOS["t_ref"]))
    return f


# uniform fields, shared by every open scenario of the same grid shape on a
# device: key -> [tensor, number of scenarios holding it]; the last scenario
# of a shape to close drops it (shared_bytes counts what is held)
UNIFORM = dict(phi=0.2, c_f=4.5e-10, c_r=1e-10, nue=1e-3, cp_f=4150.0,
               l_r=2.0, cp_r=1000.0, rho_r=2500.0, R=1.0, D_f=1e-9, zero=0.0)
_shared: dict = {}


def _acquire_uniform(hip: HipContext, name: str, value: float):
    """(key, tensor) of a shared uniform field; the caller holds one
    reference until it calls _release_uniform(key)."""
    key = (hip.device_index, hip.shape, name, float(value))
    ent = _shared.get(key)
    if ent is None:
        ent = _shared[key] = [hip.upload(np.full(hip.shape, float(value))), 0]
    ent[1] += 1
    return key, ent[0]


def _release_uniform(key) -> None:
    ent = _shared.get(key)
    if ent is None:
        return
    ent[1] -= 1
    if ent[1] <= 0:
        del _shared[key]


def shared_bytes() -> int:
    """Device bytes held by the shared uniform fields."""
    return sum(ent[0].numel() * ent[0].element_size() for ent in _shared.values())


def uniform_bytes(spec: "ScenarioSpec") -> int:
    """Upper bound of the shared uniform fields a scenario of this grid
    adds when it is the first of its shape (UNIFORM plus the two
    dispersivities)."""
    return (len(UNIFORM) + 2) * 16 * ((spec.nx + 1) // 2) * spec.ny * spec.nz


def field_bytes(spec: ScenarioSpec) -> int:
    """HBM a scenario holds: its own split fields plus both fast systems and
    the coefficient records (uniform fields are shared and not counted)."""
    n = spec.cells
    S = spec.species
    own = 16 + 3 * S                      # split-layout FP64 fields
    fast_p = 8 * 8 + 9 * 4                # bytes per cell
    fast_tr = 11 * 8 + 10 * 4
    rec = 13 * 8 + 1
    return int(n * (own * 8 + fast_p + fast_tr + rec) * 1.02) + (1 << 20)


@dataclass
class StepStats:
    sia_iters: int = 0
    dt: float = 0.0
    sweeps: dict = field(default_factory=dict)   # stage -> sweeps
    mass: tuple = ()


class DeviceScenario:
    """One scenario's fields in HBM and its SIA step as a generator."""

    def __init__(self, spec: ScenarioSpec, device: int = 0, fields=None,
                 use_graphs: bool = False, fluid_properties=None):
        """fluid_properties: None for the closed-form linear EOS on the
        device, or a callable with the contract of a config's
        fluid_properties (tse.py:178-229), called on whole fields as the SIA
        driver calls it (species axis first for C and D_f*): it may change
        rho_f, l_f, nue, c_f and cp_f; every other argument must come back
        unchanged.  It runs on the host: the Picard update then downloads
        C, T and P and uploads what changed."""
        import torch
        self.spec = spec
        self.fluid_properties = fluid_properties
        # device work between solves as hipGraphs: each segment runs eagerly
        # once (allocations happen there), is captured on its second use and
        # replayed from then on (the fields never move)
        self.use_graphs = use_graphs
        self._seen = {}
        self._graphs = {}
        d = np.full(spec.nx, float(spec.spacing))
        self.hip = HipContext(spec.nx, spec.ny, spec.nz, d,
                              np.full(spec.ny, float(spec.spacing)),
                              np.full(spec.nz, float(spec.spacing)), 9.80665,
                              spec.beta, device)
        hip = self.hip
        hf = fields if fields is not None else host_fields(spec)
        self._ukeys = {}
        u = {k: self._uniform(k, v) for k, v in UNIFORM.items()}
        if fluid_properties is not None:
            # fields a host EOS may change: this scenario's own copies, made
            # now so that their addresses never change (captured graphs)
            for k in ("nue", "c_f", "cp_f"):
                u[k] = hip.upload(np.full((spec.nx, spec.ny, spec.nz), UNIFORM[k]))
        self.u = u
        self.K = hip.upload(hf["K"])
        self.P = hip.upload(hf["P"])
        self.BP = hip.upload(hf["BP"])
        self.BC = hip.upload(hf["BC"])
        self.BT = hip.upload(hf["BT"])
        self.C = [hip.upload(c) for c in hf["C"]]
        self.T = hip.upload(hf["T"])
        self.rho = hip.upload(hf["rho_f"])
        self.l_f = hip.upload(hf["l_f"])
        self.Pn = hip.empty()
        self.Cn = [hip.empty() for _ in self.C]
        self.Tn = hip.empty()
        self.rho_prev = hip.zeros()                   # tse.py:1179
        self.Xn = hip.empty()                          # transport entry copy
        self.qx, self.qy, self.qz = hip.zeros(), hip.zeros(), hip.zeros()
        self.e = hip.zeros()
        self.fp = FastPressure(spec.nx, spec.ny, spec.nz, None, None, None,
                               None, None, hip=hip)
        self.ft = FastTransport(spec.nx, spec.ny, spec.nz, None, None, None,
                                None, None, hip=hip)
        # the autotuner restores the system it timed through `assemble`
        self.fp.assemble = self.setup_p
        self.omega_p = OmegaTuner(max(spec.nx, spec.ny, spec.nz))
        self.t = 0.0
        # time-step control (tse.py:724-740, 896-914, 1065-1079, 1486-1492):
        # dt is the running time step of the current step (a solve that does
        # not converge halves it and is repeated from its entry state), dt_max
        # caps its growth, dt_step is the dt of the last completed step
        self.dt = float(spec.dt)
        self.dt_max = float(spec.dt)
        self.dt_step = float(spec.dt)
        self.min_dt = 1e-8
        self.halvings = 0
        self._next_dt = {}
        self._outcome = None
        self.n_steps = 0
        self.history: list[StepStats] = []
        self.eos = _lib.EosParams(**EOS)
        self.lib = hip.lib
        self.stream = torch.cuda.current_stream(hip.device)

    # device pieces -----------------------------------------------------------
    def _p(self, t):
        return t.data_ptr()

    def setup_p(self):
        u = self.u
        f = _lib.PFields(P=self._p(self.P), Pn=None, c_f=self._p(u["c_f"]),
                         nue=self._p(u["nue"]), rho_f=self._p(self.rho),
                         phi=self._p(u["phi"]), Kx=self._p(self.K),
                         Ky=self._p(self.K), Kz=self._p(self.K),
                         BP=self._p(self.BP), WP=self._p(u["zero"]),
                         c_r=self._p(u["c_r"]))
        _lib.check(self.lib.tse_fast_p_setup(self.hip.ctx, ctypes.byref(f),
                                             self.dt), "tse_fast_p_setup")

    def finish_p(self):
        _lib.check(self.lib.tse_fast_p_result(self.hip.ctx, self._p(self.P)),
                   "tse_fast_p_result")

    def flow(self):
        """compute_q and flow_error on the new P (tse.py:1228-1237); the
        storage reference Pn is P at the start of the step."""
        u, L, p = self.u, self.lib, self._p
        _lib.check(L.tse_compute_q(self.hip.ctx, p(self.P), p(self.BP), p(self.K),
                                   p(self.K), p(self.K), p(self.qx), p(self.qy),
                                   p(self.qz), p(u["nue"]), p(self.rho)),
                   "tse_compute_q")
        _lib.check(L.tse_flow_error(self.hip.ctx, p(self.P), p(self.Pn), p(self.qx),
                                    p(self.qy), p(self.qz), p(u["zero"]),
                                    p(u["phi"]), p(u["c_f"]), p(u["c_r"]),
                                    p(self.rho), self.dt, p(self.e)),
                   "tse_flow_error")

    def setup_c(self, s):
        u, p, sp = self.u, self._p, self.spec
        f = _lib.CFields(C=p(self.C[s]), Cn=p(self.Xn), D_fx=p(u["D_f"]),
                         D_fy=p(u["D_f"]), D_fz=p(u["D_f"]), phi=p(u["phi"]),
                         rho_f=p(self.rho), qx=p(self.qx), qy=p(self.qy),
                         qz=p(self.qz), BC=p(self.BC), WC=p(u["zero"]),
                         RC=p(u["R"]), e=p(self.e), lf_l=p(self._disp(0)),
                         lf_t=p(self._disp(1)), imblflo=int(sp.imblflo))
        _lib.check(self.lib.tse_fast_c_setup(self.hip.ctx, ctypes.byref(f), self.dt,
                                             sp.omega), "tse_fast_c_setup")

    def setup_t(self):
        u, p, sp = self.u, self._p, self.spec
        f = _lib.TFields(T=p(self.T), Tn=p(self.Xn), l_f=p(self.l_f),
                         cp_f=p(u["cp_f"]), rho_f=p(self.rho), phi=p(u["phi"]),
                         qx=p(self.qx), qy=p(self.qy), qz=p(self.qz),
                         BT=p(self.BT), WT=p(u["zero"]), e=p(self.e),
                         l_r=p(u["l_r"]), cp_r=p(u["cp_r"]), rho_r=p(u["rho_r"]),
                         lt_l=p(self._disp(0)), lt_t=p(self._disp(1)),
                         imblflo=int(sp.imblflo))
        _lib.check(self.lib.tse_fast_t_setup(self.hip.ctx, ctypes.byref(f), self.dt,
                                             sp.omega), "tse_fast_t_setup")

    def _uniform(self, name, value):
        """A shared uniform field, referenced once per scenario."""
        k = (name, float(value))
        if k not in self._ukeys:
            self._ukeys[k] = _acquire_uniform(self.hip, name, value)
        return self._ukeys[k][1]

    def _disp(self, n):
        return self._uniform(f"disp{n}", self.spec.disp[n])

    def finish_tr(self, dst):
        _lib.check(self.lib.tse_fast_tr_result(self.hip.ctx, self._p(dst)),
                   "tse_fast_tr_result")

    def eos_update(self):
        if self.fluid_properties is not None:
            return self._host_eos()
        _lib.check(self.lib.tse_linear_eos(self.hip.ctx, self._p(self.C[0]),
                                           self._p(self.T), self._p(self.rho),
                                           self._p(self.l_f), ctypes.byref(self.eos)),
                   "tse_linear_eos")

    # tse.py:178-229 argument order; the fluid properties a host EOS may set
    _FP_ARGS = ("C", "T", "P", "nue", "rho_f", "c_f", "cp_f", "l_f", "cp_r",
                "l_r", "phi", "Kx", "Ky", "Kz", "D_fx", "D_fy", "D_fz", "rho_r")
    _FP_SET = ("nue", "rho_f", "c_f", "cp_f", "l_f")

    def _host_eos(self):
        """The Picard density update through a host fluid_properties: the
        current C, T, P and fluid properties go to the callback as whole
        fields, what it changes comes back to the device."""
        sp, h = self.spec, self.hip
        shp = (sp.nx, sp.ny, sp.nz)
        if not hasattr(self, "_fp_host"):
            const = lambda v: np.broadcast_to(np.float64(v), shp)
            self._fp_host = {k: h.download(t) for k, t in (
                ("nue", self.u["nue"]), ("c_f", self.u["c_f"]),
                ("cp_f", self.u["cp_f"]))}
            self._fp_host.update(cp_r=const(UNIFORM["cp_r"]), l_r=const(UNIFORM["l_r"]),
                                 phi=const(UNIFORM["phi"]),
                                 rho_r=const(UNIFORM["rho_r"]))
            K = h.download(self.K)
            K.flags.writeable = False
            D = np.broadcast_to(np.float64(UNIFORM["D_f"]), (sp.species,) + shp)
            self._fp_host.update(Kx=K, Ky=K, Kz=K, D_fx=D, D_fy=D, D_fz=D)
        a = dict(self._fp_host)
        a["C"] = np.stack([h.download(c) for c in self.C])
        a["T"] = h.download(self.T)
        a["P"] = h.download(self.P)
        a["rho_f"] = h.download(self.rho)
        a["l_f"] = h.download(self.l_f)
        args = [a[k] for k in self._FP_ARGS]
        for v in args:
            if v.flags.writeable:
                v.flags.writeable = False   # changes come back as new arrays
        out = self.fluid_properties(*args)
        dst = {"rho_f": self.rho, "l_f": self.l_f, "nue": self.u["nue"],
               "c_f": self.u["c_f"], "cp_f": self.u["cp_f"]}
        for k, old, new in zip(self._FP_ARGS, args, out):
            if new is old:
                continue
            if k not in self._FP_SET:
                raise NotImplementedError(
                    f"fluid_properties changed {k}; the device-resident SIA step "
                    f"takes changes of {', '.join(self._FP_SET)} only")
            v = np.ascontiguousarray(np.broadcast_to(np.asarray(new, np.float64), shp))
            h.upload(v, out=dst[k])
            if k in self._fp_host:
                self._fp_host[k] = v

    def density_norm(self) -> float:
        out = ctypes.c_double()
        _lib.check(self.lib.tse_l2_norm(self.hip.ctx, self._p(self.rho),
                                        self._p(self.rho_prev), ctypes.byref(out)),
                   "tse_l2_norm")
        return out.value

    def mass(self):
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.tse_mass_sums(self.hip.ctx, self._p(self.C[0]),
                                          self._p(self.rho), self._p(self.u["phi"]),
                                          self._p(self.qx), out), "tse_mass_sums")
        return tuple(out)

    # the same two reductions enqueued without waiting; steps() yields
    # (KIND_WAIT, None) after each, so a driver of many scenarios on one
    # stream waits once for all of them, then reads them with _reduction
    def _density_norm_start(self):
        _lib.check(self.lib.tse_l2_norm_async(self.hip.ctx, self._p(self.rho),
                                              self._p(self.rho_prev)),
                   "tse_l2_norm_async")

    def _mass_start(self):
        _lib.check(self.lib.tse_mass_sums_async(self.hip.ctx, self._p(self.C[0]),
                                                self._p(self.rho), self._p(self.u["phi"]),
                                                self._p(self.qx)), "tse_mass_sums_async")

    def _reduction(self, n):
        out = (ctypes.c_double * n)()
        _lib.check(self.lib.tse_reduction_result(self.hip.ctx, out, n),
                   "tse_reduction_result")
        return tuple(out)

    def _current(self, kind):
        """The iterate buffer a fast system's result is read from: graphs of
        the result conversions are kept per buffer."""
        if not self.use_graphs:
            return 0
        c = ctypes.c_int32()
        _lib.check(self.lib.tse_fast_current(self.hip.ctx, kind, ctypes.byref(c)),
                   "tse_fast_current")
        return c.value

    def _segment(self, name, fn, kinds=()):
        """Run a segment of device work (no synchronisation inside), from a
        captured graph once it has run eagerly before."""
        if not self.use_graphs:
            fn()
            return
        n = self._seen.get(name, 0)
        self._seen[name] = n + 1
        if n == 0:
            fn()
            return
        L, ctx = self.lib, self.hip.ctx
        g = self._graphs.get(name)
        if g is None:
            self.hip.sync_stream()
            _lib.check(L.tse_graph_begin(ctx), "tse_graph_begin")
            try:
                fn()
            finally:
                h = ctypes.c_void_p()
                rc = L.tse_graph_end(ctx, ctypes.byref(h))
            _lib.check(rc, "tse_graph_end")
            g = self._graphs[name] = h
        _lib.check(L.tse_graph_launch(g, ctx), "tse_graph_launch")
        for k in kinds:   # what the captured setup set on the host
            _lib.check(L.tse_fast_mark_loaded(ctx, k), "tse_fast_mark_loaded")

    # the SIA step ----------------------------------------------------------------
    def _save_step_start(self):
        """Pn, Cn, Tn = the fields at the start of the step."""
        self.Pn.copy_(self.P)
        for s in range(self.spec.species):
            self.Cn[s].copy_(self.C[s])
        self.Tn.copy_(self.T)

    def _sia_reset(self):
        """A rejected SIA iteration (tse.py:1430-1438): keep the density,
        restart P, T, C from the step start."""
        self.rho_prev.copy_(self.rho)
        self.P.copy_(self.Pn)
        self.T.copy_(self.Tn)
        for s in range(self.spec.species):
            self.C[s].copy_(self.Cn[s])

    def solved(self, converged: bool, sweeps: int, max_sweeps: int) -> None:
        """The driver's report on the solve the step generator last yielded,
        given before resuming it: whether it reached its tolerance, and its
        sweeps out of the max_sweeps it was allowed."""
        self._outcome = (bool(converged), int(sweeps), int(max_sweeps))

    def _solve(self, kind, name, setup, system, loaded, inc_it):
        """Yield one solve (kind, system) until it converges: a solve that
        does not halves dt and runs again from the same entry state (the
        unknown is only written back after a converged solve), as the
        reference's solvers do (tse.py:730-740); dt <= min_dt fails the run.
        On success the solver's next dt grows by inc_it when the solve took
        fewer than half its sweeps (tse.py:724-729)."""
        while True:
            self._segment(f"{name}@{self.dt!r}", setup, loaded)
            self._outcome = None
            yield kind, system
            if self._outcome is None:
                raise RuntimeError("the driver of DeviceScenario.steps must call"
                                   " solved() after every solve")
            ok, sweeps, max_sweeps = self._outcome
            if ok:
                nxt = self.dt
                if sweeps < max_sweeps / 2:
                    nxt = min(self.dt * inc_it, self.dt_max)
                self._next_dt[name] = nxt
                return
            self.dt = self.dt / 2
            self.halvings += 1
            if self.dt <= self.min_dt:
                raise RuntimeError(f"{name}: dt fell to {self.dt:.3e} s without"
                                   " convergence")

    def steps(self, nsteps: int):
        """Generator over ``nsteps`` SIA steps; yields (kind, system) for each
        solve and resumes once the system holds its solution (the driver
        reports each outcome with ``solved`` first), and (KIND_WAIT, None)
        where it is about to read a reduction from the device (resume at
        once, or after a wait that covers its stream)."""
        sp = self.spec
        for _ in range(nsteps):
            st = StepStats()
            self._next_dt = {}
            self._segment("step_begin", self._save_step_start)
            for it in range(sp.sia_maxit):
                st.sia_iters = it + 1
                yield from self._solve(0, "setup_p", self.setup_p, self.fp, (0,), 1.25)
                self._segment(f"post_p{self._current(0)}@{self.dt!r}",
                              lambda: (self.finish_p(), self.flow()))
                for s in range(sp.species):
                    self.ft.assemble = lambda s=s: self.setup_c(s)
                    yield from self._solve(1, f"setup_c{s}",
                                           lambda s=s: self.setup_c(s), self.ft,
                                           (1,), 1.05)
                    self._segment(f"finish_c{s}.{self._current(1)}",
                                  lambda s=s: self.finish_tr(self.C[s]))
                if sp.heat:
                    self.ft.assemble = self.setup_t
                    yield from self._solve(1, "setup_t", self.setup_t, self.ft,
                                           (1,), 1.05)
                    self._segment(f"finish_t{self._current(1)}",
                                  lambda: self.finish_tr(self.T))
                if self.fluid_properties is None:
                    self._segment("eos", self.eos_update)
                else:            # host work: never captured
                    self.eos_update()
                self._density_norm_start()
                yield KIND_WAIT, None
                diff = self._reduction(1)[0]
                if diff < sp.sia_conv or it == sp.sia_maxit - 1:
                    break
                self._segment("sia_reset", self._sia_reset)
            self._mass_start()
            yield KIND_WAIT, None
            st.mass = self._reduction(3)
            st.dt = self.dt
            self.t += self.dt
            self.dt_step = self.dt
            # next dt: the smallest of the solvers' next dt (tse.py:1486-1492)
            self.dt = min(self._next_dt.values()) if self._next_dt else self.dt
            self.n_steps += 1
            self.history.append(st)

    def run(self, nsteps: int, form="mixed", tol=1e-10, plan_p=None,
            plan_tr=None, sor_p=None, sor_tr=1.0, inner_p=100, inner_tr=10,
            max_sweeps=400000, strict=True, progress=None, adaptive=False):
        """Drive ``steps`` with single-system solves; returns per-solve stats.
        sor_p None: the pressure factor comes from the scenario's tuner.
        strict: raise if a solve stops at max_sweeps above its tolerance;
        else the step halves dt and repeats the solve (tse.py:730-740).
        progress: called with (kind, stats) after every solve.
        adaptive: pressure cycles sized from the convergence rate (the
        batched drivers use fixed cycles)."""
        out = []
        for kind, sysm in self.steps(nsteps):
            if kind == KIND_WAIT:
                continue
            if kind == 0:
                om = self.omega_p.omega if sor_p is None else sor_p
                r = sysm.solve(form, tol=tol, sor=om, inner=inner_p,
                               plan=plan_p, max_sweeps=max_sweeps, adaptive=adaptive)
                if sor_p is None:
                    self.omega_p.report(r.sweeps, r.contraction)
            else:
                r = sysm.solve(form, tol=tol, sor=sor_tr, inner=inner_tr,
                               plan=plan_tr, max_sweeps=max_sweeps)
            if strict and not r.converged:
                raise RuntimeError(f"{'P' if kind == 0 else 'transport'} solve did"
                                   f" not converge: rel {r.rel_residual:.3e}"
                                   f" after {r.sweeps} sweeps")
            self.solved(r.converged, r.sweeps, max_sweeps)
            if progress:
                progress(kind, r)
            out.append((kind, r))
        return out

    def snapshot(self, path: str, writer) -> None:
        """Queue the 34-tuple pickle of tse.py:322-327 for the current state
        on `writer` (a snapshot.PickleWriter).  The fields are converted to
        the reference layout on the compute stream; the copy to pinned host
        memory runs on a side stream and the writer thread waits for it, so
        the caller may go on with the next step at once."""
        import torch
        sp, hip, L = self.spec, self.hip, self.lib
        nx, ny, nz, S = sp.nx, sp.ny, sp.nz, sp.species
        N = nx * ny * nz
        own = (("P", self.P), ("T", self.T), ("qx", self.qx), ("qy", self.qy),
               ("qz", self.qz), ("K", self.K), ("rho_f", self.rho),
               ("l_f", self.l_f), ("BP", self.BP), ("BC", self.BC),
               ("BT", self.BT), ("e", self.e))
        stage = torch.empty((len(own) + S) * N, dtype=torch.float64,
                            device=hip.device)
        base = stage.data_ptr()
        hip.sync_stream()
        for n, (_, t) in enumerate(own):
            _lib.check(L.tse_from_split(hip.ctx, base + 8 * n * N, ny * nz, nz, 1,
                                        t.data_ptr()), "tse_from_split")
        c0 = len(own) * N
        for s_ in range(S):   # species last: C[i, j, k, s]
            _lib.check(L.tse_from_split(hip.ctx, base + 8 * (c0 + s_),
                                        ny * nz * S, nz * S, S,
                                        self.C[s_].data_ptr()), "tse_from_split")
        host = torch.empty(stage.numel(), dtype=torch.float64, pin_memory=True)
        if not hasattr(self, "_copy_stream"):
            self._copy_stream = torch.cuda.Stream(hip.device)   # default priority
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(hip.device))
        self._copy_stream.wait_event(ready)
        with torch.cuda.stream(self._copy_stream):
            host.copy_(stage, non_blocking=True)
        done = torch.cuda.Event()
        done.record(self._copy_stream)
        stage.record_stream(self._copy_stream)
        t_sim, dt = self.t, self.dt_step
        mass = np.array(self.history[-1].mass if self.history else (0.0, 0.0, 0.0))
        dxs = np.full(nx, float(sp.spacing))
        dys = np.full(ny, float(sp.spacing))
        dzs = np.full(nz, float(sp.spacing))

        def build():
            h = host.numpy()
            f = {name: h[n * N:(n + 1) * N].reshape(nx, ny, nz)
                 for n, (name, _) in enumerate(own)}
            C = h[c0:c0 + S * N].reshape(nx, ny, nz, S)
            shp = (nx, ny, nz)
            full = lambda name: np.full(shp, UNIFORM[name])
            K = f["K"]
            return (float(dxs.sum()), float(dys.sum()), float(dzs.sum()), nx, ny,
                    nz, dxs, dys, dzs, t_sim, dt, f["P"], C, f["T"], f["qx"],
                    f["qy"], f["qz"], full("phi"), K, K, K, f["rho_f"],
                    full("nue"), full("cp_f"), f["l_f"], f["BP"], f["BC"],
                    f["BT"], np.zeros(shp), mass, np.zeros(shp), np.zeros(shp),
                    np.zeros(shp + (S,)), f["e"])
        writer.submit(path, build, wait=done.synchronize)

    def download(self) -> dict:
        h = self.hip
        return {"P": h.download(self.P), "T": h.download(self.T),
                "C": [h.download(c) for c in self.C], "rho_f": h.download(self.rho)}

    def close(self):
        for g in self._graphs.values():
            self.lib.tse_graph_destroy(g)
        self._graphs.clear()
        self.hip.sync_stream()
        self.u = {}
        for key, _ in self._ukeys.values():
            _release_uniform(key)
        self._ukeys.clear()
        self.hip.close()
