"""Time-stepping / SIA Picard driver around the hot-path solvers.

A restatement of the TRANSPORTSE main loop (tse.py:1118-1494) that calls the
solvers through the reference's own signatures, so that config files, the
EOS callback, time-step control, chemistry callbacks, dynamic boundary
conditions, mass balance and the 34-tuple pickle output run unchanged. The
order of operations and the state handling follow the reference exactly:

  P (tse.py:1214) -> q (1228) -> e (1236) -> CFL (1241-1249) -> Peclet
  (1255-1260) -> dt_max (1268-1271) -> CFL clamp (1289-1292) -> species
  (1296-1336) -> heat (1338-1381) -> chemistry (1383-1418) -> EOS (1422)
  -> SIA test with rhof_prev starting at zeros (1179, 1425-1442)
  -> dynamic_bc (1449-1458) -> mass_balance (1461-1463) -> t_sim += dt
  -> next dt = min(dt_P, dt_T, dt_C, t - t_sim) (1486-1492)
  -> plotresults dt clamp / dump (312-329).

The solver backend is any object with compute_P, compute_q, flow_error,
compute_C, compute_T and l2_norm; ``make_hip_backend`` builds the GPU one.
The Smolarkiewicz correction passes (SMOLAR_C / SMOLAR_T, tse.py:1318-1336,
1368-1381) and NLTE two-layer heat (NLTE_SOLV, tse.py:1343-1354; the fluid
properties then act on layer 0, tse.py:207-226) follow the reference.
"""
from __future__ import annotations

import functools
import os
import sys
import time
import types

import numpy

from .config_loader import load_config
from .device import pcopy
from .snapshot import PickleWriter

# config attributes the reference binds at start-up (tse.py:81-151)
_STATE_ARRAYS = ("phi", "Kx", "Ky", "Kz", "c_r", "cp_r", "l_r", "rho_r",
                 "lf_l", "lf_t", "lt_l", "lt_t", "L", "R", "D_fx", "D_fy",
                 "D_fz", "c_f", "cp_f", "l_f", "rho_f", "nue", "P", "C", "T",
                 "BP", "BC", "BT", "WP", "WT", "WC", "qx", "qy", "qz", "mass")

_FP_NAMES = ("C", "T", "P", "nue", "rho_f", "c_f", "cp_f", "l_f", "cp_r",
             "l_r", "phi", "Kx", "Ky", "Kz", "D_fx", "D_fy", "D_fz", "rho_r")


# threads of the whole-field EOS call (numpy releases the GIL in its loops)
_EOS_THREADS = 8
_EOS_MIN_CELLS = 1 << 19     # cells per slab at least
_EOS_POOL = None


def _eos_pool():
    global _EOS_POOL
    if _EOS_POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _EOS_POOL = ThreadPoolExecutor(max_workers=_EOS_THREADS)
    return _EOS_POOL


class ConvergenceError(SystemExit):
    """The reference calls sys.exit on solver non-convergence (tse.py:1220)."""


class Simulation:
    """One TRANSPORTSE run of a config module with an injected backend."""

    def __init__(self, config, backend_factory, out_dir=None, verbose=True,
                 record_fields=False, max_steps=None, eos_mode="auto",
                 lazy_host=True):
        if isinstance(config, str):
            self.cfile = os.path.splitext(os.path.basename(config))[0]
            self.cdir = os.path.dirname(os.path.abspath(config))
            config = load_config(config)
        else:
            self.cfile = getattr(config, "__name__", "config").split(".")[-1]
            self.cdir = os.path.dirname(os.path.abspath(getattr(config, "__file__", ".")))
        self.config = config
        self.out_dir = out_dir if out_dir is not None else self.cdir
        self.verbose = verbose
        self.record_fields = record_fields
        self.max_steps = max_steps
        self.eos_mode = eos_mode
        c = config
        # tse.py:81-167 ------------------------------------------------------
        st = types.SimpleNamespace()
        st.lx, st.ly, st.lz = c.lx, c.ly, c.lz
        st.nx, st.ny, st.nz = int(c.nx), int(c.ny), int(c.nz)
        st.dx, st.dy, st.dz = c.dx, c.dy, c.dz
        st.g = c.g
        st.beta_deg = c.beta
        st.t = c.t
        st.dt_max = c.dt_max
        for name in _STATE_ARRAYS:
            setattr(st, name, getattr(c, name))
        shp = (st.nx, st.ny, st.nz)
        st.vx = numpy.zeros(shp)
        st.vy = numpy.zeros(shp)
        st.vz = numpy.zeros(shp)
        st.PeC = numpy.zeros(shp + (3,))
        st.PeT = numpy.zeros(shp + (3,))
        st.e = numpy.zeros(shp)
        st.ZEROS = numpy.zeros(shp)
        st.ONES = numpy.ones(shp)
        self.st = st
        # tse.py:231-245
        self.MAXIT_P = c.MAXIT_P if getattr(c, "MAXIT_P", 0) else 20000
        self.MAXIT_C = c.MAXIT_C if getattr(c, "MAXIT_C", 0) else 100
        self.MAXIT_T = c.MAXIT_T if getattr(c, "MAXIT_T", 0) else 100
        self.NLTE = bool(getattr(c, "NLTE_SOLV", False))   # 'NLTE_SOLV' in dir(config)
        self._eos_vectorized = None
        if c.FLUIDPROP:
            self.update_fluid_properties()
        # the backend snapshots frozen globals at first call from `st`
        self.solvers = backend_factory(self)
        if hasattr(self.solvers, "pe_writeback"):
            self.solvers.pe_writeback = False
        # deferred host copies of what the solvers write (backends that keep
        # fields on a device): the loop syncs an array before host code
        # reads it -- a callback gets its arguments current, output and the
        # recorded trajectory everything -- and copies fields on the device
        if lazy_host and hasattr(self.solvers, "lazy"):
            self.solvers.lazy = True
        self.trajectory = []
        self.writer = PickleWriter()
        self._wc_cache = {}

    # ---- EOS (tse.py:178-229) -------------------------------------------------
    def _fp_arrays(self):
        """The fluid_properties operands; a 4-D (NLTE) T contributes its
        fluid layer 0 (tse.py:207-226)."""
        st = self.st
        return [st.T[:, :, :, 0] if n == "T" and st.T.ndim == 4 else getattr(st, n)
                for n in _FP_NAMES]

    def _eos_loop(self):
        st = self.st
        arrs = self._fp_arrays()
        fp = self.config.fluid_properties
        for k in range(st.nz):
            for j in range(st.ny):
                for i in range(st.nx):
                    out = fp(*[a[i, j, k] for a in arrs])
                    for a, v in zip(arrs, out):
                        a[i, j, k] = v

    def _eos_slabs(self):
        """x-slabs of the whole-field EOS call: up to _EOS_THREADS, each of
        at least _EOS_MIN_CELLS cells (one slab for small grids)."""
        st = self.st
        cells = st.nx * st.ny * st.nz
        k = max(1, min(_EOS_THREADS, st.nx, cells // _EOS_MIN_CELLS))
        step = -(-st.nx // k)
        return [(x0, min(x0 + step, st.nx)) for x0 in range(0, st.nx, step)]

    def _eos_vector(self):
        """Call fluid_properties on whole fields (species axis first), one
        call per x-slab on the EOS threads.  The reference calls it cell by
        cell inside prange loops (tse.py:181-196), so the cells of a call
        are independent of each other and of the other slabs' calls; the
        sample check holds the assembled result to that contract.  Returns
        per output a list of (x0, x1, slab values)."""
        arrs = self._fp_arrays()
        fp = self.config.fluid_properties
        slabs = self._eos_slabs()

        def call(x0, x1):
            args = []
            for a in arrs:
                v = (numpy.moveaxis(a, -1, 0)[:, x0:x1] if a.ndim == 4
                     else a[x0:x1]).view()
                # read-only: an argument returned unchanged is known to be
                # unchanged (an in-place write raises and selects the loop)
                v.flags.writeable = False
                args.append(v)
            out = fp(*args)
            same = [v is arg for v, arg in zip(out, args)]
            vals = []
            for a, v in zip(arrs, out):
                v = numpy.asarray(v, dtype=numpy.float64)
                if a.ndim == 4 and v.ndim == 4:
                    v = numpy.moveaxis(v, 0, -1)  # back to species-last
                vals.append(numpy.broadcast_to(v, a[x0:x1].shape))
            return same, vals

        if len(slabs) == 1:
            parts = [call(*slabs[0])]
        else:
            futs = [_eos_pool().submit(call, x0, x1) for x0, x1 in slabs]
            parts = [f.result() for f in futs]
        if any(len(p[1]) != len(arrs) for p in parts):
            raise ValueError("fluid_properties returned the wrong number of values")
        self._eos_same = [all(p[0][n] for p in parts) for n in range(len(arrs))]
        return [[(x0, x1, p[1][n]) for (x0, x1), p in zip(slabs, parts)]
                for n in range(len(arrs))]

    def update_fluid_properties(self):
        """Returns the state arrays whose values may have changed."""
        if self.eos_mode == "loop" or self._eos_vectorized is False:
            self._eos_loop()
            return self._fp_arrays()
        try:
            res = self._eos_vector()
            # every call is checked against the per-cell contract: the first
            # on a broad sample, later ones on fresh random cells
            ok = self._validate_eos(res, first=not self._eos_vectorized)
        except Exception:
            ok = False
        if not ok:
            self._eos_vectorized = False
            self._eos_loop()
            return self._fp_arrays()
        self._eos_vectorized = True
        changed = []
        jobs = []
        for a, parts, same in zip(self._fp_arrays(), res, self._eos_same):
            if not same:       # the callback returned its argument unchanged
                jobs += [(a[x0:x1], v) for x0, x1, v in parts]
                changed.append(a)
        if len(jobs) > 1:
            for f in [_eos_pool().submit(numpy.copyto, d, v) for d, v in jobs]:
                f.result()
        else:
            for d, v in jobs:
                numpy.copyto(d, v)
        return changed

    def _validate_eos(self, res, first=True) -> bool:
        """Check the whole-field call against the per-cell contract
        (tse.py:181-196: fluid_properties called cell by cell) bit for bit on
        a sample of cells: the grid's corners and random cells at every call
        (16 new ones), and at the first call also the cells holding the
        minimum and maximum of every input (where value-dependent branches
        of a callback act first) and 1024 random cells.  Any difference
        selects the per-cell loop for this call and the rest of the run."""
        st = self.st
        shp = (st.nx, st.ny, st.nz)
        if not hasattr(self, "_eos_rng"):
            self._eos_rng = numpy.random.default_rng(0)
        arrs = self._fp_arrays()
        cells = [(0, 0, 0), (st.nx - 1, st.ny - 1, st.nz - 1)]
        for a in (arrs if first else ()):
            v = numpy.asarray(a)
            if v.ndim == 4:
                v = v.reshape(shp + (-1,)).max(axis=-1)
            if v.shape != shp or v.size == 0:
                continue
            for f in (numpy.argmin, numpy.argmax):
                cells.append(numpy.unravel_index(f(v), shp))
        n = min(1024 if first else 16, st.nx * st.ny * st.nz)
        cells = numpy.concatenate(
            [numpy.array(cells, dtype=numpy.int64).reshape(-1, 3),
             self._eos_rng.integers(0, shp, size=(n, 3))])
        fp = self.config.fluid_properties
        refs = [fp(*[a[i, j, k] for a in arrs]) for i, j, k in cells.tolist()]
        I, J, K = cells.T
        for n, parts in enumerate(res):
            got = numpy.empty((len(I),) + parts[0][2].shape[3:], dtype=numpy.float64)
            for x0, x1, v in parts:
                sel = (I >= x0) & (I < x1)
                got[sel] = v[I[sel] - x0, J[sel], K[sel]]
            want = numpy.array([numpy.asarray(r[n], dtype=numpy.float64)
                                for r in refs], dtype=numpy.float64)
            if got.shape != want.shape or not numpy.array_equal(
                    got.view(numpy.int64), want.view(numpy.int64)):
                return False
        return True

    # ---- diagnostics (tse.py:331-396) ---------------------------------------------
    # peclet_C / peclet_T and the CFL maxima run on the backend (the device
    # for the GPU backend, which keeps PeC / PeT there: the driver only
    # reads their maxima)
    def explicit_C(self):
        st = self.st
        inv = (1 / st.dx[:, None, None] ** 2 + 1 / st.dy[None, :, None] ** 2) \
            + 1 / st.dz[None, None, :] ** 2
        return float(numpy.min(0.45 * numpy.max(st.D_fx, axis=-1) / st.phi / inv))

    # ---- output (tse.py:312-329) ----------------------------------------------------
    def plotresults(self):
        c, st = self.config, self.st
        if not c.PLOTSAVE:
            return
        if self.PLOTINDEX <= numpy.shape(c.PLOTTIME)[0] - 1:
            t_plot = c.PLOTTIME[self.PLOTINDEX]
            if self.t_sim < t_plot:
                self.dt_prev = self.dt
                self.dt = min(self.dt, t_plot - self.t_sim)
            elif self.t_sim == t_plot:
                self._sync()
                path = os.path.join(self.out_dir, "{:s}-{:05d}.p".format(
                    self.cfile, self.PLOTINDEX))
                # the 34-tuple of tse.py:322-327; arrays are snapshot-copied so
                # the compressed dump runs on the writer thread while the time
                # loop continues
                snap = tuple(numpy.array(v, copy=True) if isinstance(v, numpy.ndarray)
                             else v for v in (
                    st.lx, st.ly, st.lz, st.nx, st.ny, st.nz, st.dx, st.dy, st.dz,
                    self.t_sim, self.dt, st.P, st.C, st.T, st.qx, st.qy, st.qz,
                    st.phi, st.Kx, st.Ky, st.Kz, st.rho_f, st.nue, st.cp_f,
                    st.l_f, st.BP, st.BC, st.BT, st.L, st.mass, st.WP, st.WT,
                    st.WC, st.e))
                self.writer.submit(path, snap)
                self.dt = self.dt_prev
                self.PLOTINDEX += 1

    def _sync(self, arrays=None):
        """Make host arrays current (all, or those overlapping `arrays`)."""
        f = getattr(self.solvers, "sync", None)
        if f is not None:
            f(arrays)

    def _copy(self, a):
        """``a.copy()``, on the backend's device where it holds the value."""
        f = getattr(self.solvers, "copy", None)
        return f(a) if f is not None else a.copy()

    def _nonzero_sum(self, a) -> bool:
        """``numpy.sum(a) != 0.0`` (tse.py:1300-1301)."""
        f = getattr(self.solvers, "nonzero_sum", None)
        return f(a) if f is not None else bool(numpy.sum(a) != 0.0)

    def _wc_nonzero(self, n) -> bool:
        """``numpy.sum(WC[:, :, :, n]) != 0.0`` (tse.py:1301), kept until a
        callback that receives WC (mass_balance, dynamic_bc) or one that may
        replace it runs."""
        k = (id(self.st.WC), n)
        v = self._wc_cache.get(k)
        if v is None:
            v = self._wc_cache[k] = bool(numpy.sum(self.st.WC[:, :, :, n]) != 0.0)
        return v

    def _store(self, dst, n, val):
        """``dst[:, :, :, n] = val`` (tse.py:1304, 1345-1353); nothing to do
        when the solver returned that very slice."""
        target = dst[:, :, :, n]
        v = numpy.asarray(val)
        if (v.__array_interface__["data"][0] == target.__array_interface__["data"][0]
                and v.strides == target.strides and v.shape == target.shape):
            return
        self._sync([v])
        target[...] = v

    def _lazy(self) -> bool:
        return bool(getattr(self.solvers, "lazy", False))

    def _host_mutated(self, arrays=None):
        """Tell the backend that host code changed `arrays` (None: any)."""
        inv = getattr(self.solvers, "invalidate", None)
        if inv is not None:
            inv(arrays)

    def _print(self, *a, **k):
        if self.verbose:
            print(*a, **k)

    # ---- main loop (tse.py:1118-1494) --------------------------------------------------
    def run(self):
        """Run the time loop; pickles still being written are finished before
        this returns (or re-raises)."""
        try:
            return self._run()
        finally:
            self._sync()
            self.writer.flush()

    def _run(self):
        c, st, sv = self.config, self.st, self.solvers
        self.dt = c.dt
        self.dt_prev = self.dt
        self.t_sim = getattr(c, "t_sim", 0)
        self.PLOTINDEX = 0
        n_tsteps = 0
        numitP = numitC = numitT = num_CFL = 0
        max_DT = 0
        Pex_T = Pey_T = Pez_T = Pex_C = Pey_C = Pez_C = 0
        bc_inc = chem_inc = dt_chem = 0
        dt_max_abs = st.dt_max
        dt_P = dt_C = dt_T = None
        Pn = Tn = Cn = None
        time_start = time.time()
        self.plotresults()
        rhof_prev = numpy.zeros_like(st.rho_f)
        solve_time = 0.0

        while self.t_sim < st.t:
            if self.max_steps is not None and n_tsteps >= self.max_steps:
                break
            SIA_filled = False
            SIA_iter = 0
            SIA_diff = 0
            while not SIA_filled and SIA_iter < c.SIA_maxit:
                Co_Filled = False
                CFL = 0
                if c.OMEGA_C == 0:
                    dt_NeC = self.explicit_C()
                num_it_P = 0
                if c.FLUIDFLOW:
                    while not Co_Filled:
                        t0 = time.perf_counter()
                        Pn = None    # replaced below; the backend need not keep it
                        self.dt, dt_P, st.P, Pn, converged_P, num_it_P = \
                            sv.compute_P(st.P, st.c_f, st.nue, st.rho_f,
                                         st.phi, st.Kx, st.Ky, st.Kz, st.qx,
                                         st.qy, st.qz, self.dt, st.BP, st.WP,
                                         max_dt=st.dt_max, inc_it=1.25,
                                         SOR=c.SOR_P, max_it=self.MAXIT_P,
                                         P_conv=c.CONV_P)
                        if not converged_P:
                            raise ConvergenceError(
                                '* P CONVERGENCE ERROR - check dt and CONV_P in compute_P')
                        st.qx, st.qy, st.qz = sv.compute_q(
                            st.P, st.BP, st.Kx, st.Ky, st.Kz, st.qx, st.qy,
                            st.qz, st.nue, st.rho_f)
                        if c.HEATTRANS or c.SPECTRANS:
                            st.e = None
                            st.e = sv.flow_error(st.P, Pn, st.qx, st.qy, st.qz,
                                                 st.WP, st.phi, st.c_f, st.c_r,
                                                 st.rho_f, self.dt)
                        solve_time += time.perf_counter() - t0
                        if self.t_sim > 0:
                            # tse.py:1241-1246 (its reshapes included)
                            m = sv.cfl_maxima(st.qx, st.qy, st.qz, st.phi)
                            with numpy.errstate(divide="ignore", invalid="ignore"):
                                dt_CFL = 0.9 / (numpy.float64(m[0]) + m[1] + m[2])
                        else:
                            dt_CFL = 1e10
                        dt_max_C = dt_max_T = dt_max_P = dt_max_abs
                        if c.HEATTRANS:
                            Pex_T, Pey_T, Pez_T = sv.peclet_T(
                                st.PeT, st.qx, st.qy, st.qz, st.rho_f, st.cp_f,
                                st.phi, st.l_r, st.l_f)
                        if c.SPECTRANS:
                            Pex_C, Pey_C, Pez_C = sv.peclet_C(
                                st.PeC, st.qx, st.qy, st.qz, st.D_fx, st.D_fy,
                                st.D_fz)
                        if c.OMEGA_C == 0:
                            dt_max_C = dt_NeC
                        if c.SIA:
                            st.dt_max = min(dt_max_C, dt_max_T, dt_max_P)
                        else:
                            st.dt_max = min(dt_CFL, dt_max_C, dt_max_T, dt_max_P)
                        if n_tsteps % c.PRINTOUT == 0 and self.verbose:
                            self._sync([st.P, st.T, st.C])
                            if c.OMEGA_C + c.OMEGA_T < 2:
                                print(' [dt_NeC] %8.2e [dt_NeT] %8.2e' % (dt_max_C, dt_max_T))
                            print(' [CFL] %6.2f%% | Pe_T %.1e/%.1e/%.1e | Pe_C %.1e/%.1e/%.1e'
                                  % (self.dt / dt_CFL * 100, Pex_T, Pey_T, Pez_T,
                                     Pex_C, Pey_C, Pez_C))
                            print('[P+] %6.2e [P-] %6.2e [T+] %6.4f [T-] %6.4f'
                                  % (numpy.max(st.P), numpy.min(st.P),
                                     numpy.max(st.T), numpy.min(st.T))
                                  + ' [C+] %6.4f [C-] %6.4f' % (numpy.max(st.C),
                                                                numpy.min(st.C)),
                                  end='')
                        Co_Filled = True
                        if self.dt > st.dt_max * 1.0:
                            num_CFL += num_it_P
                            CFL += 1
                            self.dt = st.dt_max

                num_it_C = 0
                if c.SPECTRANS:
                    Cn = self._copy(st.C)
                    for n in c.SPTRANS:
                        if self._nonzero_sum(st.C[:, :, :, n]) \
                                or self._wc_nonzero(n):
                            t0 = time.perf_counter()
                            self.dt, dt_C, res_C, converged_C, num_itC = \
                                sv.compute_C(st.C[:, :, :, n], st.D_fx[:, :, :, n],
                                             st.D_fy[:, :, :, n], st.D_fz[:, :, :, n],
                                             st.phi, st.rho_f, st.qx, st.qy, st.qz,
                                             self.dt, st.BC, st.WC[:, :, :, n], st.R,
                                             st.e, omega=c.OMEGA_C, SOR=c.SOR_C,
                                             max_it=self.MAXIT_C, C_conv=c.CONV_C)
                            self._store(st.C, n, res_C)
                            solve_time += time.perf_counter() - t0
                            num_it_C += num_itC
                            if not converged_C:
                                raise ConvergenceError(
                                    '* C CONVERGENCE ERROR- check dt and C_conv in compute_C')
                            if c.SMOLAR_C:  # tse.py:1318-1336
                                t0 = time.perf_counter()
                                st.vx, st.vy, st.vz = sv.smolarkiewicz(
                                    st.C[:, :, :, n], st.qx, st.qy, st.qz, st.vx,
                                    st.vy, st.vz, self.dt, eps=1e-15, Sc=1.08)
                                self.dt, dt_C, res_C, converged_C, num_itC = \
                                    sv.compute_C(st.C[:, :, :, n], st.D_fx[:, :, :, n],
                                                 st.D_fy[:, :, :, n], st.D_fz[:, :, :, n],
                                                 st.phi, st.rho_f, st.vx, st.vy, st.vz,
                                                 self.dt, st.BC, st.WC[:, :, :, n], st.R,
                                                 st.ZEROS, omega=c.OMEGA_C,
                                                 max_it=self.MAXIT_C, SOR=c.SOR_C,
                                                 C_conv=c.CONV_C)
                                self._store(st.C, n, res_C)
                                solve_time += time.perf_counter() - t0
                                num_it_C += num_itC
                                if not converged_C:
                                    raise ConvergenceError(
                                        '* C2 CONVERGENCE ERROR- check dt and C_conv in compute_C')

                num_it_T = 0
                if c.HEATTRANS:
                    Tn = self._copy(st.T)
                    t0 = time.perf_counter()
                    if self.NLTE:  # tse.py:1343-1354
                        # fluid layer: unit porosity, no source
                        self.dt, dt_T, res_T, converged_T, num_itT1 = \
                            sv.compute_T(st.T[:, :, :, 0], st.l_f, st.cp_f, st.rho_f,
                                         st.ONES, st.qx, st.qy, st.qz, self.dt,
                                         st.BT, st.ZEROS, st.e, omega=c.OMEGA_T,
                                         max_it=self.MAXIT_T, SOR=c.SOR_T,
                                         T_conv=c.CONV_T)
                        self._store(st.T, 0, res_T)
                        # rock layer: zero fluid properties
                        self.dt, dt_T, res_T1, converged_T, num_itT2 = \
                            sv.compute_T(st.T[:, :, :, 1], st.ZEROS, st.ZEROS,
                                         st.ZEROS, st.ZEROS, st.qx, st.qy, st.qz,
                                         self.dt, st.BT, st.WT, st.e,
                                         omega=c.OMEGA_T, max_it=self.MAXIT_T,
                                         SOR=c.SOR_T, T_conv=c.CONV_T)
                        self._store(st.T, 1, res_T1)
                        num_itT = num_itT1 + num_itT2
                    else:
                        self.dt, dt_T, st.T, converged_T, num_itT = \
                            sv.compute_T(st.T, st.l_f, st.cp_f, st.rho_f, st.phi,
                                         st.qx, st.qy, st.qz, self.dt, st.BT, st.WT,
                                         st.e, omega=c.OMEGA_T, max_it=self.MAXIT_T,
                                         SOR=c.SOR_T, T_conv=c.CONV_T)
                    solve_time += time.perf_counter() - t0
                    num_it_T += num_itT
                    if not converged_T:
                        raise ConvergenceError(
                            '* T CONVERGENCE ERROR - check dt and T_conv in compute_T')
                    if c.SMOLAR_T:  # tse.py:1368-1381
                        t0 = time.perf_counter()
                        st.vx, st.vy, st.vz = sv.smolarkiewicz(
                            st.T, st.qx, st.qy, st.qz, st.vx, st.vy, st.vz,
                            self.dt, eps=1e-15, Sc=1.08)
                        self.dt, dt_T, st.T, converged_T, num_itT = \
                            sv.compute_T(st.T, st.l_f, st.cp_f, st.rho_f, st.phi,
                                         st.vx, st.vy, st.vz, self.dt, st.BT, st.WT,
                                         st.ZEROS, omega=c.OMEGA_T,
                                         max_it=self.MAXIT_T, SOR=c.SOR_T,
                                         T_conv=c.CONV_T)
                        solve_time += time.perf_counter() - t0
                        num_it_T += num_itT
                        if not converged_T:
                            raise ConvergenceError(
                                '* T CONVERGENCE ERROR - check dt and T_conv in compute_T')

                if c.CHEMREACT:
                    if self.t_sim > 0:
                        chem_inc += self.dt
                    else:
                        chem_inc = 0
                        dt_chem = self.dt
                    if chem_inc >= c.DTCHEM:
                        dt_chem = chem_inc
                        chem_inc = 0
                    chem_list = [(i, j, k) for k in range(st.nz)
                                 for j in range(st.ny) for i in range(st.nx)
                                 if st.BC[i, j, k] != 1]
                    if chem_list != [] and chem_inc == 0:
                        self._sync()
                        self._wc_cache.clear()
                        self._chemistry(chem_list, dt_chem)
                        self._host_mutated([st.C, st.P, st.T, st.rho_f, st.phi,
                                            st.Kx, st.Ky, st.Kz])

                if c.FLUIDPROP:
                    self._sync(self._fp_arrays())
                    self._host_mutated(self.update_fluid_properties())

                if c.SIA:
                    SIA_diff = sv.l2_norm(st.rho_f, rhof_prev, 'rhof')
                    if n_tsteps % c.PRINTOUT == 0:
                        self._print(' [SIA] %4d [SIA] %6.2e' % (SIA_iter, SIA_diff))
                    if SIA_diff < c.SIA_conv or SIA_iter == c.SIA_maxit - 1:
                        SIA_filled = True
                    else:
                        rhof_prev = pcopy(st.rho_f)
                        st.P = self._copy(Pn)
                        if c.HEATTRANS:
                            st.T = self._copy(Tn)
                        if c.SPECTRANS:
                            st.C = self._copy(Cn)
                        SIA_iter += 1
                        if not self._lazy():
                            # new arrays: the backend's cache keys no longer
                            # match (a lazy backend tracks the copies itself)
                            self._host_mutated([st.P, st.T, st.C])
                else:
                    SIA_filled = True
                    if n_tsteps % c.PRINTOUT == 0:
                        self._print('')
                numitP += num_it_P
                numitC += num_it_C
                numitT += num_it_T

            if c.DYNAMICBC:
                self._sync()
                self._wc_cache.clear()
                bc_inc_ = bc_inc
                bc_inc, st.P, st.T, st.C, st.BP, st.BT, st.BC, st.WP, st.WT, \
                    st.WC, st.Kx, st.Ky, st.Kz = c.dynamic_bc(
                        self.t_sim + self.dt, bc_inc, st.P, st.T, st.C, st.BP,
                        st.BT, st.BC, st.WP, st.WT, st.WC, st.Kx, st.Ky, st.Kz,
                        st.rho_f, st.nue, st.c_f, st.cp_f, st.qx, st.phi,
                        st.l_f, st.l_r)
                # tse.py:1451-1453 also hands rho_f, nue, c_f, cp_f, qx, phi,
                # l_f and l_r to the callback, which may change any of them in
                # place: every device copy is stale from here on
                self._host_mutated(None)
                if bc_inc_ != bc_inc and c.FLUIDPROP:
                    self._host_mutated(self.update_fluid_properties())

            if c.MASSBALAN:
                self._sync([st.C, st.qx, st.qy, st.qz, st.WP, st.WC, st.WT,
                            st.D_fx, st.D_fy, st.D_fz])
                self._wc_cache.clear()
                st.mass = c.mass_balance(st.C, st.qx, st.qy, st.qz, st.WP,
                                         st.WC, st.WT, st.D_fx, st.D_fy,
                                         st.D_fz, self.dt, st.mass)

            step_dt = self.dt
            self.t_sim += self.dt
            n_tsteps += 1
            max_DT = max(max_DT, self.dt)
            rec = dict(step=n_tsteps, t_sim=self.t_sim, dt=step_dt,
                       numitP=numitP, numitC=numitC, numitT=numitT,
                       dt_P=dt_P, dt_C=dt_C, dt_T=dt_T,
                       mass=numpy.array(st.mass, dtype=float).copy())
            if self.record_fields:
                self._sync()
                rec.update(P=st.P.copy(), C=st.C.copy(), T=st.T.copy(),
                           qx=st.qx.copy(), qy=st.qy.copy(), qz=st.qz.copy(),
                           e=st.e.copy(), rho_f=st.rho_f.copy())
            self.trajectory.append(rec)

            if n_tsteps % c.PRINTOUT == 0 and self.verbose:
                print('[t] %8.2e/%8.2e [dt] %8.2e [S] %4d [P] %3d'
                      % (self.t_sim, st.t, self.dt, n_tsteps, num_it_P)
                      + ' [C] %3d [T] %3d [CFL] %d' % (num_it_C, num_it_T, CFL))

            self.dt = 1e99
            if c.FLUIDFLOW:
                self.dt = min(dt_P, st.t - self.t_sim)
            if c.HEATTRANS:
                self.dt = min(dt_T, st.t - self.t_sim, self.dt)
            if c.SPECTRANS:
                if dt_C is None:
                    raise NameError("name 'dt_C' is not defined")
                self.dt = min(dt_C, st.t - self.t_sim, self.dt)
            self.plotresults()

        self.wall_time = time.time() - time_start
        self.solve_time = solve_time
        self.summary = dict(t_sim=self.t_sim, n_tsteps=n_tsteps, numitP=numitP,
                            numitC=numitC, numitT=numitT, num_CFL=num_CFL,
                            max_DT=max_DT, wall_time=self.wall_time)
        self._print('[ SIMULATION SUMMARY ]')
        self._print('[t/s] %8.2e/%8.2e [S] %5d [P] %5d [C] %5d [T] %5d [CFL] %1d'
                    % (self.t_sim, st.t, n_tsteps, numitP, numitC, numitT, num_CFL))
        self._print('[Run time]', round(self.wall_time, 2), '(s)')
        return self.trajectory

    def _chemistry(self, chem_list, dt_chem):
        """tse.py:1384-1418: vectorised or per-cell chemistry callbacks."""
        c, st = self.config, self.st
        if hasattr(c, 'chemical_reactions_all'):
            st.C, st.P, st.T, st.rho_f, st.phi, st.Kx, st.Ky, st.Kz = \
                c.chemical_reactions_all(st.C, st.P, st.T, st.rho_f, st.phi,
                                         st.Kx, st.Ky, st.Kz, dt_chem)
            return
        from joblib import Parallel, delayed
        results = Parallel(n_jobs=c.parallel_chemistry)(
            delayed(c.chemical_reactions)(
                st.C[i, j, k], st.P[i, j, k], st.T[i, j, k], st.rho_f[i, j, k],
                st.phi[i, j, k], st.Kx[i, j, k], st.Ky[i, j, k], st.Kz[i, j, k],
                dt_chem, i, j, k) for i, j, k in chem_list)
        for num, l in enumerate(chem_list):
            r = results[num]
            st.C[l], st.P[l], st.T[l], st.rho_f[l], st.phi[l], st.Kx[l], \
                st.Ky[l], st.Kz[l] = r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]


def make_hip_backend(sim: Simulation, p_form: str = "faithful64", device: int = 0,
                     c_form: str = "faithful64", t_form: str = "faithful64"):
    """GPU backend for a Simulation (frozen globals read from its state).
    Forms as HipSolvers: faithful64 is bit-identical to the reference;
    plain64 / mixed read the config's CONV_P / CONV_C / CONV_T as relative
    residuals."""
    from .solvers import HipSolvers
    st, c = sim.st, sim.config
    return HipSolvers(st.nx, st.ny, st.nz, st.dx, st.dy, st.dz, st.g,
                      st.beta_deg, frozen_source=st, maxit_p=sim.MAXIT_P,
                      maxit_c=sim.MAXIT_C, maxit_t=sim.MAXIT_T,
                      dt_max=c.dt_max, c_imblflo=c.C_IMBLFLO,
                      t_imblflo=c.T_IMBLFLO, device=device,
                      auto_invalidate=False, verbose=sim.verbose, p_form=p_form,
                      c_form=c_form, t_form=t_form)


def main(argv=None):
    """``python -m transportse_amd.driver config.py`` (tse.py's CLI: the
    config module runs to its end time; output goes next to the config)."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m transportse_amd.driver",
                                 description="Run a TRANSPORTSE config on the GPU "
                                             "(faithful FP64 solvers by default).")
    ap.add_argument("config", help="config module (path or importable name)")
    ap.add_argument("--max-steps", type=int, default=None,
                    help="stop after this many time steps")
    ap.add_argument("--quiet", action="store_true", help="no per-step printout")
    forms = ("faithful64", "plain64", "mixed")
    for what in ("p", "c", "t"):
        ap.add_argument(f"--{what}-form", choices=forms, default="faithful64",
                        help=f"compute_{what.upper()} form (plain64 / mixed read "
                             f"CONV_{what.upper()} as a relative residual)")
    args = ap.parse_args(sys.argv[1:] if argv is None else argv[1:])
    backend = functools.partial(make_hip_backend, p_form=args.p_form,
                                c_form=args.c_form, t_form=args.t_form)
    sim = Simulation(args.config, backend, verbose=not args.quiet,
                     max_steps=args.max_steps)
    try:
        sim.run()
    except ConvergenceError as e:
        print(e)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
