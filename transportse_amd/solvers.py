"""Drop-in solvers with the reference's call signatures, run on the GPU.

``HipSolvers`` exposes ``compute_P``, ``compute_q``, ``flow_error``,
``compute_C``, ``compute_T`` and ``l2_norm`` with the argument lists, default
values and return tuples of tse.py:

* compute_P  (tse.py:647-651 -> 742): ``(t_P, dt_P, P, Pn, converged, num_it)``
* compute_q  (tse.py:746 -> 780):     ``(qx, qy, qz)`` written in place
* flow_error (tse.py:410 -> 446):     a fresh ``e``
* compute_C  (tse.py:785-787 -> 916): ``(t_C, dt_C, C, converged, num_it)``
* compute_T  (tse.py:921-923 -> 1081): ``(t_T, dt_T, T, converged, num_it)``

Arrays are numpy float64 in the reference layout; species arguments may be
strided views. Results are written into the passed arrays (P, C, T, q) and
also returned, so a caller that rebinds (``P = compute_P(...)[2]``) and one
that keeps its own reference both see the solution.

Numba freezes module globals into a kernel at its first call (tse.py
reads ``c_r`` in compute_P, ``lf_l``/``lf_t`` in compute_C and ``l_r``,
``cp_r``, ``rho_r``, ``lt_l``, ``lt_t`` in compute_T as globals). The
solvers reproduce that: they read those arrays from ``frozen_source`` at the
first call of the function that uses them and keep the device copy.

Forms (``p_form``, ``c_form``, ``t_form``): ``"faithful64"`` runs the
reference's iteration operation for operation (bit-identical); ``"plain64"``
and ``"mixed"`` solve the fixed point of the same update to the relative
residual given as P_conv / C_conv / T_conv (FP64 sweeps, or FP32 sweeps in
FP64 defect correction), with the reference's accept / grow / halve / reset
time-step control around each attempt and max_it bounding its sweeps.

Deferred downloads (``lazy = True``, off by default; the SIA driver turns
it on): the arrays a solver writes (P and the fresh Pn of compute_P, qx, qy,
qz, the fresh e of flow_error, C, T, the Smolarkiewicz velocities) keep
their values on the device, and the host arrays are filled only by
``sync`` -- which the caller invokes before host code reads them -- or when
the device buffer holding them is about to be reused.  ``copy`` copies such
an array on the device, ``nonzero_sum`` answers ``numpy.sum(a) != 0`` for
it.  Passing a pending array back to a solver uses the device copy.

Device residency: inputs are cached on the device per argument role. A
cached copy is reused while the same host array (same buffer, strides and
shape) is passed and no host mutation was signalled; ``invalidate()``
signals one. With ``auto_invalidate=True`` (the default, for stand-alone
use) every call re-uploads its inputs; the SIA driver turns it off and
invalidates after every host callback.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from .device import HipContext

_FORMS = {"faithful64": _lib.FORM_FAITHFUL64, "plain64": _lib.FORM_PLAIN64,
          "mixed": _lib.FORM_MIXED}


def _key(a: np.ndarray):
    ai = a.__array_interface__
    return (ai["data"][0], a.strides, a.shape, a.dtype.str)


class _ViewRef:
    """A weak reference to a strided view: the owning array weakly, plus the
    view's offset, shape and strides (callers make a new view object at
    every use, e.g. a species slice C[:, :, :, n], and holding the view would
    keep its owner alive)."""
    __slots__ = ("owner", "offset", "shape", "strides")

    def __init__(self, owner, a):
        self.owner = weakref.ref(owner)
        self.offset = (a.__array_interface__["data"][0]
                       - owner.__array_interface__["data"][0])
        self.shape, self.strides = a.shape, a.strides

    def __call__(self):
        o = self.owner()
        if o is None:
            return None
        flat = o.reshape(-1)
        return np.lib.stride_tricks.as_strided(flat[self.offset // flat.itemsize:],
                                               self.shape, self.strides)


def _holder(a: np.ndarray):
    """What the cache and the pending list keep of a host array: a weak
    reference (they must not keep a caller's array alive); for a view of a
    C-contiguous ndarray, a weak reference to that owner; any other view is
    held itself."""
    if a.base is None:
        return weakref.ref(a)
    o = a.base
    while isinstance(o, np.ndarray) and isinstance(o.base, np.ndarray):
        o = o.base
    if isinstance(o, np.ndarray) and o.base is None and o.flags["C_CONTIGUOUS"] \
            and o.dtype == a.dtype:
        return _ViewRef(o, a)
    return a


def _held(h):
    """The host array of a cache or pending entry (None if it is gone)."""
    if h is None or isinstance(h, np.ndarray):
        return h
    return h()


class HipSolvers:
    """GPU implementation of the reference's solver calls for one grid."""

    def __init__(self, nx, ny, nz, dx, dy, dz, g, beta_deg, frozen_source=None,
                 maxit_p=20000, maxit_c=100, maxit_t=100, dt_max=1e99,
                 c_imblflo=False, t_imblflo=False, device=0,
                 auto_invalidate=True, verbose=True, p_form="faithful64",
                 c_form="faithful64", t_form="faithful64"):
        self.hip = HipContext(nx, ny, nz, dx, dy, dz, g, beta_deg, device)
        self.lib = self.hip.lib
        self.shape = self.hip.shape
        self.frozen_source = frozen_source
        self.maxit_p, self.maxit_c, self.maxit_t = maxit_p, maxit_c, maxit_t
        self.dt_max = dt_max
        self.c_imblflo, self.t_imblflo = bool(c_imblflo), bool(t_imblflo)
        self.auto_invalidate = auto_invalidate
        self.verbose = verbose
        for name, form in (("p_form", p_form), ("c_form", c_form),
                           ("t_form", t_form)):
            if form not in _FORMS:
                raise ValueError(f"{name} must be one of {sorted(_FORMS)}")
        # faithful64: the reference's iteration, bit-identical; plain64 /
        # mixed: the same update's fixed point to the relative residual
        # P_conv / C_conv / T_conv, with the reference's dt control
        self.p_form, self.c_form, self.t_form = p_form, c_form, t_form
        self._gen = 0
        self._cache = {}   # role -> [key, gen, tensor, host_ref]
        self._frozen = {}  # name -> tensor
        self._pe = {}      # "PeC"/"PeT" -> device array in the host layout
        # peclet_C / peclet_T write PeC / PeT back into the caller's arrays;
        # a driver that only reads the returned maxima turns this off and the
        # device copy stays the only one
        self.pe_writeback = True
        self.last = None
        self.stats = {"uploads": 0, "downloads": 0, "device_copies": 0}
        self.lazy = False
        # key -> [holder, tensor]: host arrays whose values live on the
        # device only (weakly held, see _holder)
        self._pending = {}

    # residency -------------------------------------------------------------------
    def invalidate(self, arrays=None):
        """Signal that host code may have changed input arrays: all of them
        (arrays None), or those sharing memory with any of `arrays`.  Host
        code that writes an array first syncs it; its pending device value,
        if any, is dropped here (the host now holds the value)."""
        if arrays is None:
            self._gen += 1
            self._pending.clear()
            return
        arrays = [np.asarray(a) for a in arrays]
        for ent in self._cache.values():
            ref = _held(ent[3])
            if ref is not None and any(np.may_share_memory(ref, a) for a in arrays):
                ent[1] = -1
        for pid, a, _, _ in self._pending_items():
            if any(np.may_share_memory(a, b) for b in arrays):
                del self._pending[pid]

    # deferred downloads ------------------------------------------------------------
    def _pending_items(self):
        """Live pending entries as (key, host array, tensor, key)."""
        out = []
        for pk, (h, t) in list(self._pending.items()):
            a = _held(h)
            if a is None:
                del self._pending[pk]
                continue
            out.append((pk, a, t, pk))
        return out

    def _pend(self, host, tensor):
        """Record that `host`'s value is `tensor`'s (not yet downloaded)."""
        a = np.asarray(host)
        self._pending[_key(a)] = [_holder(a), tensor]

    def _pending_exact(self, a):
        k = _key(a)
        ent = self._pending.get(k)
        if ent is None:
            return None
        if _held(ent[0]) is None:
            del self._pending[k]
            return None
        return k, ent[1]

    def sync(self, arrays=None):
        """Fill the host arrays whose values are still on the device: all of
        them (arrays None), or those sharing memory with any of `arrays`."""
        sel = None if arrays is None else [np.asarray(a) for a in arrays]
        for pid, a, t, _ in self._pending_items():
            if sel is not None and not any(np.may_share_memory(a, b) for b in sel):
                continue
            del self._pending[pid]
            self.hip.download(t, out=a)
            self.stats["downloads"] += 1

    def _release(self, tensor, keep=None):
        """Before `tensor` is overwritten: download the pending host arrays
        it holds (except `keep`, whose value the write produces)."""
        kk = None if keep is None else _key(np.asarray(keep))
        for pid, a, t, k in self._pending_items():
            if t is tensor and k != kk:
                del self._pending[pid]
                self.hip.download(t, out=a)
                self.stats["downloads"] += 1

    def _out(self, role, host, tensor):
        """A solver wrote `tensor`, the value of host array `host` (None: a
        fresh array): download it, or with lazy leave it pending."""
        if host is None:
            host = np.empty(self.shape, dtype=np.float64)
        if self.lazy:
            self._pend(host, tensor)
        else:
            self.hip.download(tensor, out=host)
            self.stats["downloads"] += 1
        self._mark(role, host, tensor)
        return host

    def copy(self, arr):
        """``arr.copy()``; where arr's value (or, for a 4-D array, a species
        slice's value) is on the device -- pending or a current input copy --
        the copy is made there and the new array is pending."""
        a = np.asarray(arr)
        if not self.lazy:
            return a.copy()
        if a.ndim == 4:
            new = np.empty_like(a)
            for s_ in range(a.shape[-1]):
                src = self._device_value(a[..., s_])
                if src is None:
                    new[..., s_] = a[..., s_]
                else:
                    self._pend(new[..., s_], self._clone(src))
            return new
        src = self._device_value(a)
        if src is None:
            return a.copy()
        new = np.empty_like(a)
        self._pend(new, self._clone(src))
        return new

    def _device_value(self, a):
        """A device tensor holding a's current value, or None."""
        ex = self._pending_exact(a)
        if ex is not None:
            return ex[1]
        self.sync([a])     # a partly pending through another view
        k = _key(a)
        for ent in self._cache.values():
            if ent[0] == k and ent[1] == self._gen and _held(ent[3]) is not None:
                return ent[2]
        return None

    def _clone(self, t):
        c = self.hip.empty()
        c.copy_(t)
        self.stats["device_copies"] += 1
        return c

    def nonzero_sum(self, arr) -> bool:
        """``numpy.sum(arr) != 0.0`` (tse.py:1300): for an array whose value
        is on the device (pending, or a current input copy) from the device
        values when they decide it (all zero: False; no negative and some
        positive value: True), else from the host sum."""
        a = np.asarray(arr)
        t = self._device_value(a) if self.lazy else None
        if t is not None:
            lo, hi = (float(v) for v in t.aminmax())
            if lo == 0.0 and hi == 0.0:
                return False
            if lo >= 0.0 and hi > 0.0:
                return True
            self.sync([a])
        return bool(np.sum(a) != 0.0)

    def _dev(self, role: str, arr):
        a = np.asarray(arr)
        k = _key(a)
        ent = self._cache.get(role)
        ex = self._pending_exact(a) if self._pending else None
        # a cache hit needs the same array, unchanged; a pending array's
        # value is its pending tensor (a new array may reuse a freed one's
        # address, and with it the key of an old cache entry)
        if ent is not None and ent[0] == k and ent[1] == self._gen and \
                _held(ent[3]) is not None and (ex is None or ex[1] is ent[2]):
            return ent[2]
        out = ent[2] if ent is not None else None
        if out is not None:
            self._release(out, keep=a if ex is not None and ex[1] is out else None)
        if ex is not None and ex[1] is out:
            pass                 # the role's buffer holds the value already
        elif ex is not None:     # the value is on the device already
            t = out if out is not None else self.hip.empty()
            t.copy_(ex[1])
            self.stats["device_copies"] += 1
        else:
            self.sync([a])       # a view of it may be pending
            t = self.hip.upload(a, out=out)
            self.stats["uploads"] += 1
        self._cache[role] = [k, self._gen, t, _holder(a)]
        return t

    def _raw(self, role: str, arr):
        """Device copy of a host array in its own C-order layout, cached
        like ``_dev``."""
        import torch
        a = np.asarray(arr)
        k = _key(a)
        ent = self._cache.get(role)
        if ent is not None and ent[0] == k and ent[1] == self._gen and \
                _held(ent[3]) is not None:
            return ent[2]
        self.sync([a])
        t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(
            self.hip.device)
        self.stats["uploads"] += 1
        self._cache[role] = [k, self._gen, t, _holder(a)]
        return t

    def _pe_buffer(self, name: str, host):
        """The Peclet array's device copy: from the host array at the first
        call, device-owned afterwards (cells with B != 0 keep their values)."""
        import torch
        t = self._pe.get(name)
        if t is None:
            h = np.asarray(host, dtype=np.float64)
            if h.shape != self.shape + (3,):
                raise ValueError(f"{name} must have shape {self.shape + (3,)}")
            t = torch.from_numpy(np.ascontiguousarray(h)).to(self.hip.device)
            self._pe[name] = t
        return t

    def _mark(self, role: str, arr, tensor):
        """Record that `tensor` now equals host `arr` (after a download)."""
        a = np.asarray(arr)
        self._cache[role] = [_key(a), self._gen, tensor, _holder(a)]

    def _scratch(self, role: str):
        ent = self._cache.get(role)
        if ent is None:
            t = self.hip.empty()
            self._cache[role] = [None, -1, t, None]
            return t
        return ent[2]

    def _frozen_field(self, name: str):
        if name not in self._frozen:
            src = getattr(self.frozen_source, name)
            self._frozen[name] = self.hip.upload(np.asarray(src, dtype=np.float64))
        return self._frozen[name]

    def set_frozen(self, **arrays):
        for k, v in arrays.items():
            self._frozen[k] = self.hip.upload(np.asarray(v, dtype=np.float64))

    def _begin(self):
        if self.auto_invalidate:
            self.sync()
            self.invalidate()
        self.hip.sync_stream()

    # solvers -----------------------------------------------------------------------
    def compute_P(self, P, c_f, nue, rho_f, phi, Kx, Ky, Kz, qx, qy, qz, dt_P,
                  BP, WP, SOR=1.0, max_it=None, inc_it=1.05, max_dt=None,
                  min_dt=1e-8, P_conv=1e-10):
        self._begin()
        max_it = self.maxit_p if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        d = {n: self._dev(n, a) for n, a in (
            ("P", P), ("c_f", c_f), ("nue", nue), ("rho_f", rho_f),
            ("phi", phi), ("Kx", Kx), ("Ky", Ky), ("Kz", Kz), ("BP", BP),
            ("WP", WP))}
        d["Pn"] = self._scratch("Pn")
        d["c_r"] = self._frozen_field("c_r")
        self._release(d["P"], keep=P)
        self._release(d["Pn"])
        f = _lib.PFields(**{n: d[n].data_ptr() for n in _lib.P_FIELDS})
        prm = _lib.SorParams(dt=float(dt_P), sor=float(SOR), omega=1.0,
                             inc_it=float(inc_it), max_dt=float(max_dt),
                             min_dt=float(min_dt), conv=float(P_conv),
                             max_it=int(max_it), form=_FORMS[self.p_form])
        res = _lib.Result()
        _lib.check(self.lib.tse_compute_p(self.hip.ctx, ctypes.byref(f),
                                          ctypes.byref(prm), ctypes.byref(res)),
                   "tse_compute_p")
        self._report_halvings("P", res, dt_P)
        self._out("P", P, d["P"])
        Pn = self._out("Pn", None, d["Pn"])
        self.last = res
        return res.t, res.dt, P, Pn, bool(res.converged), int(res.num_it)

    def compute_q(self, P, BP, Kx, Ky, Kz, qx, qy, qz, nue, rho_f):
        self._begin()
        d = {n: self._dev(n, a) for n, a in (
            ("P", P), ("BP", BP), ("Kx", Kx), ("Ky", Ky), ("Kz", Kz),
            ("qx", qx), ("qy", qy), ("qz", qz), ("nue", nue),
            ("rho_f", rho_f))}
        for n, a in (("qx", qx), ("qy", qy), ("qz", qz)):
            self._release(d[n], keep=a)
        _lib.check(self.lib.tse_compute_q(
            self.hip.ctx, *[d[n].data_ptr() for n in
                            ("P", "BP", "Kx", "Ky", "Kz", "qx", "qy", "qz",
                             "nue", "rho_f")]), "tse_compute_q")
        for n, a in (("qx", qx), ("qy", qy), ("qz", qz)):
            self._out(n, a, d[n])
        return qx, qy, qz

    def flow_error(self, P, Pn, qx, qy, qz, WP, phi, c_f, c_r, rho_f, dt_P):
        self._begin()
        d = {n: self._dev(n, a) for n, a in (
            ("P", P), ("Pn", Pn), ("qx", qx), ("qy", qy), ("qz", qz),
            ("WP", WP), ("phi", phi), ("c_f", c_f), ("c_r_live", c_r),
            ("rho_f", rho_f))}
        e_dev = self._scratch("e")
        self._release(e_dev)
        _lib.check(self.lib.tse_flow_error(
            self.hip.ctx, *[d[n].data_ptr() for n in
                            ("P", "Pn", "qx", "qy", "qz", "WP", "phi", "c_f",
                             "c_r_live", "rho_f")], float(dt_P),
            e_dev.data_ptr()), "tse_flow_error")
        return self._out("e", None, e_dev)

    def compute_C(self, C, D_fx, D_fy, D_fz, phi, rho_f, qx, qy, qz, dt_C, BC,
                  WC, RC, e, omega=1, SOR=1.0, max_it=None, inc_it=1.05,
                  max_dt=None, min_dt=1e-8, C_conv=1e-10):
        self._begin()
        max_it = self.maxit_c if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        d = {n: self._dev(n, a) for n, a in (
            ("C", C), ("D_fx", D_fx), ("D_fy", D_fy), ("D_fz", D_fz),
            ("phi", phi), ("rho_f", rho_f), ("qx", qx), ("qy", qy),
            ("qz", qz), ("BC", BC), ("WC", WC), ("RC", RC), ("e", e))}
        d["Cn"] = self._scratch("Cn")
        self._release(d["C"], keep=C)
        d["lf_l"] = self._frozen_field("lf_l")
        d["lf_t"] = self._frozen_field("lf_t")
        f = _lib.CFields(**{n: d[n].data_ptr() for n in _lib.C_FIELDS},
                         imblflo=int(self.c_imblflo))
        prm = _lib.SorParams(dt=float(dt_C), sor=float(SOR), omega=float(omega),
                             inc_it=float(inc_it), max_dt=float(max_dt),
                             min_dt=float(min_dt), conv=float(C_conv),
                             max_it=int(max_it), form=_FORMS[self.c_form])
        res = _lib.Result()
        _lib.check(self.lib.tse_compute_c(self.hip.ctx, ctypes.byref(f),
                                          ctypes.byref(prm), ctypes.byref(res)),
                   "tse_compute_c")
        self._report_halvings("C", res, dt_C)
        self._out("C", C, d["C"])
        self.last = res
        return res.t, res.dt, C, bool(res.converged), int(res.num_it)

    def compute_T(self, T, l_f, cp_f, rho_f, phi, qx, qy, qz, dt_T, BT, WT, e,
                  omega=1, SOR=1.0, max_it=None, inc_it=1.05, max_dt=None,
                  min_dt=1e-8, T_conv=1e-10):
        self._begin()
        max_it = self.maxit_t if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        d = {n: self._dev(n, a) for n, a in (
            ("T", T), ("l_f", l_f), ("cp_f", cp_f), ("rho_f", rho_f),
            ("phi", phi), ("qx", qx), ("qy", qy), ("qz", qz), ("BT", BT),
            ("WT", WT), ("e", e))}
        d["Tn"] = self._scratch("Tn")
        self._release(d["T"], keep=T)
        for n in ("l_r", "cp_r", "rho_r", "lt_l", "lt_t"):
            d[n] = self._frozen_field(n)
        f = _lib.TFields(**{n: d[n].data_ptr() for n in _lib.T_FIELDS},
                         imblflo=int(self.t_imblflo))
        prm = _lib.SorParams(dt=float(dt_T), sor=float(SOR), omega=float(omega),
                             inc_it=float(inc_it), max_dt=float(max_dt),
                             min_dt=float(min_dt), conv=float(T_conv),
                             max_it=int(max_it), form=_FORMS[self.t_form])
        res = _lib.Result()
        _lib.check(self.lib.tse_compute_t(self.hip.ctx, ctypes.byref(f),
                                          ctypes.byref(prm), ctypes.byref(res)),
                   "tse_compute_t")
        self._report_halvings("T", res, dt_T)
        self._out("T", T, d["T"])
        self.last = res
        return res.t, res.dt, T, bool(res.converged), int(res.num_it)

    def smolarkiewicz(self, C, qx, qy, qz, vx, vy, vz, dt, eps=1e-15, Sc=1.08):
        """smolarkiewicz (tse.py:1083-1116): updates vx, vy, vz in place and
        returns them."""
        self._begin()
        d = {n: self._dev(n, a) for n, a in (
            ("smol_C", C), ("qx", qx), ("qy", qy), ("qz", qz),
            ("vx", vx), ("vy", vy), ("vz", vz))}
        for n, a in (("vx", vx), ("vy", vy), ("vz", vz)):
            self._release(d[n], keep=a)
        _lib.check(self.lib.tse_smolarkiewicz(
            self.hip.ctx, *[d[n].data_ptr() for n in
                            ("smol_C", "qx", "qy", "qz", "vx", "vy", "vz")],
            float(dt), float(eps), float(Sc)), "tse_smolarkiewicz")
        for n, a in (("vx", vx), ("vy", vy), ("vz", vz)):
            self._out(n, a, d[n])
        return vx, vy, vz

    # diagnostics of the SIA driver --------------------------------------------------
    def peclet_C(self, PeC, qx, qy, qz, D_fx, D_fy, D_fz):
        """peclet_C (tse.py:333-360) on the device; returns the maxima of
        PeC[..., 0..2].  BC is read from ``frozen_source`` at the first call
        (a module global of the reference kernel)."""
        self._begin()
        d = {n: self._dev(n, a) for n, a in (("qx", qx), ("qy", qy), ("qz", qz))}
        Ds = [np.asarray(D) for D in (D_fx, D_fy, D_fz)]
        nspec = Ds[0].shape[-1] if Ds[0].ndim == 4 else 1
        if any(D.shape != self.shape + ((nspec,) if Ds[0].ndim == 4 else ())
               for D in Ds):
            raise ValueError("D_fx, D_fy, D_fz must share one (nx, ny, nz[, S]) shape")
        Dd = [self._raw(f"pe_{n}", D) for n, D in zip(("D_fx", "D_fy", "D_fz"), Ds)]
        bc = self._frozen_field("BC")
        pe = self._pe_buffer("PeC", PeC)
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.tse_peclet_c(
            self.hip.ctx, d["qx"].data_ptr(), d["qy"].data_ptr(), d["qz"].data_ptr(),
            *[t.data_ptr() for t in Dd], int(nspec), bc.data_ptr(), pe.data_ptr(),
            out), "tse_peclet_c")
        if self.pe_writeback:
            PeC[...] = pe.cpu().numpy().reshape(PeC.shape)
            self.stats["downloads"] += 1
        return tuple(out)

    def peclet_T(self, PeT, qx, qy, qz, rho_f, cp_f, phi, l_r, l_f):
        """peclet_T (tse.py:364-383) on the device; returns the maxima of
        PeT[..., 0..2].  BT is read from ``frozen_source`` at the first call;
        l_r is the live argument."""
        self._begin()
        d = {n: self._dev(n, a) for n, a in (
            ("qx", qx), ("qy", qy), ("qz", qz), ("rho_f", rho_f), ("cp_f", cp_f),
            ("phi", phi), ("l_r_live", l_r), ("l_f", l_f))}
        bt = self._frozen_field("BT")
        pe = self._pe_buffer("PeT", PeT)
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.tse_peclet_t(
            self.hip.ctx, *[d[n].data_ptr() for n in
                            ("qx", "qy", "qz", "rho_f", "cp_f", "phi", "l_r_live",
                             "l_f")], bt.data_ptr(), pe.data_ptr(), out),
            "tse_peclet_t")
        if self.pe_writeback:
            PeT[...] = pe.cpu().numpy().reshape(PeT.shape)
            self.stats["downloads"] += 1
        return tuple(out)

    def cfl_maxima(self, qx, qy, qz, phi):
        """The three maxima of the SIA driver's CFL limit (tse.py:1241-1246,
        with its reshapes); dt_CFL = 0.9 / their sum."""
        self._begin()
        d = {n: self._dev(n, a) for n, a in (
            ("qx", qx), ("qy", qy), ("qz", qz), ("phi", phi))}
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.tse_cfl_maxima(
            self.hip.ctx, *[d[n].data_ptr() for n in ("qx", "qy", "qz", "phi")], out),
            "tse_cfl_maxima")
        return tuple(out)

    def l2_norm(self, New, Old, Process=""):
        """l2_norm (tse.py:273-281) on host arrays, Numba's sequential sums."""
        a = np.ascontiguousarray(New, dtype=np.float64)
        b = np.ascontiguousarray(Old, dtype=np.float64)
        flag = ctypes.c_int32(0)
        dp = ctypes.POINTER(ctypes.c_double)
        v = self.lib.tse_host_l2_norm_seq(a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                          a.size, ctypes.byref(flag))
        if flag.value:
            print("ISNAN:", Process)
            return 1
        return v

    def _report_halvings(self, what, res, dt0):
        if self.verbose and res.num_halvings:
            print(f" [{what}] {res.num_halvings} dt halving(s): {dt0} -> {res.dt}"
                  f" (s) after {res.num_it} iterations")
