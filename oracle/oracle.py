"""CPU oracle for the TRANSPORTSE SOR / transport hot path.

TEST INFRASTRUCTURE ONLY. The parity tests compare the HIP path with this
module, and it is the timed CPU baseline. Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
does.

``Oracle`` exposes the reference's solver call signatures (tse.py:647-651,
746, 410, 785-787, 921-923) on numpy arrays in the reference layout
(C order, shape (nx, ny, nz)). The arithmetic lives in ``tse_oracle.c``.

Numba freezes module globals into each kernel at its first compilation
(i.e. first call). The oracle mirrors that: ``c_r`` is copied at the first
``compute_P`` call, ``lf_l``/``lf_t`` at the first ``compute_C`` call and
``l_r``/``cp_r``/``rho_r``/``lt_l``/``lt_t`` at the first ``compute_T`` call,
read from ``frozen_source`` (any object with those attributes, e.g. a config
module).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtse_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", _dp), ("dy", _dp), ("dz", _dp),
                ("g", ctypes.c_double), ("sin_beta", ctypes.c_double),
                ("cos_beta", ctypes.c_double)]


class _Result(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("converged", ctypes.c_int), ("num_it", ctypes.c_int),
                ("num_halvings", ctypes.c_int), ("last_diff", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile libtse_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE,
                       check=True)
    return _LIB_PATH


_libs = {}


def lib(path=None):
    """The oracle library (default: libtse_oracle.so next to this file)."""
    global _lib
    if path is not None and path != _LIB_PATH:
        if path not in _libs:
            _libs[path] = _bind(ctypes.CDLL(path))
        return _libs[path]
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    if True:
        D, I, G, R = ctypes.c_double, ctypes.c_int, ctypes.POINTER(_Grid), ctypes.POINTER(_Result)
        L.tse_or_harm.restype = D
        L.tse_or_harm.argtypes = [G, _dp, I, I, I, I, I, I, I]
        L.tse_or_amean.restype = D
        L.tse_or_amean.argtypes = [G, _dp, I, I, I, I, I, I, I]
        L.tse_or_alpha.restype = D
        L.tse_or_alpha.argtypes = [G, I, I, I, I, I, I, _dp, _dp, _dp, I, I, I]
        L.tse_or_beta_g.restype = D
        L.tse_or_beta_g.argtypes = [G, I, I, I, I, I, I, _dp, I]
        L.tse_or_l2_norm.restype = D
        L.tse_or_l2_norm.argtypes = [_dp, _dp, ctypes.c_size_t, ctypes.POINTER(I)]
        L.tse_or_dispersivity.restype = None
        L.tse_or_dispersivity.argtypes = [G, _dp, _dp, _dp, _dp, _dp, _dp, I, I, I, _dp]
        L.tse_or_smolarkiewicz.restype = None
        L.tse_or_smolarkiewicz.argtypes = [G] + [_dp] * 7 + [D, D, D]
        L.tse_or_compute_p.restype = I
        L.tse_or_compute_p.argtypes = [G] + [_dp] * 12 + [D, D, I, D, D, D, D, R]
        L.tse_or_p_half_sweep.restype = None
        L.tse_or_p_half_sweep.argtypes = [G] + [_dp] * 12 + [D, D, I]
        L.tse_or_compute_q.restype = None
        L.tse_or_compute_q.argtypes = [G] + [_dp] * 10
        L.tse_or_flow_error.restype = None
        L.tse_or_flow_error.argtypes = [G] + [_dp] * 10 + [D, _dp]
        L.tse_or_compute_c.restype = I
        L.tse_or_compute_c.argtypes = [G] + [_dp] * 16 + [I, D, D, D, I, D, D, D, D, R]
        L.tse_or_compute_t.restype = I
        L.tse_or_compute_t.argtypes = [G] + [_dp] * 17 + [I, D, D, D, I, D, D, D, D, R]
        L.tse_or_set_threads.argtypes = [I]
        L.tse_or_set_progress.argtypes = [_PROGRESS, I]
        L.tse_or_set_progress.restype = None
        L.tse_or_get_threads.restype = I
    return L


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)


def _c(a) -> np.ndarray:
    """Contiguous float64 copy of a field (species views are strided)."""
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().tse_or_set_threads(int(n))


_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int)
_progress_ref = None


def set_progress(fn=None, every: int = 500) -> None:
    """Call fn(num_it) every `every` iterations of a compute_P call (for
    long runs whose callers must show that they are progressing); None
    turns it off."""
    global _progress_ref
    _progress_ref = _PROGRESS(fn) if fn is not None else _PROGRESS()
    lib().tse_or_set_progress(_progress_ref, int(every) if fn is not None else 0)


def get_threads() -> int:
    return int(lib().tse_or_get_threads())


class Oracle:
    """Reference-signature solvers backed by the C restatement."""

    def __init__(self, nx, ny, nz, dx, dy, dz, g, beta_deg, frozen_source=None,
                 maxit_p=20000, maxit_c=100, maxit_t=100, dt_max=1e99,
                 c_imblflo=False, t_imblflo=False, verbose=False,
                 lib_path=None):
        self._L = lib(lib_path)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = _c(dx)
        self.dy = _c(dy)
        self.dz = _c(dz)
        assert self.dx.shape == (self.nx,) and self.dy.shape == (self.ny,) \
            and self.dz.shape == (self.nz,)
        beta = math.radians(float(beta_deg))
        self.grid = _Grid(self.nx, self.ny, self.nz, _p(self.dx), _p(self.dy),
                          _p(self.dz), float(g), math.sin(beta), math.cos(beta))
        self.frozen_source = frozen_source
        self.maxit_p, self.maxit_c, self.maxit_t = maxit_p, maxit_c, maxit_t
        self.dt_max = dt_max
        self.c_imblflo, self.t_imblflo = bool(c_imblflo), bool(t_imblflo)
        self.verbose = verbose
        self._frozen = {}

    # frozen module globals ----------------------------------------------
    def _freeze(self, *names):
        out = []
        for n in names:
            if n not in self._frozen:
                src = getattr(self.frozen_source, n)
                self._frozen[n] = _c(src).copy()
            out.append(self._frozen[n])
        return out

    def set_frozen(self, **arrays):
        for k, v in arrays.items():
            self._frozen[k] = _c(v).copy()

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    # helpers -------------------------------------------------------------
    def harm(self, i1, j1, k1, i2, j2, k2, H, n):
        H = _c(H)
        return self._L.tse_or_harm(ctypes.byref(self.grid), _p(H), i1, j1, k1, i2, j2, k2, n)

    def amean(self, i1, j1, k1, i2, j2, k2, X, n):
        X = _c(X)
        return self._L.tse_or_amean(ctypes.byref(self.grid), _p(X), i1, j1, k1, i2, j2, k2, n)

    def alpha(self, i1, j1, k1, i2, j2, k2, nue, rho_f, K, d1, d2, n):
        nue, rho_f, K = _c(nue), _c(rho_f), _c(K)
        return self._L.tse_or_alpha(ctypes.byref(self.grid), i1, j1, k1, i2, j2, k2,
                                  _p(nue), _p(rho_f), _p(K), d1, d2, n)

    def beta_g(self, i1, j1, k1, i2, j2, k2, rho_f, d):
        rho_f = _c(rho_f)
        return self._L.tse_or_beta_g(ctypes.byref(self.grid), i1, j1, k1, i2, j2, k2,
                                   _p(rho_f), d)

    def l2_norm(self, New, Old, Process=""):
        """l2_norm (tse.py:273-281); prints ISNAN and returns 1 on NaN."""
        New, Old = _c(New), _c(Old)
        flag = ctypes.c_int(0)
        v = self._L.tse_or_l2_norm(_p(New), _p(Old), New.size, ctypes.byref(flag))
        if flag.value:
            print("ISNAN:", Process)
            return 1
        return v

    def dispersivity(self, qx, qy, qz, l_l, l_t, B, i, j, k):
        out = np.zeros(6)
        arrs = [_c(a) for a in (qx, qy, qz, l_l, l_t, B)]
        self._L.tse_or_dispersivity(ctypes.byref(self.grid), *[_p(a) for a in arrs],
                                  i, j, k, _p(out))
        return tuple(float(v) for v in out)

    def smolarkiewicz(self, C, qx, qy, qz, vx, vy, vz, dt, eps=1e-15, Sc=1.08):
        """tse.py:1083-1116; updates vx, vy, vz in place and returns them."""
        ins = [_c(a) for a in (C, qx, qy, qz)]
        outs = [_c(a) for a in (vx, vy, vz)]
        self._L.tse_or_smolarkiewicz(ctypes.byref(self.grid), *[_p(a) for a in ins],
                                     *[_p(a) for a in outs], float(dt), float(eps),
                                     float(Sc))
        for dst, src in zip((vx, vy, vz), outs):
            if dst is not src:
                dst[...] = src
        return vx, vy, vz

    # solvers -------------------------------------------------------------
    def compute_P(self, P, c_f, nue, rho_f, phi, Kx, Ky, Kz, qx, qy, qz, dt_P,
                  BP, WP, SOR=1.0, max_it=None, inc_it=1.05, max_dt=None,
                  min_dt=1e-8, P_conv=1e-10):
        max_it = self.maxit_p if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        (c_r,) = self._freeze("c_r")
        Pw = _c(P)
        Pn = np.empty_like(Pw)
        ins = [_c(a) for a in (c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP)]
        res = _Result()
        rc = self._L.tse_or_compute_p(ctypes.byref(self.grid), _p(Pw), _p(Pn),
                                    *[_p(a) for a in ins], _p(c_r), float(dt_P),
                                    float(SOR), int(max_it), float(inc_it),
                                    float(max_dt), float(min_dt), float(P_conv),
                                    ctypes.byref(res))
        if rc != 0:
            raise MemoryError("oracle compute_P failed")
        P[...] = Pw
        self.last = res
        return res.t, res.dt, P, Pn, bool(res.converged), res.num_it

    def p_half_sweep(self, P, Pn, c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP,
                     dt_P, SOR, colour):
        (c_r,) = self._freeze("c_r")
        Pw = _c(P)
        ins = [_c(a) for a in (Pn, c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP)]
        self._L.tse_or_p_half_sweep(ctypes.byref(self.grid), _p(Pw),
                                  *[_p(a) for a in ins], _p(c_r), float(dt_P),
                                  float(SOR), int(colour))
        P[...] = Pw
        return P

    def compute_q(self, P, BP, Kx, Ky, Kz, qx, qy, qz, nue, rho_f):
        outs = [_c(a) for a in (qx, qy, qz)]
        ins1 = [_c(a) for a in (P, BP, Kx, Ky, Kz)]
        ins2 = [_c(a) for a in (nue, rho_f)]
        self._L.tse_or_compute_q(ctypes.byref(self.grid), *[_p(a) for a in ins1],
                               *[_p(a) for a in outs], *[_p(a) for a in ins2])
        qx[...], qy[...], qz[...] = outs
        return qx, qy, qz

    # diagnostics of the SIA driver (numpy, as evaluated per cell by the
    # reference; maxima with numpy.max's NaN propagation) ------------------
    def peclet_C(self, PeC, qx, qy, qz, D_fx, D_fy, D_fz):
        """peclet_C (tse.py:333-360): PeC updated in place where BC == 0
        (BC frozen at the first call, a module global of the kernel); returns
        the maxima of PeC[..., 0..2]."""
        (BC,) = self._freeze("BC")
        act = BC == 0
        out = []
        for a, q, D, d in ((0, qx, D_fx, self.dx[:, None, None]),
                           (1, qy, D_fy, self.dy[None, :, None]),
                           (2, qz, D_fz, self.dz[None, None, :])):
            D = np.asarray(D, dtype=np.float64)
            mD = np.max(D, axis=-1) if D.ndim == 4 else D
            with np.errstate(divide="ignore", invalid="ignore"):
                val = np.where(mD != 0, np.abs(q) / mD * d, np.nan)
            PeC[..., a] = np.where(act, val, PeC[..., a])
            out.append(float(np.max(PeC[..., a])))
        return tuple(out)

    def peclet_T(self, PeT, qx, qy, qz, rho_f, cp_f, phi, l_r, l_f):
        """peclet_T (tse.py:364-383): PeT updated in place where BT == 0
        (BT frozen at the first call); returns the maxima of PeT[..., 0..2]."""
        (BT,) = self._freeze("BT")
        act = BT == 0
        den = (1 - phi) * l_r + phi * l_f
        out = []
        for a, q, d in ((0, qx, self.dx[:, None, None]), (1, qy, self.dy[None, :, None]),
                        (2, qz, self.dz[None, None, :])):
            with np.errstate(divide="ignore", invalid="ignore"):
                val = np.abs(q) * rho_f * cp_f / den * d
            PeT[..., a] = np.where(act, val, PeT[..., a])
            out.append(float(np.max(PeT[..., a])))
        return tuple(out)

    def cfl_maxima(self, qx, qy, qz, phi):
        """The three maxima of the CFL limit, tse.py:1241-1246 (with its
        reshapes); dt_CFL = 0.9 / sum."""
        nx, ny, nz = self.shape
        with np.errstate(divide="ignore", invalid="ignore"):
            return (float(np.max(np.reshape(np.abs(qx) / phi, (ny, nz, nx)) / self.dx)),
                    float(np.max(np.reshape(np.abs(qy) / phi, (nx, nz, ny)) / self.dy)),
                    float(np.max(np.abs(qz) / phi / self.dz)))

    def flow_error(self, P, Pn, qx, qy, qz, WP, phi, c_f, c_r, rho_f, dt_P):
        ins = [_c(a) for a in (P, Pn, qx, qy, qz, WP, phi, c_f, c_r, rho_f)]
        e = np.empty(self.shape)
        self._L.tse_or_flow_error(ctypes.byref(self.grid), *[_p(a) for a in ins],
                                float(dt_P), _p(e))
        return e

    def compute_C(self, C, D_fx, D_fy, D_fz, phi, rho_f, qx, qy, qz, dt_C, BC,
                  WC, RC, e, omega=1, SOR=1.0, max_it=None, inc_it=1.05,
                  max_dt=None, min_dt=1e-8, C_conv=1e-10):
        max_it = self.maxit_c if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        lf_l, lf_t = self._freeze("lf_l", "lf_t")
        Cw = _c(C)
        Cn = np.empty_like(Cw)
        ins = [_c(a) for a in (D_fx, D_fy, D_fz, phi, rho_f, qx, qy, qz, BC, WC, RC, e)]
        res = _Result()
        rc = self._L.tse_or_compute_c(ctypes.byref(self.grid), _p(Cw), _p(Cn),
                                    *[_p(a) for a in ins], _p(lf_l), _p(lf_t),
                                    int(self.c_imblflo), float(dt_C), float(omega),
                                    float(SOR), int(max_it), float(inc_it),
                                    float(max_dt), float(min_dt), float(C_conv),
                                    ctypes.byref(res))
        if rc != 0:
            raise MemoryError("oracle compute_C failed")
        C[...] = Cw
        self.last = res
        return res.t, res.dt, C, bool(res.converged), res.num_it

    def compute_T(self, T, l_f, cp_f, rho_f, phi, qx, qy, qz, dt_T, BT, WT, e,
                  omega=1, SOR=1.0, max_it=None, inc_it=1.05, max_dt=None,
                  min_dt=1e-8, T_conv=1e-10):
        max_it = self.maxit_t if max_it is None else max_it
        max_dt = self.dt_max if max_dt is None else max_dt
        l_r, cp_r, rho_r, lt_l, lt_t = self._freeze("l_r", "cp_r", "rho_r", "lt_l", "lt_t")
        Tw = _c(T)
        Tn = np.empty_like(Tw)
        ins = [_c(a) for a in (l_f, cp_f, rho_f, phi, qx, qy, qz, BT, WT, e)]
        res = _Result()
        rc = self._L.tse_or_compute_t(ctypes.byref(self.grid), _p(Tw), _p(Tn),
                                    *[_p(a) for a in ins], _p(l_r), _p(cp_r),
                                    _p(rho_r), _p(lt_l), _p(lt_t),
                                    int(self.t_imblflo), float(dt_T), float(omega),
                                    float(SOR), int(max_it), float(inc_it),
                                    float(max_dt), float(min_dt), float(T_conv),
                                    ctypes.byref(res))
        if rc != 0:
            raise MemoryError("oracle compute_T failed")
        T[...] = Tw
        self.last = res
        return res.t, res.dt, T, bool(res.converged), res.num_it
