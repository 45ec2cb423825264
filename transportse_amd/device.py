"""Device context: one HIP context, one grid, torch tensors as field buffers.

Fields live on the GPU in the checkerboard-split layout of
include/tse_hip.h (2*S float64 per field). Host numpy arrays in the reference
layout (C order (nx, ny, nz), or strided species views such as
``C[:, :, :, n]``) are copied into a flat staging tensor and permuted by the
``tse_to_split`` kernel; results go back through ``tse_from_split``.
PyTorch supplies device memory and the stream only.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class HipContext:
    """Owns a ``tse_ctx`` bound to one grid on one device."""

    def __init__(self, nx, ny, nz, dx, dy, dz, g, beta_deg, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("transportse_amd needs a ROCm GPU (torch.cuda is"
                               " not available); there is no CPU fallback")
        self.lib = _lib.load()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.shape = (self.nx, self.ny, self.nz)
        self.dx = np.ascontiguousarray(dx, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        self.dz = np.ascontiguousarray(dz, dtype=np.float64)
        if self.dx.shape != (self.nx,) or self.dy.shape != (self.ny,) \
                or self.dz.shape != (self.nz,):
            raise ValueError("dx/dy/dz must have lengths nx/ny/nz")
        self.device_index = int(device)
        self.device = torch.device("cuda", self.device_index)
        beta = math.radians(float(beta_deg))
        self.g = float(g)
        self.sin_beta, self.cos_beta = math.sin(beta), math.cos(beta)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib.tse_ctx_create(ctypes.byref(handle), self.device_index,
                                           ctypes.c_void_p(stream)), "tse_ctx_create")
        self.ctx = handle
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(self.lib.tse_set_grid(
            self.ctx, self.nx, self.ny, self.nz,
            self.dx.ctypes.data_as(dp), self.dy.ctypes.data_as(dp),
            self.dz.ctypes.data_as(dp), self.g, self.sin_beta, self.cos_beta),
            "tse_set_grid")
        self.F = int(self.lib.tse_field_elems(self.ctx))
        self.nxh = (self.nx + 1) // 2
        self._staging = None
        self._up = [None, None]  # device staging of uploads, one per pinned buffer
        # two pinned host staging buffers (both directions), used in turn,
        # each with the event of the last copy that reads or writes it
        self._pins = [None, None]
        self._pin_ev = [None, None]
        self._pin_turn = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.tse_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # buffers -----------------------------------------------------------------
    def empty(self):
        torch = _torch()
        return torch.empty(self.F, dtype=torch.float64, device=self.device)

    def zeros(self):
        torch = _torch()
        return torch.zeros(self.F, dtype=torch.float64, device=self.device)

    def sync_stream(self):
        torch = _torch()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.lib.tse_ctx_set_stream(self.ctx, ctypes.c_void_p(stream))

    # host <-> device -------------------------------------------------------------
    def upload(self, arr: np.ndarray, out=None):
        """Copy a host (nx,ny,nz) float64 array (any strides) into a split field."""
        torch = _torch()
        a = np.asarray(arr)
        if a.shape != self.shape:
            raise ValueError(f"field shape {a.shape} != grid {self.shape}")
        if a.dtype != np.float64:
            a = a.astype(np.float64)
        # keep the element strides of the host view when its memory is one
        # contiguous block (e.g. a species slice of a C-order 4-D array)
        base, strides = _block_and_strides(a)
        n = base.size
        if out is None:
            out = self.empty()
        self.sync_stream()
        stream = torch.cuda.current_stream(self.device)
        turn, pin = self._pinned(n)
        # the device staging of this pinned buffer: the event waited for in
        # _pinned also covers the last tse_to_split that read it, on any stream
        up = self._up[turn]
        if up is None or up.numel() < n:
            up = self._up[turn] = torch.empty(n, dtype=torch.float64, device=self.device)
        pin_np = pin.numpy()
        # chunked: host copies into pinned memory run on the copy threads,
        # and each chunk's DMA starts as soon as its copy is done
        parts = _chunks(n, n)
        if len(parts) > 1:
            futs = [_copier().submit(np.copyto, pin_np[lo:hi], base[lo:hi])
                    for lo, hi in parts]
        else:
            np.copyto(pin_np[:n], base[:n])
            futs = [None]
        for f, (lo, hi) in zip(futs, parts):
            if f is not None:
                f.result()
            up[lo:hi].copy_(pin[lo:hi], non_blocking=True)
        _lib.check(self.lib.tse_to_split(self.ctx, out.data_ptr(), up.data_ptr(),
                                          *strides), "tse_to_split")
        ev = torch.cuda.Event()
        ev.record(stream)
        self._pin_ev[turn] = ev
        return out

    def _pinned(self, n: int):
        """The next pinned staging buffer of at least n doubles, once the
        last copy that used it has finished: (index, buffer)."""
        torch = _torch()
        turn = self._pin_turn
        self._pin_turn ^= 1
        if self._pin_ev[turn] is not None:
            self._pin_ev[turn].synchronize()
            self._pin_ev[turn] = None
        if self._pins[turn] is None or self._pins[turn].numel() < n:
            self._pins[turn] = torch.empty(n, dtype=torch.float64, pin_memory=True)
        return turn, self._pins[turn]

    def download(self, field, out: np.ndarray | None = None) -> np.ndarray:
        """Copy a split field into a host (nx,ny,nz) array (new, or `out`)."""
        torch = _torch()
        n = self.nx * self.ny * self.nz
        if self._staging is None or self._staging.numel() < n:
            self._staging = torch.empty(n, dtype=torch.float64, device=self.device)
        st = self._staging[:n]
        self.sync_stream()
        _lib.check(self.lib.tse_from_split(self.ctx, st.data_ptr(),
                                            self.ny * self.nz, self.nz, 1,
                                            field.data_ptr()), "tse_from_split")
        _, pin = self._pinned(n)
        stream = torch.cuda.current_stream(self.device)
        # chunks of whole i-planes: the DMA of one chunk into pinned memory
        # overlaps the host copy of the one before out of it
        plane = self.ny * self.nz
        parts = []
        for i0, i1 in _chunks(self.nx, n):
            pin[i0 * plane:i1 * plane].copy_(st[i0 * plane:i1 * plane], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            parts.append((i0, i1, ev))
        host = pin[:n].numpy().reshape(self.shape)
        if out is None:
            out = np.empty(self.shape, dtype=np.float64)
        if len(parts) == 1:
            parts[0][2].synchronize()
            out[...] = host
            return out
        # each chunk is copied out on the copy threads once its DMA is done
        futs = []
        for i0, i1, ev in parts:
            ev.synchronize()
            futs.append(_copier().submit(np.copyto, out[i0:i1], host[i0:i1]))
        for f in futs:
            f.result()
        return out

    def synchronize(self):
        _lib.check(self.lib.tse_ctx_synchronize(self.ctx), "tse_ctx_synchronize")


_COPY_POOL = None


def _copier():
    """Threads for the host-side copies of the transfers (numpy releases the
    GIL while it copies)."""
    global _COPY_POOL
    if _COPY_POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _COPY_POOL = ThreadPoolExecutor(max_workers=4)
    return _COPY_POOL


def pcopyto(dst: np.ndarray, src) -> None:
    """``dst[...] = src`` for large arrays, in slabs along the first axis on
    the copy threads (src broadcast to dst's shape)."""
    src = np.broadcast_to(np.asarray(src), dst.shape)
    if dst.ndim == 0 or dst.shape[0] < 4 or dst.size < (1 << 21):
        dst[...] = src
        return
    k = min(4, dst.shape[0])
    step = -(-dst.shape[0] // k)
    futs = [_copier().submit(np.copyto, dst[i:i + step], src[i:i + step])
            for i in range(0, dst.shape[0], step)]
    for f in futs:
        f.result()


def pcopy(a: np.ndarray) -> np.ndarray:
    """``a.copy()`` (C order) through pcopyto."""
    out = np.empty(a.shape, dtype=a.dtype)
    pcopyto(out, a)
    return out


def _chunks(m: int, n: int):
    """Split range(m) into up to 8 consecutive parts, each standing for at
    least 2^20 of the n elements moved (one part for small copies)."""
    k = max(1, min(8, n >> 20, m))
    step = -(-m // k)
    return [(lo, min(lo + step, m)) for lo in range(0, m, step)]


def _block_and_strides(a: np.ndarray):
    """Return (contiguous 1-D buffer, element strides (si, sj, sk)).

    For a C-contiguous array this is the array itself. For a strided view
    whose elements lie in one C-contiguous parent block (a species slice),
    the smallest covering slice of that parent is used so the device gather
    reads the strided layout directly; anything else is copied contiguous.
    """
    if a.flags["C_CONTIGUOUS"]:
        return a.reshape(-1), (a.shape[1] * a.shape[2], a.shape[2], 1)
    # walk to the outermost ndarray; a chain that reaches another kind of
    # buffer (memmap file, bytes, mmap) stops there and takes the copy path
    base = a.base
    while isinstance(base, np.ndarray) and isinstance(base.base, np.ndarray):
        base = base.base
    if (isinstance(base, np.ndarray) and base.flags["C_CONTIGUOUS"]
            and base.dtype == np.float64 and all(s > 0 for s in a.strides)
            and all(s % 8 == 0 for s in a.strides)):
        start = (a.__array_interface__["data"][0]
                 - base.__array_interface__["data"][0])
        if start >= 0 and start % 8 == 0:
            flat = base.reshape(-1)
            off = start // 8
            si, sj, sk = (s // 8 for s in a.strides)
            last = off + (a.shape[0] - 1) * si + (a.shape[1] - 1) * sj \
                + (a.shape[2] - 1) * sk
            if last < flat.size:
                return flat[off:last + 1], (si, sj, sk)
    c = np.ascontiguousarray(a)
    return c.reshape(-1), (a.shape[1] * a.shape[2], a.shape[2], 1)
