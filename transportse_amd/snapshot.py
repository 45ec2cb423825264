"""Output snapshots: the 34-tuple pickle of tse.py:312-329 written off the
time loop.

``PickleWriter`` dumps (joblib, compress=3, like the reference) on one
background thread in submission order; the caller hands it a snapshot that
no later step modifies.  For device-resident scenarios the snapshot is taken
on the GPU: every field is converted to the reference layout into a device
staging buffer on the compute stream, copied to pinned host memory on a
side stream, and the writer thread waits for that copy before it builds the
tuple -- the time loop only pays for the device-side conversion.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor


class PickleWriter:
    def __init__(self, compress: int = 3):
        self.compress = compress
        self._ex = None
        self._futs = []

    def submit(self, path: str, obj, wait=None) -> None:
        """Queue a dump of `obj` (or of `obj()` if callable) to `path`, after
        `wait()` returns when given."""
        if self._ex is None:
            self._ex = ThreadPoolExecutor(max_workers=1)
        self._futs.append(self._ex.submit(self._write, path, obj, wait))

    def _write(self, path, obj, wait):
        from joblib import dump
        if wait is not None:
            wait()
        if callable(obj):
            obj = obj()
        with open(path, "wb") as f:
            dump(obj, f, compress=self.compress)

    def flush(self) -> None:
        """Wait for every queued dump; re-raises a failed write."""
        futs, self._futs = self._futs, []
        for f in futs:
            f.result()

    def close(self) -> None:
        self.flush()
        if self._ex is not None:
            self._ex.shutdown()
            self._ex = None
