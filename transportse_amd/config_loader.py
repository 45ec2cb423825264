"""Import a TRANSPORTSE config module by path (tse.py:41-45).

The reference appends the config's directory to ``sys.path`` and imports the
module by name. Configs import ``njit``/``prange`` from numba; when numba is
not installed the pure-Python stand-in in ``numba_stub/`` is put on
``sys.path`` first, so such configs import unchanged.
"""
from __future__ import annotations

import importlib
import importlib.util
import os
import sys

STUB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "numba_stub")


def ensure_numba() -> bool:
    """Make ``import numba`` work; returns True if the stand-in is used."""
    if "numba" in sys.modules:
        return getattr(sys.modules["numba"], "STUB", False)
    if importlib.util.find_spec("numba") is not None:
        try:
            importlib.import_module("numba")
            return False
        except Exception:
            pass
    if STUB_DIR not in sys.path:
        sys.path.insert(0, STUB_DIR)
    importlib.import_module("numba")
    return True


def load_config(path: str, fresh: bool = True):
    """Import the config module at `path` the way tse.py does."""
    ensure_numba()
    path = os.path.abspath(path)
    cdir = os.path.dirname(path)
    name = os.path.splitext(os.path.basename(path))[0]
    if cdir not in sys.path:
        sys.path.append(cdir)
    if fresh and name in sys.modules:
        del sys.modules[name]
    return importlib.import_module(name)
