"""Minimal pure-Python stand-in for the ``numba`` package.

TRANSPORTSE config modules decorate their callbacks with ``@njit`` and import
``njit``/``prange`` (e.g. ``from numba import njit``). Where numba is not
installed, putting this package on ``sys.path`` lets such a config import
unchanged: ``njit`` returns the function itself, ``prange`` is ``range``.
Python float arithmetic is IEEE binary64 with the same operand order, so a
callback evaluates to the same values as under Numba with fastmath=False.
"""

__version__ = "0.0+stub"
STUB = True


def njit(*args, **kwargs):
    if len(args) == 1 and callable(args[0]) and not kwargs:
        return args[0]
    if args and callable(args[0]):
        return args[0]
    return lambda f: f


jit = njit
prange = range


def set_num_threads(n):
    return None


def get_num_threads():
    return 1


def threading_layer():
    return "stub"


class _Config:
    THREADING_LAYER = "stub"
    NUMBA_NUM_THREADS = 1


config = _Config()

from . import extending, types  # noqa: E402,F401
