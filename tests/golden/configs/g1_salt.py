"""Golden case 1: variable-density flow + salt, slope 5 deg, stretched z."""
from numba import njit  # noqa: F401  (config contract: configs import njit)
from tiny_base import build

build(globals(), nx=7, ny=6, nz=5, seed=11, heat=False, nsteps=4)
