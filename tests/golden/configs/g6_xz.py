"""Golden case: x-z section (ny = 1), like the reference's 2-D cross-section examples: flow, salt and heat with slope and dispersion."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=10, ny=1, nz=8, seed=16, heat=True, species=1, sptrans=(0,), nsteps=3, dt0=3600.0, dt_max=20000.0)
