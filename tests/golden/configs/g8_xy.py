"""Golden case: map-view plane (nz = 1), like the reference's 51 x 51 x 1 examples: flow and two species."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=9, ny=8, nz=1, seed=18, heat=False, species=2, sptrans=(0, 1), nsteps=3, dt0=3600.0, dt_max=20000.0)
