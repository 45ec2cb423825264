"""Golden case: a 1 x 1 x nz column with Dirichlet pressure top and bottom, like the reference's column examples."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=1, ny=1, nz=16, seed=17, heat=True, species=1, sptrans=(0,), nsteps=3, dt0=3600.0, dt_max=20000.0, bp_faces='z', stretch=1.0)
