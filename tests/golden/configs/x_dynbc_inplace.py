"""Not a golden case (no reference trajectory): the salt config of g2 with a
dynamic_bc callback that scales phi and nue in place every step, arrays it
receives but does not return (tse.py:1451-1453).  The GPU driver, which
keeps device copies of its inputs, must follow the oracle driver bit for
bit."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=8, ny=6, nz=5, seed=21, heat=True, species=2,
      sptrans=(0, 1), sia=True, sia_maxit=3, nsteps=4, dyn_inplace=True)
