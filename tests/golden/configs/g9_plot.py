"""Golden case 9: salt + heat with PLOTSAVE on and PLOTTIME points between
the 1-day steps, so that the output clamp of tse.py:312-329 shortens dt
mid-run and a 34-tuple is written at each PLOTTIME."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=7, ny=6, nz=5, seed=19, heat=True, nsteps=4,
      plot_times=[0.0, 100000.0, 200000.0, 250000.0, 345600.0])
