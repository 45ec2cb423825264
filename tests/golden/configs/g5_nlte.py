"""Golden case 5: NLTE two-layer heat transport (tse.py:1343-1354) -- fluid
layer with unit porosity and no source, rock layer with zero fluid
properties -- with the fluid-property update on layer 0 (tse.py:207-226)."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=6, ny=5, nz=5, seed=15, heat=True, species=1,
      sptrans=(0,), c_imblflo=True, t_imblflo=True, omega_t=0.8,
      nsteps=3, dt0=3600.0, dt_max=20000.0, nlte=True)
