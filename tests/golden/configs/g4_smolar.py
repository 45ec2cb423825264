"""Golden case 4: Smolarkiewicz anti-diffusion passes for two species and
heat (tse.py:1318-1336, 1368-1381) under SIA Picard iteration."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=6, ny=5, nz=5, seed=14, heat=True, species=2,
      sptrans=(0, 1), sia=True, sia_maxit=2, sia_conv=1e-13,
      c_imblflo=True, t_imblflo=True, omega_c=1, omega_t=1, nsteps=3,
      dt0=3600.0, dt_max=20000.0, smolar_c=True, smolar_t=True)
