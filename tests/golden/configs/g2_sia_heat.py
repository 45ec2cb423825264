"""Golden case 2: flow + two species + heat under SIA Picard iteration,
OMEGA-weighted transport, SOR over-relaxation and mass-imbalance terms."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=6, ny=5, nz=5, seed=12, heat=True, species=2,
      sptrans=(0, 1), sia=True, sia_maxit=3, sia_conv=1e-13,
      c_imblflo=True, t_imblflo=True, omega_c=0.7, omega_t=0.8, sor_c=1.1,
      sor_t=1.2, nsteps=3, dt0=3600.0, dt_max=20000.0)
