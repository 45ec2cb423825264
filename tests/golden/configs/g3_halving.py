"""Golden case 3: small MAXIT_P / MAXIT_C force the dt-halving path
(compressible fluid, so halving dt strengthens the storage term)."""
from numba import njit  # noqa: F401
from tiny_base import build

build(globals(), nx=5, ny=4, nz=6, seed=13, heat=False, conv_p=1e-8,
      conv_c=1e-10, maxit_p=25, maxit_c=4, dt0=2000.0, dt_max=2000.0,
      t_end=60.0, beta=0.0, stretch=1.0, c_f=4.5e-8, salt_patch=False)
