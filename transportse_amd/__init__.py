"""MI355X (gfx950) TRANSPORTSE SOR / transport core.

Hand-written HIP kernels behind a C-ABI (include/tse_hip.h), called from
Python through ctypes, with the reference's solver call signatures:

* ``HipSolvers``   -- compute_P / compute_q / flow_error / compute_C /
  compute_T / l2_norm / smolarkiewicz / peclet_C / peclet_T for one
  grid (solvers.py)
* ``Simulation``   -- the time-step / SIA driver that calls them
  (driver.py), ``make_hip_backend`` builds its GPU backend
* ``load_config``  -- imports a TRANSPORTSE config module (config_loader.py)

Module-level drop-in: ``setup(...)`` binds a default ``HipSolvers`` and the
functions ``compute_P`` ... ``compute_T`` below forward to it, so a caller
written against tse.py can rebind its names to these.
"""
from __future__ import annotations

__version__ = "0.1.0"

_default = None


def setup(nx, ny, nz, dx, dy, dz, g, beta_deg, frozen_source=None, **kw):
    """Create the default solver object for module-level calls."""
    global _default
    from .solvers import HipSolvers
    _default = HipSolvers(nx, ny, nz, dx, dy, dz, g, beta_deg,
                          frozen_source=frozen_source, **kw)
    return _default


def _solvers():
    if _default is None:
        raise RuntimeError("transportse_amd.setup(...) has not been called")
    return _default


def compute_P(*a, **k):
    return _solvers().compute_P(*a, **k)


def compute_q(*a, **k):
    return _solvers().compute_q(*a, **k)


def flow_error(*a, **k):
    return _solvers().flow_error(*a, **k)


def compute_C(*a, **k):
    return _solvers().compute_C(*a, **k)


def compute_T(*a, **k):
    return _solvers().compute_T(*a, **k)


def l2_norm(*a, **k):
    return _solvers().l2_norm(*a, **k)


def smolarkiewicz(*a, **k):
    return _solvers().smolarkiewicz(*a, **k)


def peclet_C(*a, **k):
    return _solvers().peclet_C(*a, **k)


def peclet_T(*a, **k):
    return _solvers().peclet_T(*a, **k)


def __getattr__(name):
    if name == "HipSolvers":
        from .solvers import HipSolvers
        return HipSolvers
    if name in ("Simulation", "make_hip_backend"):
        from . import driver
        return getattr(driver, name)
    if name == "load_config":
        from .config_loader import load_config
        return load_config
    raise AttributeError(name)
