"""``numba.core`` stand-in."""
from . import cgutils  # noqa: F401
