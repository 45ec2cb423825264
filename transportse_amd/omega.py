"""Online choice of a grid's pressure over-relaxation factor.

Red/black SOR on the pressure system converges several times faster near the
optimal factor than a few hundredths away from it, and the optimum depends
on the grid size and on the permeability contrast.  A scenario solves its
pressure system a hundred times or more (every Picard iteration of every
step) with slowly changing coefficients, so the factor is tuned on the run
itself: ``OmegaTuner`` keeps omega = 2 - 2**(-u), starts at
omega = 2 - 2 / N_max (u = log2(N_max / 2)), the asymptotic form of the
optimum for a homogeneous grid of N_max cells along its longest axis, and
after every solve moves u by a step towards the side that reduced the sweep
count, reversing and halving the step when a move did not.  Its state
depends only on the scenario's own solves, so a scenario gets the same
factors, hence the same arithmetic, whatever it is batched with.
"""
from __future__ import annotations

import math


class OmegaTuner:
    def __init__(self, n_max: int, step: float = 0.5, u_min: float = 2.0,
                 u_max: float = 10.0, min_step: float = 1.0 / 32):
        self.u = min(max(math.log2(max(n_max, 4) / 2.0), u_min), u_max)
        self.step = step
        self.dir = 1.0
        self.u_min, self.u_max, self.min_step = u_min, u_max, min_step
        self.best_u = None
        self.best_s = None

    @property
    def omega(self) -> float:
        return 2.0 - 2.0 ** (-self.u)

    def report(self, sweeps: int) -> None:
        """Sweeps the last solve at ``omega`` took; picks the next omega."""
        if self.best_s is None or sweeps < self.best_s:
            self.best_u, self.best_s = self.u, sweeps
            if self.step < self.min_step:
                return
            nxt = self.u + self.dir * self.step
        else:
            self.dir = -self.dir
            self.step *= 0.5
            if self.step < self.min_step:
                self.u = self.best_u
                return
            nxt = self.best_u + self.dir * self.step
        self.u = min(max(nxt, self.u_min), self.u_max)
