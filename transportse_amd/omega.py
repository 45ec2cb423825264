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

A solve may also report the residual contraction per sweep it observed
(``lam``, over its last correction cycle after the first).  The pressure
operator is symmetric positive definite with the 7-point red/black
ordering, i.e. consistently ordered, so SOR's eigenvalues obey Young's
relation (lam + omega - 1)^2 = lam omega^2 mu^2, mu = the Jacobi
iteration's spectral radius.  Above the optimum every eigenvalue has
modulus omega - 1; below it the dominant one is real and larger.  A solve
whose contraction is clearly slower than omega - 1 therefore ran below the
optimum, and mu = (lam + omega - 1) / (omega sqrt(lam)) gives the optimum
omega_opt = 2 / (1 + sqrt(1 - mu^2)) (Young, Iterative Solution of Large
Linear Systems, 1971, ch. 6; Hageman and Young, Applied Iterative Methods,
1981, ch. 9).  The tuner moves there (rounded up to 1/128 in u), then refines
from just above it with small hill-climbing steps, where the sweep count
rises gently.  A solve whose contraction is about omega - 1 ran at or above
the optimum: until the first move to an estimate, the tuner then doubles
2 - omega (u - 1) per solve, since above the optimum the sweep count falls
monotonically towards it, and the first solve below it yields the
estimate.  The irregular sweep counts of adaptive cycles (a few percent
between neighbouring factors) mislead a hill climb alone; the estimate
does not depend on them.  Batched solves (fixed cycles) and single ones
report the same contraction, computed from the member's own residuals, so
a scenario's factors are the same alone, batched or in a server wherever
its arithmetic is; the rounding of the estimate keeps them the same where
different kernels give residuals a few ulps apart.
"""
from __future__ import annotations

import math


class OmegaTuner:
    # a contraction per sweep slower than (omega - 1)^BELOW marks a factor
    # below the optimum (the measured one above it is within ~1 % of
    # omega - 1 in rate)
    BELOW = 0.9
    MAX_JUMPS = 2             # later estimates only steer the hill climb
    GRID = 128                # estimates rounded up to 1/GRID in u (the
                              # gentle side of the optimum)

    def __init__(self, n_max: int, step: float = 0.5, u_min: float = 2.0,
                 u_max: float = 10.0, min_step: float = 1.0 / 32):
        self.u = min(max(math.log2(max(n_max, 4) / 2.0), u_min), u_max)
        self.step = step
        self.dir = 1.0
        self.u_min, self.u_max, self.min_step = u_min, u_max, min_step
        self.best_u = None
        self.best_s = None
        self.jumps = 0

    @property
    def omega(self) -> float:
        return 2.0 - 2.0 ** (-self.u)

    def is_above(self, lam: float) -> bool:
        """Whether the contraction per sweep `lam` observed at the current
        factor is about omega - 1, i.e. the factor is at or above the
        optimum."""
        om = self.omega
        return 0.0 < lam < 1.0 and 1.0 < om < 2.0 and \
            math.log(lam) <= self.BELOW * math.log(om - 1.0)

    def model_u(self, lam: float):
        """u of the optimal factor from the contraction per sweep `lam`
        observed at the current factor, if it shows that factor to be below
        the optimum (at most u_max, which is also the answer when lam is too
        close to 1 to resolve mu); else None."""
        om = self.omega
        if not (0.0 < lam < 1.0) or not (1.0 < om < 2.0) or self.is_above(lam):
            return None
        mu = (lam + om - 1.0) / (om * math.sqrt(lam))
        if mu >= 1.0:
            return self.u_max
        om_opt = 2.0 / (1.0 + math.sqrt(1.0 - mu * mu))
        return min(-math.log2(2.0 - om_opt), self.u_max)

    def report(self, sweeps: int, contraction: float = 0.0, omega=None) -> None:
        """Sweeps the last solve at ``omega`` took (and, if known, the
        residual contraction per sweep it observed); picks the next omega.
        omega: the factor the solve ended at, when it adapted its own factor
        (adaptive SOR inside the solve, which moves to the same estimate);
        the tuner continues from there as after one of its own moves to an
        estimate."""
        if omega is not None and 1.0 < omega < 2.0:
            u = min(max(-math.log2(2.0 - omega), self.u_min), self.u_max)
            if abs(u - self.u) > 1e-9:
                self.u = u
                self.jumps = max(self.jumps, 1)
                self.dir = 1.0
                self.step = min(self.step, 1.0 / 16)
        est = self.model_u(contraction) if contraction else None
        # a contraction of about omega - 1: at or above the optimum, so a
        # move that continues goes down; an estimate: below it, so up
        above = self.is_above(contraction)
        target = None if est is None else \
            min(max(math.ceil(est * self.GRID) / self.GRID, self.u_min), self.u_max)
        if target is not None and target > self.u and self.jumps < self.MAX_JUMPS:
            if self.best_s is None or sweeps < self.best_s:
                self.best_u, self.best_s = self.u, sweeps
            # below the optimum: go to the estimate, then refine
            self.jumps += 1
            self.u = target
            self.dir = 1.0
            self.step = min(self.step, 1.0 / 16)
            return
        if above and self.jumps == 0 and self.u - 1.0 >= self.u_min:
            # descending from above the optimum: every step down pays
            if self.best_s is None or sweeps < self.best_s:
                self.best_u, self.best_s = self.u, sweeps
            self.u -= 1.0
            self.dir = -1.0
            return
        if self.best_s is None or sweeps < self.best_s:
            self.best_u, self.best_s = self.u, sweeps
            if self.step < self.min_step:
                return
            if above:
                self.dir = -1.0
            elif est is not None:
                self.dir = 1.0
            nxt = self.u + self.dir * self.step
        else:
            self.dir = -self.dir
            self.step *= 0.5
            if self.step < self.min_step:
                self.u = self.best_u
                return
            nxt = self.best_u + self.dir * self.step
        self.u = min(max(nxt, self.u_min), self.u_max)
