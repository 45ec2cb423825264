"""The pressure over-relaxation tuner (host logic, no GPU)."""
import math

from transportse_amd.omega import OmegaTuner


def _sweeps(opt):
    # sweep count near the optimal factor, shaped like the measured one of
    # a 128^3 high-contrast grid: 2.3x the minimum at omega = opt - 0.01,
    # 1.9x at opt + 0.005 (relative to 2 - omega)
    def f(om):
        d = om - opt
        return int(1000 * (1 + (1.0 if d > 0 else 2.5) * abs(d) / (2 - om)))
    return f


def test_start_is_two_minus_two_over_n():
    for n in (32, 64, 128, 512):
        assert math.isclose(OmegaTuner(n).omega, 2 - 2 / n)


def test_converges_near_optimum_and_is_deterministic():
    for n, opt in ((64, 1.97), (128, 1.99), (256, 1.99), (512, 1.996)):
        f = _sweeps(opt)
        runs = []
        for _ in range(2):
            t = OmegaTuner(n)
            seq = []
            for _ in range(20):
                seq.append(t.omega)
                t.report(f(t.omega))
            runs.append(seq)
        assert runs[0] == runs[1]
        final = runs[0][-1]
        assert f(final) <= 1.1 * f(opt), (n, final, f(final), f(opt))
        # settles: the last solves use the same factor
        assert len(set(runs[0][-4:])) == 1


def test_bounds():
    t = OmegaTuner(4)
    for _ in range(30):
        t.report(10)  # never improves
        assert 1.0 < t.omega < 2.0


def _rb_operator(nx, ny, nz, seed):
    """A variable-coefficient 7-point operator (positive face weights, a
    small positive storage term), red cells first: returns D_r, D_b and the
    red-black coupling B with A = [[D_r, -B], [-B^T, D_b]]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    red = ((i + j + k) % 2 == 0).ravel()
    n = nx * ny * nz
    W = np.zeros((n, n))
    for ax in range(3):
        a = np.moveaxis(idx, ax, 0)
        lo, hi = a[:-1].ravel(), a[1:].ravel()
        w = np.exp(rng.normal(0.0, 1.0, lo.size))   # lognormal transmissibilities
        W[lo, hi] = W[hi, lo] = w
    diag = W.sum(1) + 0.1 * rng.uniform(0.5, 1.5, n)   # optimum near 1.73
    r, b = np.where(red)[0], np.where(~red)[0]
    return diag[r], diag[b], W[np.ix_(r, b)]


def _sor_radius(Dr, Db, B, om):
    """Spectral radius of red/black SOR (red first) on that operator."""
    import numpy as np
    nr, nb = len(Dr), len(Db)
    Jr = B / Dr[:, None]            # red update reads black
    Jb = B.T / Db[:, None]          # black update reads the new red
    # x_r' = (1-om) x_r + om Jr x_b ; x_b' = (1-om) x_b + om Jb x_r'
    M = np.zeros((nr + nb, nr + nb))
    M[:nr, :nr] = (1 - om) * np.eye(nr)
    M[:nr, nr:] = om * Jr
    M[nr:, :nr] = om * Jb @ M[:nr, :nr]
    M[nr:, nr:] = (1 - om) * np.eye(nb) + om * Jb @ M[:nr, nr:]
    return float(np.max(np.abs(np.linalg.eigvals(M))))


def test_model_optimum_matches_the_sor_spectrum():
    """Young's relation pinned on an actual operator: the optimal factor
    estimated from SOR's spectral radius at a factor below the optimum
    equals the factor that minimises the spectral radius (grid search over
    the eigenvalues of the SOR iteration matrix), and a factor above the
    optimum (radius omega - 1) gives no estimate."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):   # small dense eigenproblems; no BLAS threads
        _spectrum_checks(np)


def _spectrum_checks(np):
    Dr, Db, B = _rb_operator(6, 5, 4, seed=3)
    # the radius is minimal at the optimum: a coarse grid, then a fine one
    coarse = np.linspace(1.4, 1.95, 111)
    om0 = coarse[np.argmin([_sor_radius(Dr, Db, B, om) for om in coarse])]
    fine = np.linspace(om0 - 0.01, om0 + 0.01, 41)
    om_best = fine[np.argmin([_sor_radius(Dr, Db, B, om) for om in fine])]
    assert 1.6 < om_best < 1.85
    for om in (1.45, 1.55, om_best - 0.05):
        t = OmegaTuner(8)
        t.u = -math.log2(2 - om)
        u = t.model_u(_sor_radius(Dr, Db, B, om))
        assert u is not None
        assert abs((2 - 2 ** -u) - om_best) <= 2e-3, (om, 2 - 2 ** -u, om_best)
    t = OmegaTuner(8)
    om = om_best + 0.05
    t.u = -math.log2(2 - om)
    lam = _sor_radius(Dr, Db, B, om)
    assert abs(lam - (om - 1)) < 1e-6
    assert t.model_u(lam) is None


def _young(mu):
    """Contraction per sweep of red/black SOR at omega for a consistently
    ordered matrix whose Jacobi spectral radius is mu."""
    om_opt = 2 / (1 + math.sqrt(1 - mu * mu))

    def lam(om):
        if om >= om_opt:
            return om - 1
        return ((om * mu + math.sqrt(om * om * mu * mu - 4 * (om - 1))) / 2) ** 2
    return lam, om_opt


def test_contraction_reports_reach_the_optimum_in_few_solves():
    """With contractions reported, the tuner reaches the optimum from the
    starting factor in a few solves, even when the sweep counts carry
    a few percent of irregular noise (as adaptive cycles do), where the hill
    climb alone wanders; deterministic."""
    for n, mu in ((128, 1 - 3e-4), (256, 1 - 8e-5), (512, 1 - 1.4e-6)):
        lam, om_opt = _young(mu)

        def sweeps(om):
            base = math.log(1e-10) / math.log(lam(om))
            noise = 1 + 0.04 * math.sin(1e4 * om)      # irregular, +-4 %
            return int(base * noise)
        seqs = []
        for _ in range(2):
            t = OmegaTuner(n)
            seq = []
            for _ in range(6):
                seq.append(t.omega)
                t.report(sweeps(t.omega), lam(t.omega))
            seqs.append(seq)
        assert seqs[0] == seqs[1]
        t.u = t.best_u
        best = min(sweeps(o) for o in (om_opt,))
        assert sweeps(t.omega) <= 1.1 * best, (n, t.omega, om_opt)
        if seqs[0][0] < om_opt:   # started below: the second solve runs at the estimate
            assert abs(seqs[0][1] - om_opt) <= 0.1 * (2 - om_opt), (n, seqs[0][1], om_opt)


def test_no_contraction_keeps_the_hill_climb():
    """Without contractions (batched solves report none) the factors are
    the hill climb's, unchanged."""
    f = _sweeps(1.99)
    a, b = OmegaTuner(128), OmegaTuner(128)
    for _ in range(12):
        assert a.omega == b.omega
        a.report(f(a.omega))
        b.report(f(b.omega), 0.0)


def test_contraction_near_one_below_the_optimum_moves_up():
    """A contraction slower than omega - 1 but too close to 1 to resolve mu
    (mu rounds to >= 1) still means "below the optimum": the tuner moves up,
    not down."""
    t = OmegaTuner(64)
    u0 = t.u
    lam = 1.0 - 1e-17          # rounds to 1 - 2^-53 or to 1 itself
    lam = min(lam, math.nextafter(1.0, 0.0))
    assert not t.is_above(lam)
    assert t.model_u(lam) == t.u_max
    t.report(1000, lam)
    assert t.u > u0


def test_library_estimate_matches_the_tuner_model():
    """tse_sor_model (the estimate the library's adaptive-SOR solves use)
    agrees with OmegaTuner.model_u on a grid of factors and contractions,
    both in deciding "below the optimum" and in the estimate (the tuner's
    estimate is itself pinned on an operator's SOR spectrum above)."""
    import ctypes
    import numpy as np
    from transportse_amd import _lib
    lib = _lib.load()
    out = ctypes.c_double()
    for om in (1.5, 1.8, 1.95, 1.99, 1.99609375, 1.999):
        t = OmegaTuner(8)
        t.u = -math.log2(2 - om)
        for lam in np.concatenate([np.linspace(0.2, 0.999, 40), 1 - np.logspace(-9, -2, 30)]):
            rc = lib.tse_sor_model(om, float(lam), ctypes.byref(out))
            u = t.model_u(float(lam))
            assert rc in (0, 1)
            assert (rc == 1) == (u is not None), (om, lam)
            if rc == 1:
                assert math.isclose(out.value, 2 - 2 ** -u, rel_tol=0, abs_tol=1e-12), \
                    (om, lam, out.value, 2 - 2 ** -u)
    assert lib.tse_sor_model(2.0, 0.5, ctypes.byref(out)) < 0
    assert lib.tse_sor_model(1.9, 1.0, ctypes.byref(out)) < 0


def test_report_adopts_the_factor_a_solve_ended_at():
    t = OmegaTuner(512)
    t.report(7000, 0.9966, omega=1.99667)
    assert abs(t.omega - 1.99667) < 2e-4 and t.step <= 1 / 16 and t.jumps >= 1
