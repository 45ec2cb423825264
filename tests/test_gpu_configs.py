"""GPU: the benchmark configs through the SIA driver, against the oracle.

The salt config at its full size (64x32x32, 25 of its 100 steps) and the
salt + heat SIA physics at 64x64x32 (stretched z, slope, salt + heat,
dispersion, SIA; the config's 20 time steps, about a minute of oracle time
on 16 host threads) run through the same driver twice -- GPU backend and
oracle backend -- and every field of every step must be bit-identical, with
equal iteration counts and times.
The salt + heat SIA config at full size (256x256x128) runs on the GPU alone:
each step must converge, and the mass-balance output must be consistent.
"""
import os

import numpy as np
import pytest

from helpers import ROOT, oracle_backend

pytestmark = pytest.mark.gpu

CFG = os.path.join(ROOT, "transportse_amd", "configs")


def _run(cfg, backend, steps, tmp):
    from transportse_amd.driver import Simulation
    sim = Simulation(os.path.join(CFG, cfg), backend, verbose=False,
                     record_fields=True, out_dir=str(tmp), max_steps=steps)
    return sim.run()


@pytest.mark.parametrize("cfg,steps", [("salt_64x32x32.py", 25),
                                       ("salt_heat_sia_64x64x32.py", 20)])
def test_config_matches_oracle(cfg, steps, tmp_path):
    from transportse_amd.driver import make_hip_backend
    th = _run(cfg, make_hip_backend, steps, tmp_path)
    to = _run(cfg, oracle_backend, steps, tmp_path)
    assert len(th) == len(to) == steps
    for s, (a, b) in enumerate(zip(th, to)):
        for k in ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f", "mass"):
            assert np.array_equal(a[k], b[k]), (cfg, s, k)
        for k in ("numitP", "numitC", "numitT", "t_sim", "dt"):
            assert a[k] == b[k], (cfg, s, k)
    print(cfg, "bit-identical over", steps, "steps; P iterations",
          th[-1]["numitP"], flush=True)


def test_config_b_full_size(tmp_path):
    from transportse_amd.driver import make_hip_backend
    th = _run("salt_heat_sia_256x256x128.py", make_hip_backend, 2, tmp_path)
    assert len(th) == 2
    m0 = None
    for r in th:
        assert np.all(np.isfinite(r["P"])) and np.all(np.isfinite(r["T"]))
        # the reference scheme (tse.py:783-916) does not bound C and T by
        # the data range: with fluxes that are not divergence-free inside
        # the Picard loop its upwind update is no convex combination, so
        # small over/undershoots are part of its solution, large ones not
        cmin, cmax = float(r["C"].min()), float(r["C"].max())
        tmin, tmax = float(r["T"].min()), float(r["T"].max())
        assert -1e-3 < cmin and cmax < 1 + 1e-3, (cmin, cmax)
        assert 10 - 0.05 < tmin and tmax < 60 + 0.05, (tmin, tmax)
        m = r["mass"]
        assert np.all(np.isfinite(m))
        if m0 is None:
            m0 = m[0]
        # salt mass changes only by what leaves through the open faces
        assert abs(m[0] - m0) <= 0.05 * m0
    print("salt_heat_sia_256x256x128: P iterations", th[-1]["numitP"], "C",
          th[-1]["numitC"], "T", th[-1]["numitT"], flush=True)


def test_dynamic_bc_in_place_changes_reach_the_device(tmp_path):
    """A dynamic_bc that changes phi and nue in place (arrays it is handed
    but does not return): the GPU backend re-uploads them and every field of
    every step equals the oracle backend's, which keeps no device copies."""
    from helpers import GOLDEN
    from transportse_amd.driver import Simulation, make_hip_backend
    cfg = os.path.join(GOLDEN, "configs", "x_dynbc_inplace.py")
    runs = []
    for be in (make_hip_backend, oracle_backend):
        sim = Simulation(cfg, be, verbose=False, record_fields=True,
                         out_dir=str(tmp_path))
        runs.append(sim.run())
    th, to = runs
    assert len(th) == len(to) == 4
    for s, (a, b) in enumerate(zip(th, to)):
        for k in ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f", "mass"):
            assert np.array_equal(a[k], b[k]), (s, k)
        for k in ("numitP", "numitC", "numitT", "t_sim", "dt"):
            assert a[k] == b[k], (s, k)
    print("dynamic_bc in place: bit-identical over 4 steps", flush=True)


def test_deferred_host_copies_equal_eager_run(tmp_path):
    """The driver with deferred host copies (solver outputs left on the
    device until host code reads them, field copies made on the device)
    ends in the same state, bit for bit, as with every output downloaded at
    once, on the salt + heat SIA physics at 64x64x32 (SIA resets included),
    with fewer downloads."""
    from transportse_amd.driver import Simulation, make_hip_backend
    res = {}
    for lazy in (False, True):
        sim = Simulation(os.path.join(CFG, "salt_heat_sia_64x64x32.py"),
                         make_hip_backend, verbose=False, out_dir=str(tmp_path),
                         max_steps=6, lazy_host=lazy)
        sim.run()
        st = sim.st
        res[lazy] = ({k: getattr(st, k).copy() for k in
                      ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f", "l_f", "mass")},
                     sim.summary, dict(sim.solvers.stats))
    (a, sa, ta), (b, sb, tb) = res[False], res[True]
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for k in ("numitP", "numitC", "numitT", "t_sim", "n_tsteps"):
        assert sa[k] == sb[k], k
    print("eager", ta, "deferred", tb, flush=True)
    assert tb["downloads"] < ta["downloads"]
