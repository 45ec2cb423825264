"""GPU: the salt + heat SIA config at its full size (256x256x128) for one
step through the driver with the GPU backend and with the oracle backend;
every field must be bit-identical, with equal iteration counts.

The oracle needs about four minutes for the step on 16 host threads (its
compute_P alone runs 3840 iterations); the test prints a line after every
oracle solver call and every 500 compute_P iterations, past pytest's
output capture.  TSE_SKIP_FULL_B=1 skips it for quick local runs.
"""
import os
import time

import numpy as np
import pytest

from helpers import ROOT, oracle_backend

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(bool(os.environ.get("TSE_SKIP_FULL_B")),
                                 reason="TSE_SKIP_FULL_B=1: ~4 min of oracle time")]

CFG = os.path.join(ROOT, "transportse_amd", "configs", "salt_heat_sia_256x256x128.py")


class _Loud:
    """The oracle backend, printing after each solver call."""

    def __init__(self, inner, say):
        self.inner = inner
        self.say = say

    def __getattr__(self, name):
        f = getattr(self.inner, name)
        if not callable(f):
            return f

        def call(*a, **k):
            t0 = time.time()
            r = f(*a, **k)
            self.say(f"  oracle {name}: {time.time() - t0:.1f}s")
            return r
        return call


def test_config_b_full_size_matches_oracle(tmp_path, capsys):
    import oracle as orc
    from transportse_amd.driver import Simulation, make_hip_backend

    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)

    orc.set_threads(len(os.sched_getaffinity(0)))
    run = lambda backend: Simulation(CFG, backend, verbose=False, record_fields=True,
                                     out_dir=str(tmp_path), max_steps=1).run()
    th = run(make_hip_backend)
    orc.set_progress(lambda n: say(f"  oracle compute_P iteration {n}"), 500)
    try:
        to = run(lambda sim: _Loud(oracle_backend(sim), say))
    finally:
        orc.set_progress(None)
    assert len(th) == len(to) == 1
    a, b = th[0], to[0]
    for k in ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f", "mass"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("numitP", "numitC", "numitT", "t_sim", "dt"):
        assert a[k] == b[k], k
    print("256x256x128 step bit-identical; P iterations", a["numitP"], flush=True)
