"""Pickle output (tse.py:312-329) written on the writer thread (CPU only).

The driver with the oracle backend and PLOTSAVE on must write, at every
PLOTTIME, the reference's 34-tuple holding the state of that moment, in a
file named <config>-<index>.p, finished by the time run() returns.
"""
import os

import numpy as np

from helpers import GOLDEN, oracle_backend


def test_driver_pickles_match_state(tmp_path, oracle_mod):
    from joblib import load

    from transportse_amd.driver import Simulation
    sim = Simulation(f"{GOLDEN}/configs/g2_sia_heat.py", oracle_backend,
                     verbose=False, record_fields=True, out_dir=str(tmp_path))
    c = sim.config
    c.PLOTSAVE = True
    c.PLOTTIME = np.array([0.0, 3600.0, c.t])
    init = {k: np.array(getattr(sim.st, k), copy=True) for k in ("P", "C", "T")}
    traj = sim.run()
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".p"))
    assert files == [f"g2_sia_heat-{i:05d}.p" for i in range(3)]
    tup0 = load(os.path.join(tmp_path, files[0]))
    assert len(tup0) == 34
    assert tup0[9] == 0.0
    for k, idx in (("P", 11), ("C", 12), ("T", 13)):
        assert np.array_equal(tup0[idx], init[k])
    last = load(os.path.join(tmp_path, files[2]))
    assert last[9] == traj[-1]["t_sim"] == c.t
    for k, idx in (("P", 11), ("C", 12), ("T", 13), ("qx", 14), ("e", 33),
                   ("rho_f", 21)):
        assert np.array_equal(last[idx], traj[-1][k]), k
    assert (last[3], last[4], last[5]) == (c.nx, c.ny, c.nz)
