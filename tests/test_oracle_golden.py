"""The oracle against the reference's own trajectories (golden vectors).

tests/golden/<case>.npz hold what tse.py itself computed (run by
tests/golden/make_golden.py with a pure-Python numba stand-in): per time step
P, C, T, qx, qy, qz, e, rho_f, cumulative P/C/T iteration counts and the
simulated time.  The SIA driver with the oracle backend must reproduce every
number bit for bit.  This pins compute_P / compute_q / flow_error /
compute_C / compute_T / dispersivity / l2_norm and the driver's control flow
(norm schedule, accept / grow / halve / reset, SIA resets, dt selection).
"""
import numpy as np
import pytest

from helpers import CASES, GOLDEN, check_plot_files, oracle_backend


@pytest.mark.parametrize("case", CASES)
def test_oracle_driver_matches_golden(case, tmp_path, oracle_mod):
    from transportse_amd.driver import Simulation
    g = np.load(f"{GOLDEN}/{case}.npz")
    sim = Simulation(f"{GOLDEN}/configs/{case}.py", oracle_backend,
                     verbose=False, record_fields=True, out_dir=str(tmp_path))
    traj = sim.run()
    assert len(traj) == g["P"].shape[0]
    for s, r in enumerate(traj):
        for k in ("P", "C", "T", "qx", "qy", "qz", "e", "rho_f"):
            assert np.array_equal(r[k], g[k][s]), (case, s, k)
        for k in ("numitP", "numitC", "numitT"):
            assert r[k] == int(g[k][s]), (case, s, k)
        assert r["t_sim"] == g["t_end"][s]


def test_golden_cases_exercise_the_control_paths():
    """The fixtures cover SIA resets, several species, heat and dt halving."""
    g2 = np.load(f"{GOLDEN}/g2_sia_heat.npz")
    assert g2["C"].shape[-1] == 2 and g2["numitT"][-1] > 0
    log3 = open(f"{GOLDEN}/g3_halving.log.txt").read()
    assert log3.count("Reducing dt") >= 3
    log2 = open(f"{GOLDEN}/g2_sia_heat.log.txt").read()
    assert "[SIA]    1" in log2 or "[SIA]    2" in log2
    # Smolarkiewicz passes double the transport solves; NLTE keeps two layers
    g4 = np.load(f"{GOLDEN}/g4_smolar.npz")
    assert "SMOLAR_C [on]" in open(f"{GOLDEN}/g4_smolar.log.txt").read()
    assert g4["numitC"][-1] > 0 and g4["numitT"][-1] > 0
    g5 = np.load(f"{GOLDEN}/g5_nlte.npz")
    assert g5["T"].ndim == 5 and g5["T"].shape[-1] == 2
    assert not np.array_equal(g5["T"][-1, ..., 0], g5["T"][-1, ..., 1])


def test_oracle_driver_output_tuples_match_reference(tmp_path, oracle_mod):
    """PLOTSAVE with PLOTTIME points between steps (golden g9): the output
    clamp shortens dt (tse.py:316-321), and the 34-tuples written at each
    PLOTTIME equal the reference's element by element.  The reference skips
    PLOTTIME points that a step oversteps (only 3 of its 5 are written), and
    so does the driver."""
    from transportse_amd.driver import Simulation
    sim = Simulation(f"{GOLDEN}/configs/g9_plot.py", oracle_backend,
                     verbose=False, record_fields=True, out_dir=str(tmp_path))
    traj = sim.run()
    g = np.load(f"{GOLDEN}/g9_plot.npz")
    assert [r["dt"] for r in traj] == g["dt"].tolist()
    assert check_plot_files(str(tmp_path), "g9_plot") == 3

