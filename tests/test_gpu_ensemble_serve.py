"""GPU: batched ensemble (lockstep stages, hipGraph-replayed cycles) and the
serving loop (continuous batching), on small grids.

* Ensemble: every realization must end bit-identical whether the batched
  cycles are replayed from a captured hipGraph or launched directly,
  whether the realizations advance asynchronously or stage by stage, and
  whether the realization runs inside the ensemble or in an ensemble of
  its own; its fields must also match the single-system scenario driver
  (same sweeps and residual tests, residual norms summed in another order)
  to 1e-12.
* Serving: a request's results must be bit-identical when it is served
  alone and when it is served among other requests arriving together, and
  every request must complete with finite latency statistics.
"""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = ((24, 20, 16), (32, 24, 20), (40, 30, 16))


def _specs(n):
    from transportse_amd.scenario import ScenarioSpec
    out = []
    rng = np.random.default_rng(9)
    for i in range(n):
        g = SMALL[i % len(SMALL)]
        out.append(ScenarioSpec(*g, spacing=5.0, beta=float(rng.uniform(0, 10)),
                                sigma=float(rng.uniform(0.5, 2.0)), corr=4,
                                seed=100 + i, species=1, heat=True, salt_patch=0,
                                c_top=float(rng.uniform(0.2, 1.0)), dt=43200.0,
                                sia_maxit=2, sia_conv=1e-12))
    return out


def _fields(ens):
    return [s.download() for s in ens.scn]


def _same(a, b):
    return (np.array_equal(a["P"], b["P"]) and np.array_equal(a["T"], b["T"])
            and all(np.array_equal(x, y) for x, y in zip(a["C"], b["C"])))


KW = dict(plan_p=(2, 16, 1, 0), plan_tr=(2, 16, 1, 0), inner_p=20, inner_tr=4,
          sor_p=1.9, sor_tr=1.2)


def test_ensemble_graph_alone_single():
    from transportse_amd.ensemble import Ensemble
    from transportse_amd.scenario import DeviceScenario
    specs = _specs(6)
    e1 = Ensemble(specs, graph=True, **KW)
    st = e1.run(2)
    f1 = _fields(e1)
    e1.close()
    print("ensemble rounds", st.rounds, "solves", st.solves, "cycles", st.cycles,
          "P sweeps", st.sweeps_p, "TR sweeps", st.sweeps_tr, flush=True)
    assert st.solves == 6 * 2 * 2 * 3   # 6 x steps x SIA iterations x (P, C, T)
    e2 = Ensemble(specs, graph=False, **KW)
    e2.run(2)
    f2 = _fields(e2)
    e2.close()
    for a, b in zip(f1, f2):
        assert _same(a, b)
    e3 = Ensemble([specs[4]], graph=True, **KW)
    e3.run(2)
    assert _same(_fields(e3)[0], f1[4])
    e3.close()
    e4 = Ensemble(specs, graph=True, **KW)     # stage-by-stage driver
    st4 = e4.run(2, mode="lockstep")
    assert st4.solves == st.solves
    for a, b in zip(f1, _fields(e4)):
        assert _same(a, b)
    e4.close()
    scn = DeviceScenario(specs[4])
    # the pressure factor comes from the scenario's own tuner, as in the ensemble
    scn.run(2, form="mixed", tol=1e-10, plan_p=KW["plan_p"], plan_tr=KW["plan_tr"],
            sor_p=None, sor_tr=KW["sor_tr"], inner_p=KW["inner_p"],
            inner_tr=KW["inner_tr"])
    g = scn.download()
    for k in ("P", "T"):
        assert np.linalg.norm(g[k] - f1[4][k]) <= 1e-12 * np.linalg.norm(g[k])
    scn.close()


def test_ensemble_runs_continue():
    """run(1) then run(1) (the second one asynchronous) ends where run(2)
    ends: the realizations stay in their batches between runs."""
    from transportse_amd.ensemble import Ensemble
    specs = _specs(3)
    e1 = Ensemble(specs, graph=True, **KW)
    e1.run(2)
    f1 = _fields(e1)
    e1.close()
    e2 = Ensemble(specs, graph=True, **KW)
    e2.run(1)
    e2.run(1, mode="async")
    f2 = _fields(e2)
    e2.close()
    for a, b in zip(f1, f2):
        assert _same(a, b)


def test_ensemble_tuner_picks_a_plan():
    from transportse_amd.ensemble import Ensemble
    e = Ensemble(_specs(3), graph=True, **KW)
    res = e.tune_plans(cycles=1, zchunks=(0,), verbose=True)
    assert res[0] is not None and res[1] is not None
    assert e.plan_p == res[0][0] and e.plan_tr == res[1][0]
    e.run(1)
    e.close()


def test_serving_alone_equals_mix():
    from transportse_amd.serve import Server, make_requests, prepare
    reqs = make_requests(6, rate=500.0, steps=2, species=3, grids=SMALL)
    for r in reqs:
        r.spec.dt = 43200.0
    prepare(reqs)
    probe = copy.deepcopy(reqs[3])
    probe.arrival = 0.0
    kw = dict(plan_p=(2, 16, 1, 8), plan_tr=(2, 16, 1, 8), inner_p=20, inner_tr=4,
              sor_p=1.9, sor_tr=1.2, keep_results=True)
    s1 = Server(**kw)
    st1 = s1.run([probe])
    s1.close()
    s2 = Server(**kw)
    st2 = s2.run(reqs)
    s2.close()
    print("alone p50", st1.p50, "mix p50", st2.p50, "p99", st2.p99,
          "max concurrent", st2.max_concurrent, flush=True)
    assert st1.completed == 1 and st2.completed == 6
    assert st2.max_concurrent >= 2
    assert np.isfinite(st2.p99) and st2.p99 >= st2.p50 > 0
    mixed = [r for r in reqs if r.rid == 3][0]
    assert mixed.solves == probe.solves and mixed.sweeps == probe.sweeps
    assert _same(mixed.result, probe.result)
