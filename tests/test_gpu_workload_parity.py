"""GPU: the ensemble and serving workloads against the oracle.

Members of a batched ensemble of the ensemble workload's kind (heat + one
species, sigma_lnK in [0.5, 2], slope in [0, 10] deg, top salt Dirichlet in
[0.2, 1], uneven grids with odd and even nx, the cell-pair sweep on the
bfloat16 inner operator, hipGraph-replayed cycles), and requests of the serving
workload's kind (heat + three species, sigma_lnK in [0.5, 2.5]) served
continuously in a mix of four, are compared with the oracle's SIA steps
(tests/helpers.py oracle_sia_steps: the reference functions in tse.py
order, every solve converged to an update norm of 1e-15).  Tolerances of
the mixed form at relative residual 1e-10, relative L2 per field: P 1e-9,
T and every C 1e-7.

A small HBM budget makes admission bind: requests wait for earlier ones to
finish, at most two run at once, and every result still equals the
request's run alone bit for bit.
"""
import copy

import numpy as np
import pytest

from helpers import oracle_sia_steps

pytestmark = pytest.mark.gpu

BOUND_P, BOUND_X = 1e-9, 1e-7


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check(got, ref, label):
    errs = {"P": _rel(got["P"], ref["P"]), "T": _rel(got["T"], ref["T"])}
    for s, (a, b) in enumerate(zip(got["C"], ref["C"])):
        errs[f"C{s}"] = _rel(a, b)
    print(label, "relative L2 vs oracle", errs, flush=True)
    assert errs["P"] <= BOUND_P, (label, errs)
    for k, v in errs.items():
        assert v <= BOUND_X, (label, k, v)


def _ensemble_specs():
    from transportse_amd.scenario import ScenarioSpec
    rng = np.random.default_rng(41)
    out = []
    for g in ((21, 18, 14), (26, 15, 12), (17, 22, 10)):
        out.append(ScenarioSpec(*g, spacing=5.0, beta=float(rng.uniform(0, 10)),
                                sigma=float(rng.uniform(0.5, 2.0)), corr=4,
                                seed=int(rng.integers(1, 2 ** 31 - 1)), species=1,
                                heat=True, salt_patch=0,
                                c_top=float(rng.uniform(0.2, 1.0)), dt=3600.0,
                                sia_maxit=2, sia_conv=1e-12))
    return out


def test_ensemble_members_match_oracle():
    from transportse_amd.ensemble import Ensemble
    specs = _ensemble_specs()
    # the plans the ensemble benchmark runs: the cell-pair sweep on the
    # bfloat16 inner operator for every member (odd nx: one pad column)
    ens = Ensemble(specs, graph=True, plan_p=(3, 32, 2, 0, 5),
                   plan_tr=(2, 16, 1, 0, 2), inner_p=51, inner_tr=6)
    st = ens.run(2)
    assert st.solves == 3 * 2 * 2 * 3
    for i, (scn, spec) in enumerate(zip(ens.scn, specs)):
        assert scn.halvings == 0 and scn.t == 2 * spec.dt
        _check(scn.download(), oracle_sia_steps(spec, 2), f"ensemble member {i}")
    ens.close()


def _requests(n, grids, species=3):
    from transportse_amd.scenario import ScenarioSpec
    from transportse_amd.serve import Request, prepare
    rng = np.random.default_rng(43)
    reqs = []
    for i in range(n):
        g = grids[i % len(grids)]
        spec = ScenarioSpec(*g, spacing=5.0, beta=0.0,
                            sigma=float(rng.uniform(0.5, 2.5)), corr=4,
                            seed=int(rng.integers(1, 2 ** 31 - 1)),
                            species=species, heat=True, salt_patch=3, dt=3600.0,
                            sia_maxit=2, sia_conv=1e-12)
        reqs.append(Request(i, 0.002 * i, spec, 2))
    prepare(reqs)
    return reqs


SERVE_KW = dict(plan_tr=(2, 16, 1, 0, 2), inner_p=51, inner_tr=6,
                keep_results=True)


def test_served_requests_match_oracle():
    from transportse_amd.serve import Server
    reqs = _requests(4, ((20, 16, 12), (24, 18, 14), (22, 20, 10), (18, 14, 16)))
    srv = Server(**SERVE_KW)
    st = srv.run(reqs)
    srv.close()
    print("served 4: max concurrent", st.max_concurrent, "p50", st.p50, flush=True)
    assert st.completed == 4 and st.max_concurrent >= 2
    for r in reqs[:2]:
        _check(r.result, oracle_sia_steps(r.spec, 2), f"request {r.rid}")


def test_admission_budget_binds_and_results_equal_alone():
    from transportse_amd.scenario import field_bytes, uniform_bytes
    from transportse_amd.serve import Server
    grids = ((24, 18, 14), (22, 20, 10), (20, 16, 12))
    reqs = _requests(5, grids)
    alone = []
    for r in reqs:
        p = copy.deepcopy(r)
        p.arrival = 0.0
        srv = Server(**SERVE_KW)
        srv.run([p])
        srv.close()
        alone.append(p)
    # room for the two largest requests (with their shapes' shared fields),
    # never for three
    need = sorted(field_bytes(r.spec) + uniform_bytes(r.spec) for r in reqs)
    budget = need[-1] + need[-2] + 1
    assert need[0] + need[1] + need[2] > budget
    srv = Server(budget_bytes=budget, **SERVE_KW)
    st = srv.run(reqs)
    srv.close()
    waited = [r.admitted - r.arrival for r in reqs]
    print("budget", budget, "max concurrent", st.max_concurrent,
          "admission waits", waited, "peak bytes", st.peak_bytes, flush=True)
    assert st.completed == 5
    assert st.max_concurrent <= 2 and st.peak_bytes <= budget
    # the last request admitted waited for a finished one
    assert max(r.admitted for r in reqs) >= min(r.done for r in reqs)
    for r, a in zip(reqs, alone):
        assert r.solves == a.solves and r.sweeps == a.sweeps
        for k in ("P", "T"):
            assert np.array_equal(r.result[k], a.result[k]), (r.rid, k)
        for x, y in zip(r.result["C"], a.result["C"]):
            assert np.array_equal(x, y), r.rid
