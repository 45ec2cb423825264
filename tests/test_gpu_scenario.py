"""GPU: device-resident SIA steps on the performance forms (scenario engine).

The oracle runs the same SIA steps with the reference functions (tse.py
order: compute_P, compute_q, flow_error, compute_C per species, compute_T,
fluid properties, density norm, reset on a rejected Picard iteration) at
fixed dt with update-norm tolerances near machine precision, so its result
is the exact FP64 fixed point of each solve.  The scenario engine solves the
same systems to a relative residual; every field must agree within what the
residual target leaves:
  * plain FP64 to 1e-12: P within 1e-11, C and T within 1e-9 (relative L2),
  * FP32-in-FP64 to 1e-10: P within 1e-9, C and T within 1e-7.
The linear EOS and the mass sums run on the device and must equal the same
expressions evaluated by numpy on the downloaded fields.
"""
import types

import numpy as np
import pytest

from helpers import oracle_sia_steps

pytestmark = pytest.mark.gpu


def _spec():
    from transportse_amd.scenario import ScenarioSpec
    return ScenarioSpec(24, 20, 16, spacing=5.0, beta=3.0, sigma=1.5, corr=4,
                        seed=7, species=2, heat=True, salt_patch=3,
                        dt=43200.0, sia_maxit=2, sia_conv=1e-12)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("form,tol,bp,bx", [("plain64", 1e-12, 1e-11, 1e-9),
                                            ("mixed", 1e-10, 1e-9, 1e-7)])
def test_scenario_steps_match_oracle(form, tol, bp, bx):
    from transportse_amd.scenario import DeviceScenario
    spec = _spec()
    ref = oracle_sia_steps(spec, 2)
    scn = DeviceScenario(spec)
    scn.run(2, form=form, tol=tol, plan_p=(2, 16, 1, 0), plan_tr=(2, 16, 1, 0),
            sor_p=1.9, sor_tr=1.2, inner_p=20, inner_tr=4)
    got = scn.download()
    errs = {"P": _rel(got["P"], ref["P"]), "T": _rel(got["T"], ref["T"])}
    for s in range(spec.species):
        errs[f"C{s}"] = _rel(got["C"][s], ref["C"][s])
    print(form, "relative L2 error vs oracle:", errs,
          "SIA iterations", [h.sia_iters for h in scn.history], flush=True)
    assert errs["P"] <= bp
    for k, v in errs.items():
        assert v <= bx, k
    # device EOS = numpy on the downloaded fields, bit for bit
    assert np.array_equal(got["rho_f"], 1000.0 + 700.0 * got["C"][0])
    # device mass sums = numpy on the downloaded fields
    m = scn.history[-1].mass
    qx = scn.hip.download(scn.qx)
    V = spec.spacing ** 3
    area = spec.spacing ** 2
    C0 = got["C"][0]
    exp = (float(np.sum(C0 * 0.2 * V)), float(np.sum(got["rho_f"] * 0.2 * V)),
           float(np.sum(np.maximum(qx[-2], 0.0) * C0[-2] * area)
                 + np.sum(np.maximum(-qx[0], 0.0) * C0[1] * area)))
    for a, b in zip(m, exp):
        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300)
    scn.close()


def test_scenario_snapshot_pickle(tmp_path):
    """The device-side snapshot writes the 34-tuple of the state at the time
    of the call while the scenario goes on stepping."""
    from joblib import load

    from transportse_amd.scenario import DeviceScenario
    from transportse_amd.snapshot import PickleWriter
    spec = _spec()
    scn = DeviceScenario(spec)
    kw = dict(form="mixed", tol=1e-10, plan_p=(2, 16, 1, 0), plan_tr=(2, 16, 1, 0),
              sor_p=1.9, sor_tr=1.0, inner_p=20, inner_tr=4)
    scn.run(1, **kw)
    ref = scn.download()
    qx = scn.hip.download(scn.qx)
    e = scn.hip.download(scn.e)
    w = PickleWriter()
    path = str(tmp_path / "scn-00000.p")
    scn.snapshot(path, w)
    scn.run(1, **kw)            # the next step overwrites the device fields
    w.close()
    tup = load(path)
    assert len(tup) == 34
    assert tup[9] == spec.dt and (tup[3], tup[4], tup[5]) == (spec.nx, spec.ny, spec.nz)
    assert np.array_equal(tup[11], ref["P"]) and np.array_equal(tup[13], ref["T"])
    assert tup[12].shape == (spec.nx, spec.ny, spec.nz, spec.species)
    for s_ in range(spec.species):
        assert np.array_equal(tup[12][..., s_], ref["C"][s_])
    assert np.array_equal(tup[14], qx) and np.array_equal(tup[33], e)
    assert np.array_equal(tup[21], ref["rho_f"])
    assert not np.array_equal(scn.download()["P"], ref["P"])
    scn.close()


def test_scenario_segment_graphs_bit_identical():
    """The device work between solves (setups, result conversions, fluxes,
    EOS), replayed from captured hipGraphs, gives the same bits as the same
    calls issued directly, over steps whose solves end in either iterate
    buffer."""
    import torch

    from transportse_amd.scenario import DeviceScenario
    spec = _spec()
    kw = dict(form="mixed", tol=1e-10, plan_p=(2, 16, 1, 0), plan_tr=(2, 16, 1, 0),
              sor_p=1.9, sor_tr=1.0, inner_p=20, inner_tr=4)
    out = []
    for graphs in (False, True):
        stream = torch.cuda.Stream()          # capture needs a non-null stream
        with torch.cuda.stream(stream):
            scn = DeviceScenario(spec, use_graphs=graphs)
            scn.hip.sync_stream()
            res = scn.run(3, **kw)
            got = scn.download()
            got["qx"] = scn.hip.download(scn.qx)
            got["e"] = scn.hip.download(scn.e)
            got["graphs"] = sorted(scn._graphs)
            got["sweeps"] = [r.sweeps for _, r in res]
            scn.close()
        out.append(got)
    eager, replay = out
    print("captured segments:", replay["graphs"], flush=True)
    assert not eager["graphs"]
    # segments that depend on dt are kept per dt value
    dt = repr(spec.dt)
    assert {f"setup_p@{dt}", f"setup_c0@{dt}", f"setup_t@{dt}", "eos"} \
        <= set(replay["graphs"])
    assert replay["sweeps"] == eager["sweeps"]
    for k in ("P", "T", "rho_f", "qx", "e"):
        assert np.array_equal(replay[k], eager[k]), k
    for a, b in zip(replay["C"], eager["C"]):
        assert np.array_equal(a, b)


def test_scenario_halves_dt_on_a_failed_solve():
    """The first pressure solve is given too few sweeps to converge: the
    step halves dt and repeats it from the same entry state, the rest of the
    step runs at dt/2, and the next step's dt is the smallest of the solvers'
    next dt, min(dt/2 * 1.25, dt/2 * 1.05) (tse.py:724-740, 1486-1492).  The
    fields then match the oracle's SIA steps at that dt sequence within the
    mixed form's tolerances (P 1e-9, C and T 1e-7 relative L2)."""
    from transportse_amd.scenario import KIND_WAIT, DeviceScenario
    spec = _spec()
    scn = DeviceScenario(spec)
    big = 10 ** 6
    first = True
    for kind, sysm in scn.steps(2):
        if kind == KIND_WAIT:
            continue
        if kind == 0:
            budget = 3 if first else big
            r = sysm.solve("mixed", tol=1e-10, sor=1.9, inner=3, plan=(3, 16, 1, 0),
                           max_sweeps=budget)
            if first:
                assert not r.converged
            first = False
        else:
            r = sysm.solve("mixed", tol=1e-10, sor=1.2, inner=4, plan=(2, 16, 1, 0),
                           max_sweeps=big)
        scn.solved(r.converged, r.sweeps, budget if kind == 0 else big)
    dts = [spec.dt / 2, spec.dt / 2 * 1.05]
    assert scn.halvings == 1
    assert [h.dt for h in scn.history] == dts
    assert scn.t == dts[0] + dts[1]
    ref = oracle_sia_steps(spec, 2, dts=dts)
    got = scn.download()
    errs = {"P": _rel(got["P"], ref["P"]), "T": _rel(got["T"], ref["T"])}
    for s in range(spec.species):
        errs[f"C{s}"] = _rel(got["C"][s], ref["C"][s])
    print("halving: relative L2 error vs oracle", errs, flush=True)
    assert errs["P"] <= 1e-9
    for k, v in errs.items():
        assert v <= 1e-7, k
    scn.close()


def _linear_fp(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
               D_fx, D_fy, D_fz, rho_r):
    """The scenario's linear EOS written as a config's fluid_properties."""
    rho_f = 1000.0 + 700.0 * C[0]
    l_f = 2.5 * (1.0 + 0.002 * (T - 10.0))
    return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
            D_fx, D_fy, D_fz, rho_r)


def _nonlinear_fp(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky,
                  Kz, D_fx, D_fy, D_fz, rho_r):
    """Density with thermal expansion and salt, temperature-dependent
    conductivity and viscosity."""
    rho_f = 1000.0 + 700.0 * C[0] - 0.2 * (T - 10.0) - 0.004 * (T - 10.0) ** 2
    l_f = 0.56 + 0.0018 * T
    nue = 1e-3 * 2.0 ** (-(T - 10.0) / 30.0)
    return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
            D_fx, D_fy, D_fz, rho_r)


def test_scenario_host_fluid_properties():
    """A config-style fluid_properties drives the Picard update of the
    device-resident SIA step on the host.  Written as the linear EOS it
    gives the device EOS's fields bit for bit; a nonlinear one (density of
    salt and temperature, conductivity and viscosity of temperature) matches
    the oracle's SIA steps with the same callback within the mixed form's
    tolerances (P 1e-9, C and T 1e-7 relative L2)."""
    from transportse_amd.scenario import DeviceScenario
    spec = _spec()
    kw = dict(form="mixed", tol=1e-10, plan_p=(2, 16, 1, 0), plan_tr=(2, 16, 1, 0),
              sor_p=1.9, sor_tr=1.2, inner_p=20, inner_tr=4)
    a = DeviceScenario(spec)
    a.run(2, **kw)
    b = DeviceScenario(spec, fluid_properties=_linear_fp)
    b.run(2, **kw)
    ga, gb = a.download(), b.download()
    for k in ("P", "T", "rho_f"):
        assert np.array_equal(ga[k], gb[k]), k
    for x, y in zip(ga["C"], gb["C"]):
        assert np.array_equal(x, y)
    a.close()
    b.close()
    c = DeviceScenario(spec, fluid_properties=_nonlinear_fp)
    c.run(2, **kw)
    got = c.download()
    ref = oracle_sia_steps(spec, 2, eos=_nonlinear_fp)
    errs = {"P": _rel(got["P"], ref["P"]), "T": _rel(got["T"], ref["T"]),
            "rho_f": _rel(got["rho_f"], ref["rho_f"])}
    for s in range(spec.species):
        errs[f"C{s}"] = _rel(got["C"][s], ref["C"][s])
    print("host EOS: relative L2 error vs oracle", errs, flush=True)
    assert errs["P"] <= 1e-9
    for k, v in errs.items():
        assert v <= 1e-7, k
    c.close()
