"""GPU: batched fast-form solves (persistent tile-queue engine).

A member of a batch must follow exactly the arithmetic of its own
single-system solve: same sweeps, same cycles, bit-identical iterate -- for
pressure and transport, FP64 and mixed forms, grids of uneven sizes, with
and without hipGraph replay of the cycle.  A member solved alone in a batch
and solved in a mix of other members must agree bit for bit, residual
included, and re-arming a member for a second solve must reproduce the
first solve.
"""
import numpy as np
import pytest

from helpers import Fields

pytestmark = pytest.mark.gpu

SHAPES = [(37, 23, 19), (64, 40, 33), (20, 50, 17), (90, 31, 12), (45, 45, 45)]
DT = 86400.0 * 5


def _stream():
    import torch
    return torch.cuda.Stream()


def _pressure(shape, seed):
    from transportse_amd.pressure import FastPressure
    f = Fields(*shape, seed=seed)
    fp = FastPressure(*f.grid_args())
    fp.load(f.P, DT, **{n: getattr(f, n) for n in
                        ("c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz",
                         "BP", "WP", "c_r")})
    return fp


def _transport(shape, seed, which):
    from transportse_amd.transport import FastTransport
    f = Fields(*shape, seed=seed)
    f.qx *= 30.0
    f.qy *= 30.0
    f.qz *= 30.0
    ft = FastTransport(*f.grid_args())
    if which == "C":
        ft.load_c(f.C[..., 0], f.D_fx[..., 0], f.D_fy[..., 0], f.D_fz[..., 0],
                  f.phi, f.rho_f, f.qx, f.qy, f.qz, DT, f.BC, f.WC[..., 0], f.R,
                  f.e, 1.0, lf_l=f.lf_l, lf_t=f.lf_t, imblflo=True)
    else:
        ft.load_t(f.T, f.l_f, f.cp_f, f.rho_f, f.phi, f.qx, f.qy, f.qz, DT,
                  f.BT, f.WT, f.e, 0.8, l_r=f.l_r, cp_r=f.cp_r, rho_r=f.rho_r,
                  lt_l=f.lt_l, lt_t=f.lt_t, imblflo=False)
    return ft


def _systems(kind):
    if kind == 0:
        return [_pressure(s, 50 + n) for n, s in enumerate(SHAPES)]
    return [_transport(s, 60 + n, "CT"[n % 2]) for n, s in enumerate(SHAPES)]


# SOR factor: pressure is symmetric positive definite (any 0 < w < 2
# converges); the upwind transport operator is not, so it stays near 1.
SOR = {0: 1.9, 1: 1.2}
CASES = [(0, "mixed", (2, 16, 1, 0), 1e-10),
         (0, "mixed", (2, 24, 2, 6, 1), 1e-10),
         (0, "mixed", (2, 16, 1, 0, 2), 1e-10),
         (0, "plain64", (2, 16, 1, 9, 2), 1e-11),
         (0, "plain64", (1, 16, 1, 7), 1e-11),
         (1, "mixed", (3, 16, 1, 8), 1e-11),
         (1, "mixed", (2, 16, 1, 5, 2), 1e-11),
         (1, "plain64", (2, 16, 1, 0), 1e-12)]


@pytest.mark.parametrize("kind,form,plan,tol", CASES)
def test_batch_matches_single_solves(kind, form, plan, tol):
    import torch
    from transportse_amd.batch import Batch
    with torch.cuda.stream(_stream()):
        systems = _systems(kind)
        alone = []
        for s in systems:
            st = s.solve(form, tol=tol, sor=SOR[kind], inner=20, plan=plan,
                         max_sweeps=20000)
            assert st.converged
            alone.append((st, s.result()))
            s.assemble()
        for graph in (False, True):
            b = Batch(kind, form, plan, inner=20, tol=tol, sor=SOR[kind],
                      max_sweeps=20000)
            slots = [b.add(s) for s in systems]
            cycles = b.run(graph=graph)
            for (st, x), s, slot in zip(alone, systems, slots):
                ms = b.status(slot)
                b.writeback(slot)
                y = s.result()
                print(kind, form, plan, "graph", graph, "sweeps", ms.sweeps,
                      st.sweeps, "rel", ms.rel_residual, st.rel_residual,
                      flush=True)
                assert ms.converged and not ms.active
                assert ms.sweeps == st.sweeps and ms.cycles == st.cycles
                assert np.array_equal(x, y)
                assert abs(ms.rel_residual - st.rel_residual) <= 1e-9 * st.rel_residual
            assert cycles == max(st.cycles for st, _ in alone)
            b.close()
            for s in systems:
                s.assemble()


@pytest.mark.parametrize("kind", [0, 1])
def test_alone_equals_mix_and_rearm(kind):
    import torch
    from transportse_amd.batch import Batch
    form, plan, tol = "mixed", (2, 16, 1, 0), 1e-10
    with torch.cuda.stream(_stream()):
        systems = _systems(kind)
        probe = systems[1]
        b1 = Batch(kind, form, plan, inner=20, tol=tol, sor=SOR[kind])
        s1 = b1.add(probe)
        b1.run(graph=True)
        st1 = b1.status(s1)
        b1.writeback(s1)
        x1 = probe.result()
        b1.close()
        for s in systems:
            s.assemble()
        b2 = Batch(kind, form, plan, inner=20, tol=tol, sor=SOR[kind])
        slots = [b2.add(s) for s in systems]
        b2.run(graph=True)
        st2 = b2.status(slots[1])
        b2.writeback(slots[1])
        x2 = probe.result()
        assert np.array_equal(x1, x2)
        assert (st1.sweeps, st1.cycles, st1.rel_residual) == \
            (st2.sweeps, st2.cycles, st2.rel_residual)
        # second solve of the same members: re-arm, replay the captured cycle
        first = {}
        for s, slot in zip(systems, slots):
            b2.writeback(slot)
            first[slot] = s.result()
            s.assemble()
            b2.rearm(slot)
        b2.run(graph=True)
        for s, slot in zip(systems, slots):
            b2.writeback(slot)
            assert np.array_equal(first[slot], s.result())
        print("alone == mix, re-arm reproduces; kind", kind, "sweeps",
              st2.sweeps, flush=True)
        b2.close()


def test_per_member_sor_and_rearm_many():
    """Per-member over-relaxation factors (set after add: uploaded with the
    member table before the next cycle), then a second solve of every member
    re-armed at once (one initial-residual launch for all), then a re-arm of
    two members only: each solve must match the member's single-system solve
    at its own factor bit for bit, and the members left alone keep their
    iterate."""
    import torch
    from transportse_amd.batch import Batch
    form, plan, tol = "mixed", (2, 16, 1, 0), 1e-10
    sors = [1.85, 1.9, 1.95, 1.8, 1.93]
    with torch.cuda.stream(_stream()):
        systems = _systems(0)
        alone = []
        for s, w in zip(systems, sors):
            st = s.solve(form, tol=tol, sor=w, inner=20, plan=plan, max_sweeps=20000)
            assert st.converged
            alone.append((st, s.result()))
            s.assemble()
        b = Batch(0, form, plan, inner=20, tol=tol, sor=1.9, max_sweeps=20000)
        slots = [b.add(s) for s in systems]
        for slot, w in zip(slots, sors):
            b.set_sor(slot, w)

        def check(which):
            for n in which:
                st, x = alone[n]
                ms = b.status(slots[n])
                b.writeback(slots[n])
                assert ms.converged and ms.sweeps == st.sweeps
                assert np.array_equal(x, systems[n].result())

        b.run(graph=True)
        check(range(len(systems)))
        for s in systems:
            s.assemble()
        b.rearm_many(slots)                       # every member at once
        b.run(graph=True)
        check(range(len(systems)))
        kept = {n: systems[n].result() for n in (1, 3, 4)}
        for n in (0, 2):
            systems[n].assemble()
        b.rearm_many([slots[0], slots[2]])        # two members only
        b.run(graph=True)
        check((0, 2))
        for n, x in kept.items():
            b.writeback(slots[n])
            assert np.array_equal(x, systems[n].result())
        print("per-member sor and rearm_many ok", flush=True)
        b.close()


EVEN_SHAPES = [(36, 23, 19), (64, 40, 33), (20, 50, 17), (90, 31, 12), (46, 45, 45)]


@pytest.mark.parametrize("plan", [(3, 32, 2, 0, 3), (2, 32, 2, 7, 3), (3, 32, 2, 0, 5),
                                  (2, 32, 2, 7, 5), (3, 32, 2, 9, 5)])
def test_batch_pair_variant_matches_single_solves(plan):
    """Sweep variants 3 (cell pairs) and 5 (the same on the bfloat16 inner
    operator) in the persistent batch kernel: each member bit-identical to
    its single-system solve, with and without graph replay."""
    import torch
    from transportse_amd.batch import Batch
    tol = 1e-10
    with torch.cuda.stream(_stream()):
        systems = [_pressure(sh, 70 + n) for n, sh in enumerate(EVEN_SHAPES)]
        alone = []
        for s in systems:
            st = s.solve("mixed", tol=tol, sor=1.9, inner=18, plan=plan,
                         max_sweeps=20000)
            assert st.converged
            alone.append((st, s.result()))
            s.assemble()
        for graph in (False, True):
            b = Batch(0, "mixed", plan, inner=18, tol=tol, sor=1.9, max_sweeps=20000)
            slots = [b.add(s) for s in systems]
            b.run(graph=graph)
            for (st, x), s, slot in zip(alone, systems, slots):
                ms = b.status(slot)
                b.writeback(slot)
                assert ms.converged and ms.sweeps == st.sweeps
                assert np.array_equal(x, s.result())
            b.close()
            for s in systems:
                s.assemble()
        print("batch pair variant", plan, "ok", flush=True)


@pytest.mark.parametrize("ns,zchunk", [(2, 0), (2, 7), (3, 0), (3, 9)])
@pytest.mark.parametrize("var", [3, 5])
def test_batch_odd_and_even_nx(ns, zchunk, var):
    """Members with odd and even nx in one batch of a cell-pair plan (odd
    rows get an inactive pad column): each member bit-identical to its
    single-system solve with the same plan, with graph replay."""
    import torch
    from transportse_amd.batch import Batch
    tol = 1e-10
    shapes = [(36, 23, 19), (37, 23, 19), (64, 40, 33), (45, 45, 45), (90, 31, 12)]
    plan = (ns, 32, 2, zchunk, var)
    with torch.cuda.stream(_stream()):
        systems = [_pressure(sh, 80 + n) for n, sh in enumerate(shapes)]
        alone = []
        for s in systems:
            st = s.solve("mixed", tol=tol, sor=1.9, inner=18, plan=plan, max_sweeps=20000)
            assert st.converged
            alone.append((st, s.result()))
            s.assemble()
        b = Batch(0, "mixed", plan, inner=18, tol=tol, sor=1.9, max_sweeps=20000)
        slots = [b.add(s) for s in systems]
        b.run(graph=True)
        for (st, x), s, slot in zip(alone, systems, slots):
            ms = b.status(slot)
            b.writeback(slot)
            assert ms.converged and ms.sweeps == st.sweeps
            assert np.array_equal(x, s.result())
        b.close()
        print("batch odd and even nx", plan, "ok", flush=True)
