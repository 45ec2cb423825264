"""CPU: host-side scheduling helpers (no GPU calls).

* ``scenario.resume_all`` advances step generators past their reduction
  reads (KIND_WAIT) to the next solve, resuming every waiting generator of a
  round before the next round, and reports ended generators as None.
* ``device._chunks`` splits a transfer into consecutive parts that cover the
  range exactly once.
"""
from transportse_amd.device import _chunks
from transportse_amd.scenario import KIND_WAIT, resume_all


def _gen(log, name, script):
    for item in script:
        log.append((name, item))
        yield item


def test_resume_all_skips_waits_and_reports_ends():
    log = []
    gens = {
        0: _gen(log, 0, [(KIND_WAIT, None), (0, "p0")]),
        1: _gen(log, 1, [(1, "t1")]),
        2: _gen(log, 2, [(KIND_WAIT, None), (KIND_WAIT, None)]),
        3: _gen(log, 3, []),
    }
    out = resume_all(gens, [0, 1, 2, 3])
    assert out == {0: (0, "p0"), 1: (1, "t1"), 2: None, 3: None}
    # the first pass takes one item of every generator before any is resumed
    # again, so all reductions of a pass are enqueued before the first read
    first_pass = [name for name, _ in log[:3]]
    assert first_pass == [0, 1, 2]
    assert log[3:] == [(0, (0, "p0")), (2, (KIND_WAIT, None))]


def test_resume_all_subset_leaves_others_untouched():
    log = []
    gens = {0: _gen(log, 0, [(0, "a"), (0, "b")]), 1: _gen(log, 1, [(1, "c")])}
    assert resume_all(gens, [1]) == {1: (1, "c")}
    assert log == [(1, (1, "c"))]
    assert resume_all(gens, [0]) == {0: (0, "a")}
    assert resume_all(gens, [0]) == {0: (0, "b")}
    assert resume_all(gens, [0]) == {0: None}


def test_chunks_cover_the_range_once():
    for m, n in [(1, 1), (64, 64 * 32 * 32), (256, 256 * 256 * 128), (5, 5 << 22),
                 (3, 3 << 21), (512, 1 << 20), (7, 7 << 25), (1000, 1000)]:
        parts = _chunks(m, n)
        assert 1 <= len(parts) <= 8
        assert parts[0][0] == 0 and parts[-1][1] == m
        for (a, b), (c, d) in zip(parts, parts[1:]):
            assert b == c
        assert all(lo < hi for lo, hi in parts)
        if n < 1 << 21:
            assert len(parts) == 1


def test_strided_views_of_foreign_buffers_take_the_copy_path(tmp_path):
    """A species slice of an array whose base chain ends in something other
    than an ndarray (np.frombuffer over bytes, np.memmap over a file) is
    gathered through a contiguous copy instead of failing on the walk."""
    import numpy as np

    from transportse_amd.device import _block_and_strides
    shp = (4, 3, 5, 2)
    ref = np.arange(np.prod(shp), dtype=np.float64).reshape(shp)
    raw = np.frombuffer(ref.tobytes(), dtype=np.float64).reshape(shp)
    mm = np.memmap(tmp_path / "f.bin", dtype=np.float64, mode="w+", shape=shp)
    mm[...] = ref
    for a in (raw, mm, ref):
        view = a[:, :, :, 1]
        flat, (si, sj, sk) = _block_and_strides(view)
        i, j, k = np.indices(view.shape)
        assert np.array_equal(flat[i * si + j * sj + k * sk], ref[:, :, :, 1])


def test_weak_view_references_follow_their_owner():
    """The solvers' cache and pending list hold host arrays weakly; a species
    slice is held through its owner, and rebuilt as the same strided view
    while the owner lives."""
    import gc

    import numpy as np

    from transportse_amd.solvers import _held, _holder, _key
    C = np.arange(4 * 3 * 5 * 2, dtype=np.float64).reshape(4, 3, 5, 2)
    v = C[:, :, :, 1]
    h = _holder(v)
    w = _held(h)
    assert _key(w) == _key(v) and np.array_equal(w, v)
    w[0, 0, 0] = -1.0                   # a write through it reaches C
    assert C[0, 0, 0, 1] == -1.0
    P = np.zeros((3, 3, 3))
    hp = _holder(P)
    assert _held(hp) is P
    del C, v, w, P
    gc.collect()
    assert _held(h) is None and _held(hp) is None


def test_vectorised_eos_falls_back_when_cells_disagree(tmp_path, oracle_mod):
    """A fluid_properties whose whole-field call differs from its per-cell
    call (here: a density that uses the field mean, which per cell is the
    cell itself) is caught by the sample check, and the driver runs the
    per-cell loop, as the reference does (tse.py:181-196)."""
    import numpy as np

    from helpers import GOLDEN, oracle_backend
    from transportse_amd.driver import Simulation
    sim = Simulation(f"{GOLDEN}/configs/g1_salt.py", oracle_backend,
                     verbose=False, out_dir=str(tmp_path), max_steps=1)
    c = sim.config

    def fp(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
           D_fx, D_fy, D_fz, rho_r):
        rho_f = 1000.0 + 700.0 * np.mean(C[0])
        return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
                D_fx, D_fy, D_fz, rho_r)
    c.fluid_properties = fp
    sim._eos_vectorized = None
    sim.update_fluid_properties()
    assert sim._eos_vectorized is False
    expect = 1000.0 + 700.0 * sim.st.C[..., 0]
    assert np.array_equal(sim.st.rho_f, expect)


def test_slabbed_eos_equals_the_per_cell_loop(tmp_path, oracle_mod, monkeypatch):
    """The whole-field EOS call split into x-slabs on threads (forced here on
    a small grid) gives the per-cell loop's fields bit for bit, species-last
    C included, and a callback that differs per cell still falls back."""
    import numpy as np

    from helpers import GOLDEN, oracle_backend
    from transportse_amd import driver
    from transportse_amd.driver import Simulation
    monkeypatch.setattr(driver, "_EOS_MIN_CELLS", 1)
    sim = Simulation(f"{GOLDEN}/configs/g1_salt.py", oracle_backend,
                     verbose=False, out_dir=str(tmp_path), max_steps=1)
    assert len(sim._eos_slabs()) == min(driver._EOS_THREADS, sim.st.nx) > 1
    c = sim.config
    rng = np.random.default_rng(3)
    st = sim.st
    st.C[...] = rng.random(st.C.shape)
    st.T[...] = 10.0 + 5.0 * rng.random(st.T.shape)

    def fp(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
           D_fx, D_fy, D_fz, rho_r):
        rho_f = 1000.0 + 700.0 * C[0] - 0.2 * (T - 4.0) ** 2
        nue = 1.0e-3 * np.exp(-0.02 * T) * (1.0 + 0.5 * C[0])
        C = C * 0.999
        return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
                D_fx, D_fy, D_fz, rho_r)
    c.fluid_properties = fp
    names = ("C", "T", "rho_f", "nue", "l_f")
    before = {n: getattr(st, n).copy() for n in names}
    sim._eos_loop()
    loop = {n: getattr(st, n).copy() for n in names}
    for n in names:
        getattr(st, n)[...] = before[n]
    sim._eos_vectorized = None
    changed = sim.update_fluid_properties()
    assert sim._eos_vectorized is True
    assert any(a is st.rho_f for a in changed) and any(a is st.C for a in changed)
    assert not any(a is st.T for a in changed)
    for n in names:
        assert np.array_equal(getattr(st, n), loop[n]), n

    def fp_mean(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky,
                Kz, D_fx, D_fy, D_fz, rho_r):
        rho_f = 1000.0 + 700.0 * np.mean(C[0])
        return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky, Kz,
                D_fx, D_fy, D_fz, rho_r)
    c.fluid_properties = fp_mean
    sim._eos_vectorized = None
    sim.update_fluid_properties()
    assert sim._eos_vectorized is False
    assert np.array_equal(st.rho_f, 1000.0 + 700.0 * st.C[..., 0])
