"""Oracle restatement of the SIA driver's diagnostics: peclet_C / peclet_T
(tse.py:333-383) and the CFL maxima (tse.py:1241-1246), pinned by closed
forms on non-uniform spacing, by the masking and frozen-flag semantics, by
numpy.max's NaN propagation, and, for the CFL reshape, by a per-element loop
that divides C-order element f by dx[f % nx] / dy[f % ny]."""
import types

import numpy as np
import pytest

NX, NY, NZ = 5, 4, 3


def _oracle(BC=None, BT=None):
    from oracle import Oracle
    dx = np.array([1.0, 2.0, 4.0, 8.0, 16.0])
    dy = np.array([0.5, 1.5, 2.5, 3.5])
    dz = np.array([3.0, 5.0, 7.0])
    src = types.SimpleNamespace(BC=np.zeros((NX, NY, NZ)) if BC is None else BC,
                                BT=np.zeros((NX, NY, NZ)) if BT is None else BT)
    return Oracle(NX, NY, NZ, dx, dy, dz, 9.81, 0.0, frozen_source=src), src


def test_peclet_c_closed_form_species_max_and_mask():
    BC = np.zeros((NX, NY, NZ))
    BC[0, :, :] = 1.0
    o, src = _oracle(BC=BC)
    qx, qy, qz = np.full((NX, NY, NZ), -3.0), np.full((NX, NY, NZ), 2.0), \
        np.full((NX, NY, NZ), 0.5)
    D = np.empty((NX, NY, NZ, 2))
    D[..., 0], D[..., 1] = 1e-9, 4e-9          # the max over species is 4e-9
    Pe = np.full((NX, NY, NZ, 3), 7.0)
    m = o.peclet_C(Pe, qx, qy, qz, D, 2 * D, 3 * D)
    for i in range(NX):
        for j in range(NY):
            for k in range(NZ):
                if i == 0:
                    assert np.all(Pe[i, j, k] == 7.0)       # masked: unchanged
                    continue
                assert Pe[i, j, k, 0] == 3.0 / 4e-9 * o.dx[i]
                assert Pe[i, j, k, 1] == 2.0 / 8e-9 * o.dy[j]
                assert Pe[i, j, k, 2] == 0.5 / 12e-9 * o.dz[k]
    assert m == (3.0 / 4e-9 * 16.0, 2.0 / 8e-9 * 3.5, 0.5 / 12e-9 * 7.0)


def test_peclet_c_zero_diffusivity_gives_nan_and_nan_max():
    o, _ = _oracle()
    q = np.ones((NX, NY, NZ))
    D = np.full((NX, NY, NZ, 1), 1e-9)
    D[2, 1, 0, 0] = 0.0
    Pe = np.zeros((NX, NY, NZ, 3))
    m = o.peclet_C(Pe, q, q, q, D, D, D)
    assert np.isnan(Pe[2, 1, 0]).all() and np.isnan(m).all()
    assert np.isfinite(np.delete(Pe.reshape(-1, 3), 2 * NY * NZ + NZ, axis=0)).all()


def test_peclet_flags_frozen_at_first_call():
    BT = np.zeros((NX, NY, NZ))
    o, src = _oracle(BT=BT)
    f = np.ones((NX, NY, NZ))
    Pe = np.full((NX, NY, NZ, 3), -1.0)
    o.peclet_T(Pe, f, f, f, f, f, 0.5 * f, f, f)
    src.BT[:] = 1.0                 # later changes do not reach the kernel
    Pe[:] = -1.0
    o.peclet_T(Pe, f, f, f, f, f, 0.5 * f, f, f)
    assert np.all(Pe != -1.0)


def test_peclet_t_closed_form():
    BT = np.zeros((NX, NY, NZ))
    BT[:, :, 2] = 1.0
    o, _ = _oracle(BT=BT)
    sh = (NX, NY, NZ)
    qx, qy, qz = np.full(sh, 2e-6), np.full(sh, -1e-6), np.full(sh, 4e-6)
    rho, cp, phi = np.full(sh, 1000.0), np.full(sh, 4000.0), np.full(sh, 0.25)
    l_r, l_f = np.full(sh, 3.0), np.full(sh, 0.6)
    Pe = np.full(sh + (3,), 5.0)
    m = o.peclet_T(Pe, qx, qy, qz, rho, cp, phi, l_r, l_f)
    den = 0.75 * 3.0 + 0.25 * 0.6
    i, j, k = 3, 2, 1
    assert Pe[i, j, k, 0] == 2e-6 * 1000.0 * 4000.0 / den * o.dx[i]
    assert Pe[i, j, k, 1] == 1e-6 * 1000.0 * 4000.0 / den * o.dy[j]
    assert Pe[i, j, k, 2] == 4e-6 * 1000.0 * 4000.0 / den * o.dz[k]
    assert np.all(Pe[:, :, 2] == 5.0)
    assert m[2] == 4e-6 * 1000.0 * 4000.0 / den * o.dz[1]   # k = 2 is masked


def test_cfl_reshape_divides_by_flat_index_spacing():
    """The reference reshapes |qx|/phi to (ny, nz, nx) before dividing by dx:
    C-order element f = (i*ny + j)*nz + k meets dx[f % nx], not dx[i]."""
    o, _ = _oracle()
    rng = np.random.default_rng(3)
    qx, qy, qz = (rng.normal(size=(NX, NY, NZ)) for _ in range(3))
    phi = rng.uniform(0.1, 0.4, size=(NX, NY, NZ))
    m = o.cfl_maxima(qx, qy, qz, phi)
    ex = ey = ez = -np.inf
    for f in range(NX * NY * NZ):
        i, j, k = f // (NY * NZ), (f // NZ) % NY, f % NZ
        ex = max(ex, abs(qx[i, j, k]) / phi[i, j, k] / o.dx[f % NX])
        ey = max(ey, abs(qy[i, j, k]) / phi[i, j, k] / o.dy[f % NY])
        ez = max(ez, abs(qz[i, j, k]) / phi[i, j, k] / o.dz[k])
    assert m == (ex, ey, ez)
    # with non-uniform dx the quirk differs from the cell's own spacing
    own = max(abs(qx[i, j, k]) / phi[i, j, k] / o.dx[i]
              for i in range(NX) for j in range(NY) for k in range(NZ))
    assert own != ex
