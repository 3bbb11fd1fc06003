"""Potential coefficients P = W^T M0 W / (4 pi eps0 er) on the device (src/pcbqs/mqs.py:532-556)
against the oracle's scalar fill and the explicit sparse product, with merged charge nodes; the
parallel-plate bracket eps A / h < C < 1.5 eps A / h (the reference's tests/test_peec.py:31-33) of the
capacitance that a charge solve on P gives; and the device capacitance matrix C = B^T P^-1 B
(src/pcbqs/peec.py:95-123) with its dense symmetric positive definite solve."""
import numpy as np
import pytest

from oracle import fill as F
from tests.meshes import two_layer

pytestmark = pytest.mark.gpu


def oracle_potential(nodes, tris, z, node_of, er=1.0):
    from pcbqs_amd.geometry import EPS0
    g = F.prepare(nodes, tris, z, 1)
    planes, _, _ = F.cr_fill(g, 3.0, 1, True)
    M0 = planes[0]
    n = node_of.max() + 1
    area = np.bincount(node_of, weights=g.area, minlength=n)
    W = np.zeros((len(tris), n))
    W[np.arange(len(tris)), node_of] = 1 / area[node_of]
    return W.T @ M0 @ W / (4 * np.pi * EPS0 * er)


@pytest.mark.parametrize('precision,tol', [(64, 1e-10), (32, 1e-5)])
def test_potential_matches_the_oracle_with_merged_nodes(precision, tol):
    from pcbqs_amd.static import potential
    nodes, tris, z = two_layer(10, 6)
    nt = len(tris)
    node_of = np.arange(nt)
    node_of[:7] = 0                           # a merged pad of seven triangles
    node_of[nt - 5:] = node_of[nt - 6]        # and one at the end
    node_of = np.unique(node_of, return_inverse=True)[1]
    want = oracle_potential(nodes, tris, z, node_of, er=4.4)
    got = potential(nodes, tris, z, node_of, er=4.4, precision=precision)
    assert got.shape == want.shape
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
    assert np.array_equal(got, got.T)


def test_parallel_plate_capacitance_bracket():
    from pcbqs_amd.geometry import EPS0
    from pcbqs_amd.static import potential
    nodes, tris, z = two_layer(12, 8)
    P = potential(nodes, tris, z, np.arange(len(tris)))
    top = z == z.max()
    q = np.linalg.solve(P, np.where(top, 0.5, -0.5))
    C = q[top].sum()
    g = F.prepare(nodes, tris, z, 1)
    A = min(g.area[top].sum(), g.area[~top].sum())
    h = z.max() - z.min()
    assert EPS0 * A / h < C < 1.5 * EPS0 * A / h


def test_capacitance_matrix_matches_the_oracle_and_the_parallel_plate_bracket():
    """C = B^T P^-1 B on the device (P factored as L D L^T, refined in FP64) against the oracle's
    P solved by numpy, and the plates' capacitance inside eps A / h < C < 1.5 eps A / h as the
    reference's floating-pair drive computes it (tests/test_peec.py:25-33)."""
    from pcbqs_amd.geometry import EPS0
    from pcbqs_amd.static import capacitance_matrix
    nodes, tris, z = two_layer(12, 8)
    node_of = np.arange(len(tris))
    labels = np.where(z == z.max(), 0, 1)
    info = {}
    C = capacitance_matrix(nodes, tris, z, node_of, labels, er=2.0, info=info)
    P = oracle_potential(nodes, tris, z, node_of, er=2.0)
    B = np.eye(2)[labels]
    Cref = B.T @ np.linalg.solve(P, B)
    Cref = (Cref + Cref.T) / 2
    assert np.abs(C - Cref).max() <= 1e-9 * np.abs(Cref).max()
    assert info['relative_residual'] <= 1e-12
    drive = np.array([1.0, -1.0])
    pair = 1 / (drive @ np.linalg.solve(C, drive))
    g = F.prepare(nodes, tris, z, 1)
    top = z == z.max()
    A = min(g.area[top].sum(), g.area[~top].sum())
    h = z.max() - z.min()
    assert 2.0 * EPS0 * A / h < pair < 1.5 * 2.0 * EPS0 * A / h
    assert C[0, 0] > 0 and C[1, 1] > 0 and C[0, 1] < 0      # Maxwell signs


@pytest.mark.parametrize('n,nrhs', [(1, 1), (63, 3), (130, 20), (300, 9)])
def test_spd_solve_matches_numpy(n, nrhs):
    """pq_spd_solve on random symmetric positive definite matrices: every ragged panel shape,
    right-hand sides in groups of 16 with padding."""
    import ctypes
    from pcbqs_amd._lib import DeviceArray, call, ptr
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.eye(n)
    Bh = rng.standard_normal((n, nrhs))
    dA = DeviceArray.from_host(np.ascontiguousarray(A))
    X = np.empty((n, nrhs))
    steps, res, fms = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    try:
        call('pq_spd_solve', 0, n, dA.ptr, n, nrhs, ptr(np.ascontiguousarray(Bh)), 1e-12, 12, ptr(X),
             ctypes.byref(steps), ctypes.byref(res), ctypes.byref(fms))
    finally:
        dA.close()
    want = np.linalg.solve(A, Bh)
    assert res.value <= 1e-12
    assert np.abs(X - want).max() <= 1e-10 * np.abs(want).max()


def test_spd_solve_rejects_a_singular_matrix():
    import ctypes
    from pcbqs_amd._lib import DeviceArray, call, ptr
    n = 80
    A = np.eye(n)
    A[40, 40] = 0.0
    dA = DeviceArray.from_host(A)
    X = np.empty((n, 1))
    try:
        with pytest.raises(ValueError, match='pivot'):
            call('pq_spd_solve', 0, n, dA.ptr, n, 1, ptr(np.ones((n, 1))), 1e-12, 12, ptr(X), None, None, None)
    finally:
        dA.close()


def test_board_capacitance_matches_the_oracle():
    """A seeded board's conductors (its copper and vias joined into connected pieces): C from the
    device against the oracle's P solved by numpy with the same conductor labels."""
    from pcbqs_amd import board as boards
    from pcbqs_amd.static import board_capacitance
    b = boards.board_c1()
    C, labels = board_capacitance(b, er=4.4)
    node_of = np.asarray(b.node_of)
    P = oracle_potential(b.nodes, b.tris, b.z_of, node_of, er=4.4)
    B = np.eye(labels.max() + 1)[labels]
    Cref = B.T @ np.linalg.solve(P, B)
    Cref = (Cref + Cref.T) / 2
    print(f'board C1: {b.nt} triangles, {len(labels)} charge nodes, {C.shape[0]} conductors', flush=True)
    assert C.shape == Cref.shape
    assert np.abs(C - Cref).max() <= 1e-9 * np.abs(Cref).max()
    assert (np.diag(C) > 0).all()
