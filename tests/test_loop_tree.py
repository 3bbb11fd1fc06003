"""The loop basis's spanning tree, as the device projection reads it (pq_loop_tree): prefix sums of a
symmetric branch matrix's rows down the tree give C^T A C exactly as the sparse loop matrix C does,
for boards with vias and several ports."""
import numpy as np
import pytest

from pcbqs_amd import board as boards
from pcbqs_amd.loops import LoopModel
from pcbqs_amd.solve import tree_arrays


def tree_projection(A, nbr, arrays, n_nodes):
    """C^T A C from the tree arrays, by the device kernels' formulas (rows >= nbr carry nothing)."""
    parent, upb, ups, level_ptr, level_nodes, cot, ea, eb = arrays

    def prefix(R):
        S = np.zeros((n_nodes, R.shape[1]))
        for d in range(len(level_ptr) - 1):
            for x in level_nodes[level_ptr[d]:level_ptr[d + 1]]:
                S[x] = S[parent[x]]
                if 0 <= upb[x] < nbr:
                    S[x] += ups[x] * R[upb[x]]
        return S

    def close(R, S):
        out = S[ea] - S[eb]
        ok = cot < nbr
        out[ok] += R[cot[ok]]
        return out

    U = close(A, prefix(A)).T               # U[k][j] = A[cot j][k] + S_k(a_j) - S_k(b_j): A symmetric
    return close(U, prefix(U))


@pytest.mark.parametrize('make', [lambda: boards.board_c1(target=400), lambda: boards.board_batch(3, (300, 500)),
                                  lambda: boards.board_batch(0, (800, 1200))])
def test_tree_prefix_sums_give_the_loop_projection(make):
    b = make()
    m = LoopModel(b)
    nbr = m.ne + m.nv
    _, arrays = tree_arrays(m.basis)
    rng = np.random.default_rng(0)
    A = rng.standard_normal((nbr, nbr))
    A = A + A.T
    C = m.basis.C.tocsc()[:nbr]
    ref = (C.T @ (C.T @ A).T).T
    got = tree_projection(A, nbr, arrays, b.n_nodes)
    assert got.shape == (m.n_loops, m.n_loops)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_tree_levels_list_every_non_root_after_its_parent():
    b = boards.board_batch(2, (300, 500))
    m = LoopModel(b)
    tree, (parent, upb, ups, level_ptr, level_nodes, cot, ea, eb) = tree_arrays(m.basis)
    seen = parent < 0
    assert level_ptr[0] == 0 and level_ptr[-1] == len(level_nodes) == (~seen).sum()
    for d in range(tree.n_levels):
        lvl = level_nodes[level_ptr[d]:level_ptr[d + 1]]
        assert seen[parent[lvl]].all() and not seen[lvl].any()
        seen[lvl] = True
    assert seen.all()
    assert set(np.unique(ups[parent >= 0])) <= {-1.0, 1.0}


def test_automatic_solver_choice_follows_the_reference_rule():
    """solver='auto': direct while the free loops number at most min(8000, 60 x ports x frequencies)
    (src/pcbqs/mqs_iterative.py:260-272)."""
    from pcbqs_amd.solve import port_method
    assert port_method(2446, 16, 20) == 'direct'
    assert port_method(8000, 8, 20) == 'direct' and port_method(8001, 8, 20) == 'iterative'
    assert port_method(240, 4, 1) == 'direct' and port_method(241, 4, 1) == 'iterative'
    assert port_method(9639, 4, 1) == 'iterative' and port_method(32211, 8, 100) == 'iterative'


@pytest.mark.parametrize('make', [lambda: boards.board_c1(target=600), lambda: boards.board_batch(3, (300, 500)),
                                  lambda: boards.board_batch(5, (1500, 2500))])
def test_library_topology_equals_the_numpy_construction(make):
    """pq_loop_basis_create / pq_sheet_resistance / pq_loop_resistance against loop_basis,
    sheet_resistance and scipy's C^T R C: the same tree and loops bitwise, every loop a cycle of the
    branch graph, Re bitwise, Rl exactly symmetric and within rounding of the scipy product."""
    from scipy import sparse
    b = make()
    n, p = LoopModel(b, native=True), LoopModel(b)
    a, q = n.basis, p.basis
    assert a.n_free == q.n_free and np.array_equal(n.port_cols, p.port_cols)
    for name in ('tree', 'parent', 'upb', 'ups', 'depth', 'cotree', 'positions', 'cot_ends'):
        assert np.array_equal(getattr(a, name), getattr(q, name)), name
    assert (a.C != q.C).nnz == 0 and np.array_equal(a.C.indptr, q.C.indptr) and np.array_equal(a.C.indices, q.C.indices)
    # every column of C is a cycle: the node-branch incidence annihilates it
    ends = n.ends
    nb = len(ends)
    inc = sparse.csr_matrix((np.r_[np.ones(nb), -np.ones(nb)], (np.r_[ends[:, 0], ends[:, 1]],
                                                                  np.r_[np.arange(nb), np.arange(nb)])),
                            shape=(b.n_nodes, nb))
    assert abs(inc @ a.C).max() == 0
    assert (n.Re != p.Re).nnz == 0
    assert (n.Rl != n.Rl.T).nnz == 0
    ref = (a.C.T @ n.R @ a.C).tocsr()
    assert abs(n.Rl - ref).max() <= 1e-13 * abs(ref).max()
