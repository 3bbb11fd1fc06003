"""The sparse hat construction of the coarse space equals a direct dense evaluation of the same
hats (every hat over every node), bitwise, on one- and multi-layer boards and grid sizes; the
normalised loop-space columns agree to rounding."""
import numpy as np
import pytest

from pcbqs_amd import board as boards
from pcbqs_amd.coarse import boundary_vertices, edge_signs, loop_restrict, stream_patterns
from pcbqs_amd.loops import LoopModel


def dense_hats(model, grid):
    b, r = model.board, model.rwg
    s = edge_signs(r)
    u, v = r.ends[:, 0], r.ends[:, 1]
    interior = ~boundary_vertices(b.tris, len(b.nodes))
    cols = []
    for lay in np.unique(b.layer_of):
        ln = np.unique(b.tris[b.layer_of == lay])
        x0, y0 = b.nodes[ln].min(0)
        x1, y1 = b.nodes[ln].max(0)
        gx, gy = np.linspace(x0, x1, grid + 1), np.linspace(y0, y1, grid + 1)
        hx, hy = max(gx[1] - gx[0], 1e-30), max(gy[1] - gy[0], 1e-30)
        on = np.zeros(len(b.nodes), bool)
        on[ln] = True
        edges_on = on[u] & on[v]
        on &= interior
        for i in range(grid + 1):
            for j in range(grid + 1):
                if i == grid and j == grid:
                    continue
                psi = (np.maximum(0, 1 - np.abs(b.nodes[:, 0] - gx[i]) / hx) *
                       np.maximum(0, 1 - np.abs(b.nodes[:, 1] - gy[j]) / hy) * on)
                ib = np.zeros(model.nb)
                ib[:model.ne] = np.where(edges_on, s * (psi[v] - psi[u]), 0.0)
                if np.abs(ib).max() > 0:
                    cols.append(ib)
    return np.array(cols).T


@pytest.mark.parametrize('name,grid', [('c1', 4), ('c2', 6), ('c2', 11), ('batch', 5), ('batch', 8)])
def test_sparse_hats_equal_the_dense_evaluation(name, grid):
    b = {'c1': lambda: boards.board_c1(), 'c2': lambda: boards.board_c2(2, target=3000),
         'batch': lambda: boards.board_batch(3)}[name]()
    m = LoopModel(b)
    sp = stream_patterns(m, grid)
    dn = dense_hats(m, grid)
    assert sp.shape == dn.shape
    assert np.array_equal(sp.toarray(), dn)
    # the columns' norms are summed in a different order: equal to rounding
    np.testing.assert_allclose(loop_restrict(m, sp), loop_restrict(m, dn), rtol=1e-13, atol=1e-15)


def test_sparse_loop_columns_equal_the_dense_ones():
    """loop_restrict and via_loop_columns give the same columns as scipy CSC (dense=False)."""
    from pcbqs_amd.coarse import via_loop_columns
    b = boards.board_c2(2, target=3000)
    m = LoopModel(b)
    sp = stream_patterns(m, 8)
    assert np.array_equal(loop_restrict(m, sp, dense=False).toarray(), loop_restrict(m, sp))
    assert np.array_equal(via_loop_columns(m, dense=False).toarray(), via_loop_columns(m))


def test_unique_columns_drops_repeats_up_to_sign_and_rounding():
    """A column equal to an earlier one up to sign and 1e-12 is dropped; a column with the same
    rows but other values, or other rows, is kept; the order of the kept columns is unchanged."""
    from scipy import sparse
    from pcbqs_amd.coarse import unique_columns
    rng = np.random.default_rng(3)
    a = np.zeros((20, 6))
    a[[1, 4, 7], 0] = rng.normal(size=3)
    a[:, 1] = -a[:, 0] * (1 + 1e-14)              # a repeat up to sign and rounding
    a[[1, 4, 7], 2] = rng.normal(size=3)          # same rows, other values
    a[[2, 4, 7], 3] = a[[1, 4, 7], 0]             # same values, other rows
    a[:, 4] = a[:, 2]                             # an exact repeat
    a[[0, 19], 5] = 1.0
    a /= np.linalg.norm(a, axis=0)
    u = unique_columns(sparse.csc_matrix(a))
    assert np.array_equal(u.toarray(), a[:, [0, 2, 3, 5]])
    # a matrix without repeats comes back unchanged
    assert np.array_equal(unique_columns(sparse.csc_matrix(a[:, [0, 2, 3, 5]])).toarray(), a[:, [0, 2, 3, 5]])
