"""Host side of the learned correction: the branch graph and the placement of learned columns."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')


def test_branch_graph_and_column_placement():
    from pcbqs_amd import board as boards
    from pcbqs_amd.learned import K_CHANNELS, N_IN, BranchGraph, learned_columns
    from pcbqs_amd.loops import LoopModel
    m = LoopModel(boards.board_c2(3, target=3000))
    g = BranchGraph(m)
    assert g.static.shape == (m.nb, N_IN) and np.isfinite(g.static).all()
    # mean aggregation: rows of A sum to one (every branch has a neighbour), A^T is its transpose
    assert np.allclose(np.asarray(g.A.sum(1)).ravel(), 1.0)
    assert abs(g.AT - g.A.T).max() == 0
    # symmetric structure: neighbours both ways
    assert ((g.A > 0) != (g.A.T > 0)).nnz == 0
    assert g.n_cells == m.nv + m.np and g.cell.max() < g.n_cells
    E = torch.arange(m.nb * 16, dtype=torch.float64).reshape(m.nb, 16)
    Phi = learned_columns(torch, g, E).numpy()
    assert Phi.shape == (m.n_free, g.n_cells * K_CHANNELS)
    for l in (0, m.n_free // 2, m.n_free - 1):
        b = g.cot[l]
        nz = np.flatnonzero(Phi[l])
        cols = g.cell[b] * K_CHANNELS + np.arange(K_CHANNELS)
        assert np.array_equal(nz[nz >= 0], cols[E[b, :K_CHANNELS].numpy() != 0])
        assert np.array_equal(Phi[l, cols], E[b, :K_CHANNELS].numpy())
