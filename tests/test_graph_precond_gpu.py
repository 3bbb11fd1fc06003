"""GPU parity of the graph preconditioner (the node Laplacian factored and solved on the device)
against the oracle's restatement of the reference's (scipy splu): one application agrees to 1e-9
on boards with vias, many ports and merged terminals, from DC to 100 MHz, with and without coupled
branch pairs; and GMRES on the device operator converges with it to the dense solution."""
import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, gmres

from oracle import precond as O
from pcbqs_amd import board as boards

pytestmark = pytest.mark.gpu

BOARDS = {'c1': boards.board_c1, 'c2-small': lambda: boards.board_c2(1, target=3000),
          'batch-11-ports': lambda: boards.board_batch(0)}


@pytest.mark.parametrize('name', list(BOARDS))
def test_device_graph_preconditioner_matches_the_oracle(name):
    from pcbqs_amd.precond import branch_diagonal, graph, graph_setup
    from pcbqs_amd.solve import PortSolver
    b = BOARDS[name]()
    s = PortSolver(b, device=0)
    m = s.model
    p = graph_setup(m)
    po = O.setup(b.n_nodes, m.ends, m.is_src, m.basis.C, m.basis.tree, np.arange(m.n_free))
    rng = np.random.default_rng(0)
    y = rng.normal(size=(m.n_free, 3)) + 1j * rng.normal(size=(m.n_free, 3))
    rdiag = m.R.diagonal()
    for f in (0.0, 1e3, 1e6, 1e8):
        w = 2 * np.pi * f
        d = branch_diagonal(rdiag, s.branch_ldiag, w)
        assert np.array_equal(d, O.diagonal(rdiag, s.branch_ldiag, w))
        M, size = graph(p, d)
        want = O.graph(po, d) @ y
        got = M @ y
        print(f'graph {name} f={f:g}: {p["keep"].sum()} nodes, factor {M.laplacian.factor_ms:.2f} ms, '
              f'{size / 1e6:.1f} MB, diff {np.abs(got - want).max() / np.abs(want).max():.2e}', flush=True)
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
        assert np.allclose(M @ y[:, 0], want[:, 0], rtol=1e-9, atol=1e-9 * np.abs(want).max())
        M.laplacian.close()
    # coupled branch pairs (the face-sheet blocks), among branches that carry loop current
    carrying = np.flatnonzero(np.asarray(abs(m.basis.C.tocsr()[:, :m.n_free]).sum(1)).ravel() > 0)
    cand = rng.permutation(np.intersect1d(carrying, p['passive']))[:200]
    pairs = np.stack([cand[:100], cand[100:]])
    d = branch_diagonal(rdiag, s.branch_ldiag, 2 * np.pi * 1e6)
    zpair = 0.2 * d[cand[:100]]
    M, _ = graph(dict(p, pairs=pairs), d, zpair)
    want = O.graph(dict(po, pairs=pairs), d, zpair) @ y
    np.testing.assert_allclose(M @ y, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    M.laplacian.close()
    s.close()


def test_gmres_with_the_device_graph_preconditioner_reaches_the_dense_solution():
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c2(1, target=3000)
    s = PortSolver(b, device=0)
    m = s.model
    f = 1e6
    w = 2 * np.pi * f
    n = m.n_free
    Ll = s.loop_matrix(n)[:, :n].astype(np.float64)
    Z = m.Rl.tocsr()[:n, :n].toarray() + 1j * w * Ll
    rng = np.random.default_rng(3)
    rhs = rng.normal(size=n) + 1j * rng.normal(size=n)
    x_ref = np.linalg.solve(Z, rhs)
    A = LinearOperator((n, n), matvec=lambda x: s.operator(x, f).ravel(), dtype=complex)
    M, _ = s.graph_preconditioner(f)
    counts = {}
    for label, pre in (('none', None), ('graph', M)):
        its = []
        x, info = gmres(A, rhs, M=pre, rtol=1e-9, restart=100, maxiter=50,
                        callback=lambda r: its.append(r), callback_type='pr_norm')
        assert info == 0
        counts[label] = len(its)
        assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    print(f'GMRES iterations at 1 MHz, {n} loops: {counts}', flush=True)
    assert counts['graph'] < counts['none']
    M.laplacian.close()
    s.close()
