"""Oracle for the graph preconditioner of the iterative loop solve (test infrastructure only; see
oracle/__init__.py).

Restates, with numpy/scipy:
  setup      the preconditioner's part of _prepare (src/pcbqs/mqs_iterative.py:39-74): the node-branch
             incidence A (+1 at a branch's first node, -1 at its second), the passive (non-source)
             branches, each free loop's cotree branch (its non-tree branch of coefficient +1), and the
             nodes kept (all but the first node of each connected piece of the passive graph)
  diagonal   the branch impedance it is built from, d = diag(R) + j w ldiag with the floor at 1e-12 of
             its largest magnitude (src/pcbqs/mqs_iterative.py:337-339)
  graph      (C_f^T D C_f)^-1 applied exactly for a block-diagonal D, by one sparse LU of the graph
             Laplacian A D^-1 A^T per frequency (src/pcbqs/mqs_precond.py:27-59)

Pinned by tests/test_graph_precond_oracle.py: one application inverts C_f^T D C_f formed densely
(the exactness that defines the preconditioner), for diagonal D and for D with coupled branch pairs.
"""
import numpy as np
from scipy.sparse import coo_matrix, csgraph, diags
from scipy.sparse.linalg import LinearOperator, splu


def setup(n_nodes, ends, is_src, C, tree, free):
    ends = np.asarray(ends, np.int64).reshape(-1, 2)
    nb = len(ends)
    a, b = ends[:, 0], ends[:, 1]
    branch = np.arange(nb)
    A = coo_matrix((np.r_[np.ones(nb), -np.ones(nb)], (np.r_[a, b], np.r_[branch, branch])),
                   shape=(n_nodes, nb)).tocsc()
    passive = np.flatnonzero(~np.asarray(is_src, bool))
    C = C.tocsc()
    non_tree = np.flatnonzero(~np.asarray(tree, bool))
    ident = C[non_tree][:, free].tocoo()
    one = ident.data == 1
    cotree = np.empty(len(free), np.int64)
    cotree[ident.col[one]] = non_tree[ident.row[one]]
    g = coo_matrix((np.ones(len(passive)), (a[passive], b[passive])), shape=(n_nodes, n_nodes))
    _, labels = csgraph.connected_components(g, directed=False)
    roots = np.unique(labels, return_index=True)[1]
    keep = np.ones(n_nodes, bool)
    keep[roots] = False
    return dict(A=A, passive=passive, cotree=cotree, keep=keep, pairs=np.zeros((2, 0), np.int64))


def diagonal(Rdiag, ldiag, w):
    d = (np.asarray(Rdiag) + 1j * w * np.asarray(ldiag)).astype(complex)
    scale = max(float(np.max(abs(d))), np.finfo(float).tiny)
    return np.where(abs(d) < scale * 1e-12, scale * 1e-12, d)


def dinv(p, d, zpair):
    """D^-1 restricted to the passive branches (the 2 x 2 blocks of coupled pairs inverted whole)."""
    nb = len(d)
    passive = p['passive']
    inv = np.zeros(nb, complex)
    inv[passive] = 1 / d[passive]
    rows, cols, vals = [np.arange(nb)], [np.arange(nb)], [inv]
    pairs = p['pairs']
    if pairs.shape[1]:
        e, g = pairs
        det = d[e] * d[g] - zpair ** 2
        inv[e], inv[g] = d[g] / det, d[e] / det
        rows += [e, g]
        cols += [g, e]
        vals += [-zpair / det, -zpair / det]
    D = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nb, nb)).tocsr()
    return diags(np.isin(np.arange(nb), passive).astype(float)) @ D


def graph(p, d, zpair=np.zeros(0, complex)):
    A, keep, cotree = p['A'], p['keep'], p['cotree']
    nb = len(d)
    Dinv = dinv(p, d, zpair)
    lap = (A @ Dinv @ A.T).tocsc()[keep][:, keep]
    lu = splu(lap.tocsc())

    def apply(y):
        force = np.zeros((nb,) + np.shape(y)[1:], complex)
        force[cotree] = y
        trial = Dinv @ force
        v = np.zeros((A.shape[0],) + np.shape(y)[1:], complex)
        v[keep] = lu.solve(np.ascontiguousarray(np.asarray(A @ trial)[keep]))
        return (trial - Dinv @ np.asarray(A.T @ v))[cotree]
    return LinearOperator((len(cotree), len(cotree)), matvec=apply, matmat=apply, dtype=complex)
