"""Oracle for the surface-charge (PEEC) solve (test infrastructure only; see oracle/__init__.py).

Restates, in numpy / scipy with dense FP64 matrices, for a copper model with current sources and
no passive parts or frequency-dependent resistance:
  prepare       src/pcbqs/peec.py:164-201 (incidence A, basis K = [loops | tree], T = A K_tree,
                pscale, the electrostatic block T^T P T / pscale, the projected R and L, the copper
                components, the tree's sparse LU)
  system        src/pcbqs/kernels.py:63-78 (peec_system: Z / zscale on the loops' columns,
                j w Z / pscale on the tree's, the electrostatic block added)
  solve         src/pcbqs/peec.py:204-292 (the components' equilibrium, the tree baseline, one dense
                LU solve per frequency, currents, charges, potentials, source impedances)
  potential     src/pcbqs/mqs.py:532-556 (P = W^T M0 W / (4 pi eps0 er), M0 the oracle's scalar fill)
on the oracle's own fill, combine and loop basis (oracle/fill.py, oracle/mqs.py).

Pinned by: tests/test_peec_oracle.py: two plates without copper between them carry displacement
current only, and the capacitance that impedance gives equals 1 / (d^T C^-1 d) of the Maxwell
capacitance matrix C = B^T P^-1 B (tests/test_peec.py:21-30 of the reference; 2e-5) inside the
parallel-plate bracket; KCL (j w q + A i = 0) and potential = P q at every node; total charge zero;
linearity in the source amplitude; the impedance scales as 1 / er; a closed copper path's PEEC
impedance approaches the inductive solve's (tests/test_peec.py:86-101; 1e-5).
"""
import numpy as np
from scipy import sparse
from scipy.linalg import lu_factor, lu_solve
from scipy.sparse.csgraph import connected_components
from scipy.sparse.linalg import splu

from . import fill as F
from . import mqs as M

MU0 = 4e-7 * np.pi
EPS0 = 1.0 / (MU0 * 299792458.0 ** 2)


def copper_model(board, ideal_terminals=True):
    return M.CopperModel(board.nodes, board.tris, board.z_of, board.t_of, board.node_of, board.terminals,
                         board.vseg, board.via_nodes, board.ports, ideal_terminals=ideal_terminals)


def potential(q, g, M0, er=1.0):
    n = q.n_nodes
    area = np.bincount(q.node_of, weights=g.area, minlength=n)
    W = np.zeros((len(q.tris), n))
    W[np.arange(len(q.tris)), q.node_of] = 1.0 / area[q.node_of]
    return W.T @ M0 @ W / (4 * np.pi * EPS0 * er)


def prepare(board, er=1.0, ideal_terminals=True):
    q = copper_model(board, ideal_terminals)
    g, m0, mx, my, rr = q.fill()
    mats = q.assemble((g, m0, mx, my, rr))
    P = potential(q, g, m0, er)
    ends = mats['ends']
    nb = len(ends)
    branch = np.arange(nb)
    a, b = np.asarray(ends, int).reshape(-1, 2).T
    A = sparse.coo_matrix((np.r_[np.ones(nb), -np.ones(nb)], (np.r_[a, b], np.r_[branch, branch])),
                          shape=(q.n_nodes, nb)).tocsc()
    C, src_cols, is_src, tree = mats['C'], mats['src'], mats['is_src'], mats['tree']
    passive, sources = np.flatnonzero(~is_src), np.flatnonzero(is_src)
    cycles = [k for k in range(C.shape[1]) if k not in src_cols.values()]
    K = sparse.hstack([C[passive][:, cycles], sparse.identity(nb, format='csc')[passive][:, np.flatnonzero(tree)]],
                      format='csc')
    nl = len(cycles)
    T = A[:, passive] @ K[:, nl:]
    pscale = np.max(np.abs(P))
    electro = M.qform(T, P, T) / pscale
    kc = K.tocoo()
    full = sparse.coo_matrix((kc.data, (passive[kc.row], kc.col)), shape=(nb, K.shape[1])).tocsc()
    R = (full.T @ mats['R'] @ full).toarray()
    ne, nv = q.ne, q.nv
    L = M.qform(full[:ne], mats['Le'], full[:ne])
    if nv:
        L = L + M.qform(full[ne:ne + nv], mats['Lv'], full[ne:ne + nv])
    graph = sparse.coo_matrix((np.ones(len(passive)), (a[passive], b[passive])), shape=(q.n_nodes, q.n_nodes))
    nc, labels = connected_components(graph, directed=False)
    B = np.eye(nc)[labels]
    roots = np.unique(labels, return_index=True)[1]
    keep = np.ones(q.n_nodes, bool)
    keep[roots] = False
    tree_lu = splu(T[keep].astype(complex).tocsc()) if T.shape[1] else None
    names = [name for name, _, _ in board.ports]
    return dict(A=A, K=K, T=T, P=P, pscale=pscale, electro=np.asarray(electro), R=R, L=np.asarray(L), nl=nl,
                passive=passive, sources=sources, B=B, keep=keep, tree_lu=tree_lu, names=names,
                source_components=A[:, sources].T @ B, n_nodes=q.n_nodes, nb=nb)


def system(R, L, jw, pscale, nl, electro):
    Z = R + jw * L
    zscale = max(np.max(np.abs(Z)), np.finfo(float).tiny) if Z.size else np.finfo(float).tiny
    S = Z / zscale
    S[:, nl:] = jw * Z[:, nl:] / pscale
    S[nl:, nl:] += electro
    return S, float(zscale)


def solve(m, freqs, amps=None):
    """(currents (F, nb), charge (F, n_nodes), potential (F, n_nodes), impedance {name: (F,)})."""
    f = np.atleast_1d(np.asarray(freqs, float))
    A, K, T, P, B = m['A'], m['K'], m['T'], m['P'], m['B']
    nl, pscale, passive, sources, names = m['nl'], m['pscale'], m['passive'], m['sources'], m['names']
    excitation = np.array([(amps or {}).get(n, 1.0) for n in names], complex)
    drive = np.asarray(A[:, sources] @ excitation).ravel()
    totals = m['source_components'].T @ excitation
    qpart = np.zeros(m['n_nodes'], complex)
    vpart = np.zeros(B.shape[1], complex)
    if np.any(totals):
        PinvB = np.linalg.solve(P, B)
        vpart = np.linalg.solve(B.T @ PinvB, -totals)
        qpart = PinvB @ vpart
    y0 = m['tree_lu'].solve((-drive - qpart)[m['keep']]) if m['tree_lu'] is not None else np.zeros(0, complex)
    baseline = np.r_[np.zeros(nl, complex), y0]
    source_voltage = -m['source_components'] @ vpart
    rb, lb = m['R'] @ baseline, m['L'] @ baseline
    currents = np.zeros((len(f), m['nb']), complex)
    charge = np.zeros((len(f), m['n_nodes']), complex)
    potential_ = np.zeros_like(charge)
    impedance = {n: np.zeros(len(f), complex) for n in names}
    for fi, freq in enumerate(f):
        jw = 2j * np.pi * freq
        if K.shape[1]:
            S, zscale = system(m['R'], m['L'], jw, pscale, nl, m['electro'])
            x = lu_solve(lu_factor(S), -(rb + jw * lb) / zscale)
            alpha = zscale / pscale
            currents[fi, passive] = K[:, :nl] @ x[:nl] + K[:, nl:] @ (y0 + jw * alpha * x[nl:])
            charge[fi] = qpart / jw - alpha * (T @ x[nl:])
            potential_[fi] = B @ vpart / jw - alpha * (P @ (T @ x[nl:]))
            voltage = source_voltage / jw + alpha * (A[:, sources].T @ (P @ (T @ x[nl:])))
        else:
            charge[fi] = qpart / jw
            potential_[fi] = B @ vpart / jw
            voltage = source_voltage / jw
        currents[fi, sources] = excitation
        for j, name in enumerate(names):
            impedance[name][fi] = voltage[j] / excitation[j] if excitation[j] else np.nan
    return currents, charge, potential_, impedance


def capacitance(m):
    """Maxwell C = B^T P^-1 B of the copper components (src/pcbqs/peec.py:95-123), symmetrised."""
    C = m['B'].T @ np.linalg.solve(m['P'], m['B'])
    return (C + C.T) / 2
