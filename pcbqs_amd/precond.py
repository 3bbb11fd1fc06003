"""The graph preconditioner of the iterative loop solve, with its Laplacian solve on the device.

graph(p, d, zpair) is the drop-in for mqs_precond.graph (src/pcbqs/mqs_precond.py:27-59): the
loop system C_f^T D C_f inverted exactly for a block-diagonal branch impedance D (each branch's
own d, and the 2 x 2 block of each coupled branch pair). Under a force y on the free loops'
cotree branches the branch currents are i = D^-1 (f - A^T v) with Kirchhoff's current law A i = 0,
so the node potentials solve (A D^-1 A^T) v = A D^-1 f, one node grounded per connected piece of
copper. That Laplacian is complex symmetric with a positive definite real part; it is factored
once per frequency on the device (pq_csym_create: FP32 L D L^T, no pivoting) and every
application solves it there with FP64 refinement to 1e-12 (pq_csym_solve), in place of SuperLU.
The incidence products around the solve are O(branches) sparse host work, as in the reference.
"""
import ctypes

import numpy as np
from scipy import sparse
from scipy.sparse import csgraph
from scipy.sparse.linalg import LinearOperator

from ._lib import call, ptr

REFINE_TOL, REFINE_STEPS = 1e-12, 12


def graph_setup(model):
    """A (nodes x branches incidence, +1 at a branch's first end), the passive branches, the free
    loops' cotree branches (device loop order) and the kept nodes, for a LoopModel
    (src/pcbqs/mqs_iterative.py:39-74)."""
    ends = np.asarray(model.ends, np.int64)
    nb, n_nodes = len(ends), model.board.n_nodes
    k = np.arange(nb)
    A = sparse.csc_matrix((np.r_[np.ones(nb), -np.ones(nb)], (np.r_[ends[:, 0], ends[:, 1]], np.r_[k, k])),
                          shape=(n_nodes, nb))
    passive = np.flatnonzero(~model.is_src)
    pe = ends[passive]
    copper = sparse.coo_matrix((np.ones(len(pe)), (pe[:, 0], pe[:, 1])), shape=(n_nodes, n_nodes))
    _, piece = csgraph.connected_components(copper, directed=False)
    grounded = np.unique(piece, return_index=True)[1]             # the first node of each piece
    keep = np.ones(n_nodes, bool)
    keep[grounded] = False
    cotree = np.asarray(model.basis.cotree[:model.n_free], np.int64)
    return dict(A=A, passive=passive, cotree=cotree, keep=keep, pairs=np.zeros((2, 0), np.int64))


def branch_diagonal(Rdiag, ldiag, w):
    """d = diag(R) + j w diag(L), with magnitudes floored at 1e-12 of the largest
    (src/pcbqs/mqs_iterative.py:337-339)."""
    d = np.asarray(Rdiag, float) + 1j * w * np.asarray(ldiag, float)
    floor = max(float(np.abs(d).max()), np.finfo(float).tiny) * 1e-12
    return np.where(np.abs(d) < floor, floor, d)


def _inverse_impedance(p, d, zpair):
    """D^-1 (sparse), coupled pairs' 2 x 2 blocks inverted whole, its rows kept on the passive
    branches only (a source branch carries no current of its own)."""
    nb = len(d)
    inv = 1 / d
    r, c, v = [np.arange(nb)], [np.arange(nb)], []
    e, g = p['pairs']
    if len(e):
        det = d[e] * d[g] - zpair ** 2
        inv[e], inv[g] = d[g] / det, d[e] / det
        r += [e, g]
        c += [g, e]
        v = [-zpair / det] * 2
    on = np.zeros(nb)
    on[p['passive']] = 1.0
    D = sparse.csr_matrix((np.concatenate([inv] + v), (np.concatenate(r), np.concatenate(c))), shape=(nb, nb))
    return sparse.diags(on) @ D


# The Laplacian is factored dense (FP32 complex, n^2 x 8 bytes: 12.8 GB at this bound, and O(n^3)
# work per frequency), where the reference's SuperLU is sparse; larger node counts are refused.
MAX_LAPLACIAN_NODES = 40000


class _DeviceLaplacian:
    """The kept-node Laplacian factored on the device (pq_csym_*)."""

    def __init__(self, lap, device):
        if lap.shape[0] > MAX_LAPLACIAN_NODES:
            raise ValueError(f'graph preconditioner: {lap.shape[0]:,} kept nodes exceed the dense device '
                             f'factor\'s limit of {MAX_LAPLACIAN_NODES:,} (use block-Jacobi or two-level)')
        low = sparse.tril(lap, format='coo')
        rows, cols = np.ascontiguousarray(low.row, np.int32), np.ascontiguousarray(low.col, np.int32)
        vals = np.ascontiguousarray(low.data, np.complex128)
        self.n = lap.shape[0]
        self._h = None
        self.factor_ms, self.bytes = 0.0, 0
        if not self.n:
            return
        self._h = ctypes.c_void_p()
        ms, nbytes = ctypes.c_double(), ctypes.c_int64()
        call('pq_csym_create', device, self.n, len(rows), ptr(rows), ptr(cols), ptr(vals), ctypes.byref(self._h),
             ctypes.byref(ms), ctypes.byref(nbytes))
        self.factor_ms, self.bytes = ms.value, nbytes.value

    def solve(self, b):
        b2 = np.ascontiguousarray(np.asarray(b, np.complex128).reshape(self.n, -1))
        x = np.empty_like(b2)
        steps, res = ctypes.c_int64(), ctypes.c_double()
        call('pq_csym_solve', self._h, b2.shape[1], ptr(b2), REFINE_TOL, REFINE_STEPS, ptr(x), ctypes.byref(steps),
             ctypes.byref(res))
        return x.reshape(np.shape(b))

    def close(self):
        from . import _lib
        if getattr(self, '_h', None) is not None and _lib._lib is not None:
            _lib._lib.pq_csym_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass


def graph(p, d, zpair=None, device=0):
    """(M, factor bytes): M applies (C_f^T D C_f)^-1 to (loops,) or (loops, k) arrays."""
    A, keep, cotree = p['A'], p['keep'], p['cotree']
    nb = len(d)
    d = np.asarray(d, complex)
    zpair = np.zeros(0, complex) if zpair is None else np.asarray(zpair, complex)
    Dinv = _inverse_impedance(p, d, zpair)
    lap = (A @ Dinv @ A.T).tocsr()[keep][:, keep]
    lu = _DeviceLaplacian(lap, device)

    def apply(y):
        y = np.asarray(y)
        force = np.zeros((nb,) + y.shape[1:], complex)
        force[cotree] = y
        trial = Dinv @ force
        v = np.zeros((A.shape[0],) + y.shape[1:], complex)
        if lu.n:
            v[keep] = lu.solve(np.asarray(A @ trial)[keep])
        return (trial - Dinv @ np.asarray(A.T @ v))[cotree]
    M = LinearOperator((len(cotree), len(cotree)), matvec=apply, matmat=apply, dtype=complex)
    M.block = apply
    M.laplacian = lu
    return M, lu.bytes
