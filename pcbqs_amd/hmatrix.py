"""Device product of a block-compressed (H-matrix) operator: the drop-in for HMatrix's product.

`DeviceHMatrix(blocks, n)` takes the reference's block list (src/pcbqs/hmatrix.py:80-83: tuples
(rows a, columns b, U, V or None, same), each off-diagonal block standing for its transpose too)
and keeps it in HBM; `H @ X` runs the deterministic two-phase product of HMatrix._matmat /
hmatrix_product (src/pcbqs/hmatrix.py:223-297, src/pcbqs/core/hmatrix_kernels.hpp:20-101) on the
device (pq_hmatrix_apply). As in the reference, a real operator applies to complex vectors as
independent real and imaginary columns, and the adjoint of the transpose-symmetric operator is
conj(A conj(x)). Low-rank blocks are limited to rank 256 (the reference's ACA stops at 128).

`RwgEntries(nodes, tris, z)` and `build(entries)` construct the compressed RWG inductance itself on
the device, without a dense matrix: the entry provider of edge_inductance (src/pcbqs/compressed.py:
59-73), HMatrix's level-by-level native construction (src/pcbqs/hmatrix.py:88-95,179-210) with its
tree and admissibility on the host, and the checked ACA and dense leaves of aca_blocks /
dense_blocks (src/pcbqs/core/hmatrix_build.hpp:143-352) as device kernels (pq_aca_blocks,
pq_dense_blocks) whose factors stay in HBM and become the product handle directly.
"""
import argparse
import ctypes
import json
import time

import numpy as np
from scipy.sparse.linalg import LinearOperator

from ._lib import call, ptr


class _Cluster:
    def __init__(self, ids, lo, hi):
        self.ids, self.lo, self.hi, self.children = ids, lo, hi, ()


def _tree(ids, lo, hi, leaf):
    """Bisection along the widest spread of the support-box centres (src/pcbqs/hmatrix.py:88-95)."""
    node = _Cluster(ids, lo[ids].min(0), hi[ids].max(0))
    if len(ids) > leaf:
        centres = (lo[ids] + hi[ids]) / 2
        axis = int(np.argmax(np.ptp(centres, axis=0)))
        order = ids[np.argsort(centres[:, axis], kind='stable')]
        mid = len(ids) // 2
        node.children = (_tree(order[:mid], lo, hi, leaf), _tree(order[mid:], lo, hi, leaf))
    return node


def _low_rank(block, tol, rng):
    """U V with |block - U V|_F <= tol |block|_F and rank <= min(128, 0.7 m n / (m + n)), or None.

    Truncated SVD; above 256 rows and columns the SVD is taken of the block's projection on a
    randomised range of cap + 16 columns (Halko, Martinsson & Tropp 2011, algorithm 4.1)."""
    m, n = block.shape
    cap = min(128, int(.7 * m * n / (m + n)))
    if cap < 1:
        return None
    if min(m, n) > 256:
        om = rng.standard_normal((n, cap + 16))
        if block.dtype.kind == 'c':
            om = om + 1j * rng.standard_normal((n, cap + 16))
        q, _ = np.linalg.qr(block @ om)
        w, s, vt = np.linalg.svd(q.conj().T @ block, full_matrices=False)
        u = q @ w
        resid2 = max(np.linalg.norm(block) ** 2 - np.sum(s ** 2), 0.0)
    else:
        u, s, vt = np.linalg.svd(block, full_matrices=False)
        resid2 = 0.0
    tail = np.sqrt(np.r_[np.cumsum((s ** 2)[::-1])[::-1], 0.0] + resid2)    # |block - rank-r part|
    scale = max(np.linalg.norm(block), np.finfo(float).tiny)
    ok = tail <= tol * scale
    if not ok.any():
        return None
    r = int(np.argmax(ok))
    if r > cap:
        return None
    return np.ascontiguousarray(u[:, :r] * s[:r]), np.ascontiguousarray(vt[:r])


def compress(entry, lo, hi, tol=1e-6, leaf=64, eta=1.0, dtype=float, seed=0):
    """Blocks (a, b, U, V, same) of a transpose-symmetric matrix given by entry(rows, cols), on the
    reference's cluster tree and admissibility rule (src/pcbqs/hmatrix.py:88-95, 152-177): far
    blocks (gap > eta * diameter) low rank when that is economical, diagonal leaves symmetrised,
    everything else subdivided down to dense leaf pairs. The factors come from truncated SVDs to
    the relative tolerance, where the reference runs checked partial-pivot ACA (hmatrix.py:103-150).
    A host-side construction, as the reference's; the product is the device part."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    rng = np.random.default_rng(seed)
    blocks = []
    if len(lo):
        root = _tree(np.arange(len(lo)), lo, hi, leaf)
        pending = [(root, root)]
    else:
        pending = []
    while pending:
        a, b = pending.pop()
        same = a is b
        if _admissible(a, b, eta):
            f = _low_rank(np.asarray(entry(a.ids, b.ids), dtype), tol, rng)
            if f is not None:
                blocks.append((a.ids, b.ids, f[0], f[1], False))
                continue
        if not a.children and not b.children:
            block = np.asarray(entry(a.ids, b.ids), dtype)
            if same:
                block = (block + block.T) / 2
            blocks.append((a.ids, b.ids, np.ascontiguousarray(block), None, same))
            continue
        pending.extend(reversed(_children(a, b)))
    return blocks


def _admissible(a, b, eta):
    """Far enough apart for a low-rank block: gap > eta * the larger diameter (hmatrix.py:154-156,185-187)."""
    gap = np.linalg.norm(np.maximum(0, np.maximum(a.lo - b.hi, b.lo - a.hi)))
    diameter = max(np.linalg.norm(a.hi - a.lo), np.linalg.norm(b.hi - b.lo))
    return a is not b and gap > eta * diameter


def _children(a, b):
    """The pairs a non-admissible, non-leaf pair splits into (hmatrix.py:172-177)."""
    if a is b:
        x, y = a.children
        return [(x, x), (x, y), (y, y)]
    if a.children and (not b.children or len(a.ids) >= len(b.ids)):
        return [(c, b) for c in a.children]
    return [(a, c) for c in b.children]


class DeviceHMatrix(LinearOperator):
    def __init__(self, blocks, n, dtype=None, device=0):
        blocks = list(blocks)
        complex_ = (np.dtype(dtype).kind == 'c' if dtype is not None else False) or any(
            np.asarray(u).dtype.kind == 'c' or (v is not None and np.asarray(v).dtype.kind == 'c')
            for _, _, u, v, _ in blocks)
        et = np.complex128 if complex_ else np.float64
        super().__init__(np.dtype(dtype if dtype is not None else et), (int(n), int(n)))
        meta = np.zeros((len(blocks), 6), np.int64)
        idx, factors, io = [], [], 0
        for q, (a, b, u, v, same) in enumerate(blocks):
            a, b = np.asarray(a, np.int64).ravel(), np.asarray(b, np.int64).ravel()
            u = np.asarray(u)
            if v is None:
                if u.shape != (len(a), len(b)):
                    raise ValueError('H-matrix block layout shape: dense block must be |a| x |b|')
                rank = -1
                factors.append(np.ascontiguousarray(u, et).ravel())
            else:
                v = np.asarray(v)
                if u.ndim != 2 or v.ndim != 2 or u.shape[0] != len(a) or v.shape != (u.shape[1], len(b)):
                    raise ValueError('H-matrix block layout shape: low-rank block must be |a| x r times r x |b|')
                rank = u.shape[1]
                factors += [np.ascontiguousarray(u, et).ravel(), np.ascontiguousarray(v, et).ravel()]
            if same and len(a) != len(b):
                raise ValueError('H-matrix block layout: a diagonal block must be square')
            meta[q] = (io, len(a), io + len(a), len(b), rank, int(bool(same)))
            idx += [a, b]
            io += len(a) + len(b)
        self._idx = np.ascontiguousarray(np.concatenate(idx) if idx else np.zeros(0), np.int32)
        if len(self._idx) and (self._idx.min() < 0 or self._idx.max() >= n):
            raise ValueError('H-matrix block layout: index out of range')
        f = np.ascontiguousarray(np.concatenate(factors) if factors else np.zeros(0), et)
        self.complex_ = complex_
        self.device = device
        self._handle = None
        h = ctypes.c_void_p()
        call('pq_hmatrix_create', device, int(n), len(blocks), ptr(meta), ptr(self._idx), len(self._idx), ptr(f),
             len(f), int(complex_), ctypes.byref(h))
        self._handle = h
        info = (ctypes.c_int64 * 5)()
        call('pq_hmatrix_info', h, info)
        self.info = dict(blocks=len(blocks), low_rank_blocks=int(info[4]), dense_blocks=int(info[3]),
                         slots=int(info[1]), bytes=int(info[2]))
        self.storage_bytes = int(info[2])

    @classmethod
    def _adopt(cls, handle, n, device):
        """The operator around a product handle built on the device (real FP64)."""
        self = cls.__new__(cls)
        LinearOperator.__init__(self, np.dtype(np.float64), (int(n), int(n)))
        self.complex_, self.device, self._handle = False, device, handle
        info = (ctypes.c_int64 * 5)()
        call('pq_hmatrix_info', handle, info)
        self.info = dict(blocks=int(info[3] + info[4]), low_rank_blocks=int(info[4]), dense_blocks=int(info[3]),
                         slots=int(info[1]), bytes=int(info[2]))
        self.storage_bytes = int(info[2])
        return self

    @classmethod
    def from_hmatrix(cls, H, device=0):
        """The device copy of a reference HMatrix (its `blocks`, `shape` and `dtype`)."""
        return cls(H.blocks, H.shape[0], H.dtype, device)

    def _apply(self, x):
        """x (n, k) in the operator's element type, C order."""
        y = np.empty_like(x)
        call('pq_hmatrix_apply', self._handle, x.shape[1], ptr(x), ptr(y))
        return y

    def _matmat(self, x):
        x = np.asarray(x)
        n, k = x.shape
        out_dtype = np.result_type(self.dtype, x.dtype)
        if self.complex_:
            y = self._apply(np.ascontiguousarray(x, np.complex128))
        elif x.dtype.kind == 'c':
            # a real operator acts on the real and imaginary parts as independent columns
            xr = np.ascontiguousarray(x, np.complex128).view(np.float64)
            y = self._apply(np.ascontiguousarray(xr)).view(np.complex128)
        else:
            y = self._apply(np.ascontiguousarray(x, np.float64))
        return y.astype(out_dtype, copy=False)

    def _matvec(self, x):
        return self._matmat(np.asarray(x).reshape(-1, 1)).ravel()

    def _adjoint(self):
        return LinearOperator(self.shape, matvec=lambda x: np.conj(self._matvec(np.conj(x))),
                              matmat=lambda x: np.conj(self._matmat(np.conj(x))), dtype=self.dtype)

    def time_product(self, k, reps=20):
        """Mean device milliseconds of one product with k columns (HBM-resident vectors)."""
        x = np.ones((self.shape[0], k), np.complex128 if self.complex_ else np.float64)
        self._apply(x)
        ms = ctypes.c_double()
        call('pq_hmatrix_time', self._handle, int(reps), ctypes.byref(ms))
        return ms.value

    def close(self):
        from . import _lib
        if getattr(self, '_handle', None) is not None and _lib._lib is not None:
            _lib._lib.pq_hmatrix_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass


def _lists(rows, cols):
    """(offsets, concatenated int32 ids) of the row clusters and of the column clusters."""
    rp = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int64)
    cp = np.r_[0, np.cumsum([len(c) for c in cols])].astype(np.int64)
    return (rp, np.ascontiguousarray(np.concatenate(rows), np.int32),
            cp, np.ascontiguousarray(np.concatenate(cols), np.int32))


class RwgEntries:
    """The RWG inductance entries of a triangle mesh (nodes, tris, z) on the device: entry (i, j) =
    mu0 / 4 pi sum over the support triangles of w_i w_j int int (r - v_i).(r' - v_j) / R, every pair
    integrated in canonical order (compressed.py:59-73, compressed_kernels.hpp:9-39). The geometry is
    centred as the fill's; `lo`, `hi` are each edge's support box (compressed.py:54-57,132).
    Also the store of every block built through it (pq_aca_blocks / pq_dense_blocks)."""

    def __init__(self, nodes, tris, z, near=3.0, outer_levels=1, device=0):
        from .backend import Mi355xBackend
        from .geometry import MU0, RWG, FillGeometry
        g = FillGeometry.prepare(nodes, tris, z, outer_levels)
        self.geometry, self.device = g, device
        self._handle = None
        self._backend = Mi355xBackend(g, 1, 16384, device)
        r = RWG(g.nodes, g.tris)
        self.n = int(r.Ne)
        self.support = np.ascontiguousarray(np.stack((r.tp, r.tm), 1), np.int32)
        self.vertices = np.ascontiguousarray(np.stack((r.vp, r.vm), 1), np.float64)
        self.alpha = np.ascontiguousarray(np.stack((1 / (2 * r.area[r.tp]), -1 / (2 * r.area[r.tm])), 1))
        v = g.nodes[g.tris]
        lo, hi = np.c_[v.min(1), g.z], np.c_[v.max(1), g.z]
        self.lo, self.hi = lo[self.support].min(1), hi[self.support].max(1)
        h = ctypes.c_void_p()
        call('pq_rwg_entries_create', self._backend._handle, self.n, ptr(self.support), ptr(self.vertices),
             ptr(self.alpha), MU0 / (4 * np.pi), float(near), int(outer_levels), ctypes.byref(h))
        self._handle = h

    def __call__(self, rows, cols):
        """The (len(rows), len(cols)) entry block (rwg_inductance, compressed_kernels.hpp:127-144)."""
        rows = np.ascontiguousarray(rows, np.int32).ravel()
        cols = np.ascontiguousarray(cols, np.int32).ravel()
        out = np.empty((len(rows), len(cols)))
        call('pq_entries_block', self._handle, ptr(rows), len(rows), ptr(cols), len(cols), ptr(out))
        return out

    def aca(self, rows, cols, tol):
        """(kept ids, ranks) of the checked ACA of the blocks rows[q] x cols[q] (edge index arrays);
        -1 where it did not converge."""
        rp, r, cp, c = _lists(rows, cols)
        ids, ranks = np.empty(len(rows), np.int64), np.empty(len(rows), np.int64)
        call('pq_aca_blocks', self._handle, len(rows), ptr(rp), ptr(r), ptr(cp), ptr(c), float(tol),
             ptr(ids), ptr(ranks), None)
        return ids, ranks

    def dense(self, rows, cols, same):
        """Kept ids of the dense blocks rows[q] x cols[q] (diagonal ones, same[q], symmetrised)."""
        rp, r, cp, c = _lists(rows, cols)
        same = np.ascontiguousarray(same, np.int32)
        ids = np.empty(len(rows), np.int64)
        call('pq_dense_blocks', self._handle, len(rows), ptr(rp), ptr(r), ptr(cp), ptr(c), ptr(same), ptr(ids))
        return ids

    def factors(self, kept, m, n, rank):
        """Host copies of a kept block: (dense, None) or (U, V)."""
        if rank < 0:
            u = np.empty((m, n))
            call('pq_entries_factors', self._handle, int(kept), ptr(u), None)
            return u, None
        u, v = np.empty((m, rank)), np.empty((rank, n))
        call('pq_entries_factors', self._handle, int(kept), ptr(u), ptr(v))
        return u, v

    def stats(self):
        entries, reruns, nbytes = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        ms = (ctypes.c_double * 2)()
        call('pq_entries_stats', self._handle, ctypes.byref(entries), ctypes.byref(reruns), ctypes.byref(nbytes), ms)
        return dict(evaluated_entries=entries.value, aca_reruns=reruns.value, kept_bytes=nbytes.value,
                    aca_ms=ms[0], dense_ms=ms[1])

    def close(self):
        from . import _lib
        if getattr(self, '_handle', None) is not None and _lib._lib is not None:
            _lib._lib.pq_entries_destroy(self._handle)
        self._handle = None
        if getattr(self, '_backend', None) is not None:
            self._backend.close()
            self._backend = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass


class _Flat:
    """The cluster tree as arrays (preorder): boxes, children (-1 at leaves), sizes, index arrays."""

    def __init__(self, root):
        nodes, stack = [], [root]
        while stack:
            node = stack.pop()
            node.k = len(nodes)
            nodes.append(node)
            stack.extend(reversed(node.children))
        self.nodes = nodes
        self.lo = np.array([n.lo for n in nodes])
        self.hi = np.array([n.hi for n in nodes])
        self.kids = np.array([[c.k for c in n.children] if n.children else [-1, -1] for n in nodes], np.int64)
        self.size = np.array([len(n.ids) for n in nodes], np.int64)
        self.ids = [np.ascontiguousarray(n.ids, np.int32) for n in nodes]

    def admissible(self, A, B, eta):
        """_admissible over arrays of node pairs; pairs within 1e-9 of the threshold are decided by
        the scalar rule itself, so the decisions are _admissible's."""
        d = np.maximum(0, np.maximum(self.lo[A] - self.hi[B], self.lo[B] - self.hi[A]))
        gap = np.sqrt((d * d).sum(1))
        da, db = self.hi[A] - self.lo[A], self.hi[B] - self.lo[B]
        bound = eta * np.maximum(np.sqrt((da * da).sum(1)), np.sqrt((db * db).sum(1)))
        adm = (A != B) & (gap > bound)
        for k in np.flatnonzero((A != B) & (np.abs(gap - bound) <= 1e-9 * np.maximum(gap, bound))):
            adm[k] = _admissible(self.nodes[A[k]], self.nodes[B[k]], eta)
        return adm

    def children(self, A, B):
        """_children of every pair, concatenated in pair order."""
        kids, size = self.kids, self.size
        same = A == B
        xa, ya, xb, yb = kids[A, 0], kids[A, 1], kids[B, 0], kids[B, 1]
        split_a = ~same & (xa >= 0) & ((xb < 0) | (size[A] >= size[B]))
        split_b = ~same & ~split_a
        ca, cb = np.full((len(A), 3), -1, np.int64), np.full((len(A), 3), -1, np.int64)
        ca[same], cb[same] = np.c_[xa, xa, ya][same], np.c_[xa, ya, ya][same]
        ca[split_a, :2], cb[split_a, :2] = np.c_[xa, ya][split_a], B[split_a, None]
        ca[split_b, :2], cb[split_b, :2] = A[split_b, None], np.c_[xb, yb][split_b]
        keep = ca >= 0
        return ca[keep], cb[keep]


def build(entries, tol=1e-6, leaf=64, eta=1.0):
    """The compressed operator of `entries` (an RwgEntries) on the device: per level of the cluster
    tree, the admissible pairs by ACA, the failed ones subdivided (or dense at the leaves), then the
    level's dense leaf pairs, diagonal ones symmetrised; blocks in the reference's order
    (hmatrix.py:179-210). The level's classification runs over arrays of node pairs. Returns a
    DeviceHMatrix with `blocks` ((a, b, kept id, rank, same) per block, rank -1 dense) and `info`
    (counts, device milliseconds, host seconds)."""
    if not np.isfinite(tol) or not 0 < tol < 1 or not np.isfinite(eta) or eta <= 0:
        raise ValueError('invalid compression tolerance or admissibility')
    if not isinstance(leaf, (int, np.integer)) or leaf < 1:
        raise ValueError('leaf must be a positive integer')
    t0 = time.perf_counter()
    layout, levels = [], []
    if entries.n:
        T = _Flat(_tree(np.arange(entries.n), entries.lo, entries.hi, leaf))
        A = B = np.zeros(1, np.int64)
    else:
        A = B = np.zeros(0, np.int64)
    while len(A):
        adm = T.admissible(A, B, eta)
        leaves = (T.kids[A, 0] < 0) & (T.kids[B, 0] < 0)
        dense_a, dense_b = A[~adm & leaves], B[~adm & leaves]
        next_a, next_b = T.children(A[~adm & ~leaves], B[~adm & ~leaves])
        level = dict(pairs=len(A), aca_blocks=int(adm.sum()), aca_s=0.0, dense_s=0.0)
        if adm.any():
            aa, ab = A[adm], B[adm]
            level['largest_aca_block'] = int((T.size[aa] * T.size[ab]).max())
            t1 = time.perf_counter()
            kept, ranks = entries.aca([T.ids[k] for k in aa], [T.ids[k] for k in ab], tol)
            level['aca_s'] = time.perf_counter() - t1
            level['aca_converged'] = int((kept >= 0).sum())
            ok = kept >= 0
            layout += [(T.ids[a], T.ids[b], int(k), int(r), False) for a, b, k, r in
                       zip(aa[ok], ab[ok], kept[ok], ranks[ok])]
            failed_leaves = ~ok & (T.kids[aa, 0] < 0) & (T.kids[ab, 0] < 0)
            dense_a, dense_b = np.r_[dense_a, aa[failed_leaves]], np.r_[dense_b, ab[failed_leaves]]
            fa, fb = T.children(aa[~ok & ~failed_leaves], ab[~ok & ~failed_leaves])
            next_a, next_b = np.r_[next_a, fa], np.r_[next_b, fb]
        level['dense_blocks'] = len(dense_a)
        if len(dense_a):
            same = dense_a == dense_b
            t1 = time.perf_counter()
            kept = entries.dense([T.ids[k] for k in dense_a], [T.ids[k] for k in dense_b], same)
            level['dense_s'] = time.perf_counter() - t1
            layout += [(T.ids[a], T.ids[b], int(k), -1, bool(s)) for a, b, k, s in zip(dense_a, dense_b, kept, same)]
        levels.append(level)
        A, B = next_a, next_b
    order = np.array([k for _, _, k, _, _ in layout], np.int64)
    h = ctypes.c_void_p()
    call('pq_entries_hmatrix', entries._handle, len(order), ptr(order), ctypes.byref(h))
    H = DeviceHMatrix._adopt(h, entries.n, entries.device)
    H.blocks = layout
    H.info.update(entries.stats(), tolerance=tol, leaf=leaf, eta=eta, dense_bytes=entries.n ** 2 * 8,
                  build_seconds=time.perf_counter() - t0, levels=levels)
    return H


def rwg_inductance(board, device=0):
    """The dense FP64 RWG inductance of a board (device fill and combine, as the loop solve's) and
    each RWG's support box (its two triangles): (L, lo, hi)."""
    from .backend import Mi355xBackend, combine_device
    from .geometry import MU0, RWG, FillGeometry
    g = FillGeometry.prepare(board.nodes, board.tris, board.z_of, 1)
    backend = Mi355xBackend(g, 1, 16384, device)
    planes, _ = backend.fill_device(3.0, 1, False, 64)
    rwg = RWG(g.nodes, board.tris)
    Ld = combine_device(planes, 64, rwg, g.cent, MU0 / (4 * np.pi), 1.0 / rwg.l, 64, device)
    L = np.ascontiguousarray(Ld.host()[:, :Ld.shape[0]])
    Ld.close()
    planes.close()
    backend.close()
    xy = board.nodes[board.tris]                                                       # (nt, 3, 2)
    pts = np.concatenate([xy, np.broadcast_to(board.z_of[:, None, None], xy.shape[:2] + (1,))], axis=2)
    both = np.concatenate([pts[rwg.tp], pts[rwg.tm]], axis=1)
    return L, both.min(1), both.max(1)


def main(argv=None):
    """Build the compressed RWG inductance of a seeded board and time its device product, at the
    reference's default options (tol 1e-6, leaf 64, eta 1). --builder device: the device
    construction (no dense matrix; --check compares products with the dense FP64 inductance of the
    device fill); --builder svd: the host builder over the dense inductance."""
    from . import board as boards
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument('--board', choices=['c2', 'c5', 'small'], default='small')
    ap.add_argument('--triangles', type=int, default=5800, help='triangles of the small board')
    ap.add_argument('--builder', choices=['device', 'svd'], default='device')
    ap.add_argument('--check', action='store_true')
    ap.add_argument('--columns', type=int, nargs='+', default=[2, 8])
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--leaf', type=int, default=64)
    ap.add_argument('--reps', type=int, default=50)
    args = ap.parse_args(argv)
    if args.board == 'c2':
        b, name = boards.board_c2(0), 'config-2 board (2 layers, 60 x 60 mm)'
    elif args.board == 'c5':
        b, name = boards.board_c5(0), 'config-5 board (4 layers, 120 x 120 mm, 200 vias)'
    else:
        b, name = boards.board_c2(0, size=30e-3, target=args.triangles), 'seeded 2-layer board'
    out = dict(builder=args.builder, tol=args.tol, leaf=args.leaf, products={})
    L = None
    t0 = time.perf_counter()
    if args.builder == 'device':
        E = RwgEntries(b.nodes, b.tris, b.z_of)
        t1 = time.perf_counter()
        H = build(E, tol=args.tol, leaf=args.leaf)
        t2 = time.perf_counter()
        n = E.n
        out.update(provider_s=t1 - t0, build_s=t2 - t1, **{k: v for k, v in H.info.items() if k not in ('tolerance', 'leaf', 'eta')})
        E.close()
        if args.check:
            L, _, _ = rwg_inductance(b)
    else:
        L, lo, hi = rwg_inductance(b)
        t1 = time.perf_counter()
        blocks = compress(lambda a, c: L[np.ix_(a, c)], lo, hi, tol=args.tol, leaf=args.leaf)
        t2 = time.perf_counter()
        H = DeviceHMatrix(blocks, len(L))
        n = len(L)
        out.update(fill_combine_s=t1 - t0, compress_s=t2 - t1, blocks=H.info['blocks'],
                   low_rank_blocks=H.info['low_rank_blocks'], bytes=H.storage_bytes, dense_bytes=L.nbytes)
    out['workload'] = f'RWG inductance, {name}: {b.nt} triangles, {n} RWG'
    rng = np.random.default_rng(0)
    for k in args.columns:
        x = rng.standard_normal((n, k))
        y = H @ x
        ms = H.time_product(k, args.reps)
        rec = dict(ms=ms, factor_GBs=H.storage_bytes * ((k + 7) // 8) / ms / 1e6)
        if L is not None:
            ref = L @ x
            rec['rel_err_vs_dense'] = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        out['products'][k] = rec
    print(json.dumps(out), flush=True)
    H.close()


if __name__ == '__main__':
    main()
