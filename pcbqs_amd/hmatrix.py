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


def _lists(pairs):
    """(offsets, concatenated int32 ids) of the row clusters and of the column clusters."""
    rows = [a.ids for a, _ in pairs]
    cols = [b.ids for _, b in pairs]
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

    def aca(self, pairs, tol):
        """(kept ids, ranks) of the checked ACA of cluster pairs; -1 where it did not converge."""
        rp, rows, cp, cols = _lists(pairs)
        ids, ranks = np.empty(len(pairs), np.int64), np.empty(len(pairs), np.int64)
        call('pq_aca_blocks', self._handle, len(pairs), ptr(rp), ptr(rows), ptr(cp), ptr(cols), float(tol),
             ptr(ids), ptr(ranks), None)
        return ids, ranks

    def dense(self, pairs):
        """Kept ids of dense blocks (diagonal ones symmetrised)."""
        rp, rows, cp, cols = _lists(pairs)
        same = np.array([a is b for a, b in pairs], np.int32)
        ids = np.empty(len(pairs), np.int64)
        call('pq_dense_blocks', self._handle, len(pairs), ptr(rp), ptr(rows), ptr(cp), ptr(cols), ptr(same),
             ptr(ids))
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


def build(entries, tol=1e-6, leaf=64, eta=1.0):
    """The compressed operator of `entries` (an RwgEntries) on the device: per level of the cluster
    tree, the admissible pairs by ACA, the failed ones subdivided (or dense at the leaves), then the
    level's dense leaf pairs, diagonal ones symmetrised; blocks in the reference's order
    (hmatrix.py:179-210). Returns a DeviceHMatrix with `blocks` ((a, b, kept id, rank, same) per
    block, rank -1 dense) and `info` (counts, device milliseconds, host seconds)."""
    if not np.isfinite(tol) or not 0 < tol < 1 or not np.isfinite(eta) or eta <= 0:
        raise ValueError('invalid compression tolerance or admissibility')
    if not isinstance(leaf, (int, np.integer)) or leaf < 1:
        raise ValueError('leaf must be a positive integer')
    t0 = time.perf_counter()
    layout = []
    pending = [(_tree(np.arange(entries.n), entries.lo, entries.hi, leaf),) * 2] if entries.n else []
    while pending:
        adm, dense, following = [], [], []
        for a, b in pending:
            if _admissible(a, b, eta):
                adm.append((a, b))
            elif not a.children and not b.children:
                dense.append((a, b))
            else:
                following.extend(_children(a, b))
        if adm:
            ids, ranks = entries.aca(adm, tol)
            for (a, b), k, r in zip(adm, ids, ranks):
                if k >= 0:
                    layout.append((a.ids, b.ids, int(k), int(r), False))
                elif not a.children and not b.children:
                    dense.append((a, b))
                else:
                    following.extend(_children(a, b))
        if dense:
            for (a, b), k in zip(dense, entries.dense(dense)):
                layout.append((a.ids, b.ids, int(k), -1, a is b))
        pending = following
    order = np.array([k for _, _, k, _, _ in layout], np.int64)
    h = ctypes.c_void_p()
    call('pq_entries_hmatrix', entries._handle, len(order), ptr(order), ctypes.byref(h))
    H = DeviceHMatrix._adopt(h, entries.n, entries.device)
    H.blocks = layout
    H.info.update(entries.stats(), tolerance=tol, leaf=leaf, eta=eta, dense_bytes=entries.n ** 2 * 8,
                  build_seconds=time.perf_counter() - t0)
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
    """Time the device product on the RWG inductance of a seeded two-layer board compressed at the
    reference's default options (tol 1e-6, leaf 64, eta 1)."""
    from . import board as boards
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument('--triangles', type=int, default=5800)
    ap.add_argument('--columns', type=int, nargs='+', default=[2, 8])
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--leaf', type=int, default=64)
    ap.add_argument('--reps', type=int, default=50)
    args = ap.parse_args(argv)
    b = boards.board_c2(0, size=30e-3, target=args.triangles)
    t0 = time.perf_counter()
    L, lo, hi = rwg_inductance(b)
    t1 = time.perf_counter()
    blocks = compress(lambda a, c: L[np.ix_(a, c)], lo, hi, tol=args.tol, leaf=args.leaf)
    t2 = time.perf_counter()
    H = DeviceHMatrix(blocks, len(L))
    rng = np.random.default_rng(0)
    out = dict(workload=f'RWG inductance, seeded 2-layer board: {b.nt} triangles, {len(L)} RWG',
               tol=args.tol, leaf=args.leaf, blocks=H.info['blocks'], low_rank_blocks=H.info['low_rank_blocks'],
               factor_bytes=H.storage_bytes, dense_bytes=L.nbytes, fill_combine_s=t1 - t0, compress_s=t2 - t1,
               products={})
    for k in args.columns:
        x = rng.standard_normal((len(L), k))
        y = H @ x
        ref = L @ x
        ms = H.time_product(k, args.reps)
        # algorithmic bytes: every factor once, X gathered per block row/column, the slots written and read
        slot_bytes = 2 * H.info['slots'] * 8 * min(k, 8) * ((k + 7) // 8)
        out['products'][k] = dict(ms=ms, rel_err_vs_dense=float(np.linalg.norm(y - ref) / np.linalg.norm(ref)),
                                  factor_GBs=H.storage_bytes * ((k + 7) // 8) / ms / 1e6,
                                  slot_bytes=slot_bytes)
    print(json.dumps(out), flush=True)
    H.close()


if __name__ == '__main__':
    main()
