"""Oracle for the compressed (H-matrix) operator: its product and its construction (test
infrastructure only; see oracle/__init__.py).

Restates, in numpy:
  matmat       HMatrix._matmat, the per-block Python product (src/pcbqs/hmatrix.py:223-233)
  pack         HMatrix._pack's slot numbering and row order (src/pcbqs/hmatrix.py:235-270)
  two_phase    hmatrix_product's deterministic two-phase sum (src/pcbqs/core/hmatrix_kernels.hpp:20-101)
  dense        the operator the blocks stand for, as a dense array
  coverage     how often each entry is stood for by a block or a block's transpose
  canonical_moments  canonical_pair (src/pcbqs/core/compressed_kernels.hpp:9-21, compressed.py:36-51)
  rwg_block    rwg_entry / rwg_block, the RWG inductance entries of an index block
               (compressed_kernels.hpp:23-39,73-108; compressed.py:79-92), with rounded products
  splitmix64, draw, aca   the compiled ACA with its SplitMix64 check-row draws
               (src/pcbqs/core/hmatrix_build.hpp:143-270), in the same arithmetic order
  build        HMatrix's level-by-level native construction (src/pcbqs/hmatrix.py:88-95,179-210):
               cluster tree, admissibility, ACA for admissible blocks, dense (diagonal: symmetrised)
               leaves, failed ACA blocks subdivided or made dense

Pinned by tests/test_hmatrix_oracle.py: matmat equals dense(blocks) @ x (a different computation)
and the matrix the blocks were built from (the reference's own known-answer case: 512 points,
1/sqrt(r^2 + 0.005^2), leaf 16, tol 1e-7, tests/test_compressed.py:10-19); two_phase equals
matmat; a product that drops the transposed contributions fails these checks.
tests/test_aca_oracle.py: rwg_block equals the reference core's rwg_inductance bit for bit and the
dense RWG inductance of the oracle's fill and combine to 1e-12; aca equals the reference core's
aca_blocks bit for bit (factors and failures) where oracle/_ref is built; aca recovers a random
rank-5 matrix at rank 5 and gives up on a full-rank one; the built operator's product is within
the tolerance of the dense matrix and of the reference core's own native build (50 tol, the
reference's bound in tests/test_compressed.py:300-319), and its partition covers each entry once.
"""
import numpy as np

from . import fill as F

MU0 = 4e-7 * np.pi
TINY = np.finfo(float).tiny


def matmat(blocks, n, x):
    """The reference's Python product (hmatrix.py:226-233)."""
    x = np.asarray(x)
    dt = np.result_type(*[u.dtype for _, _, u, _, _ in blocks], x.dtype) if blocks else x.dtype
    out = np.zeros((n,) + x.shape[1:], dt)
    for a, b, u, v, same in blocks:
        out[a] += u @ x[b] if v is None else u @ (v @ x[b])
        if not same:
            out[b] += u.T @ x[a] if v is None else v.T @ (u.T @ x[a])
    return out


def pack(blocks, n):
    """Slot numbering (forward rows, then transposed rows, block by block) and each row's slots in
    slot order (hmatrix.py:235-270)."""
    targets, starts, so = [], [], 0
    for a, b, u, v, same in blocks:
        fwd = so
        so += len(a)
        tr = -1 if same else so
        if not same:
            so += len(b)
        starts.append((fwd, tr))
        targets.append(np.asarray(a, np.int64))
        if not same:
            targets.append(np.asarray(b, np.int64))
    t = np.concatenate(targets) if targets else np.zeros(0, np.int64)
    slots = np.argsort(t, kind='stable')
    row_ptr = np.r_[0, np.cumsum(np.bincount(t, minlength=n))].astype(np.int64)
    return starts, slots, row_ptr, so


def two_phase(blocks, n, x):
    """hmatrix_product: every block's contribution rows into its slots, then each row's slots
    summed in slot order (hmatrix_kernels.hpp:20-101)."""
    x = np.asarray(x)
    x2 = x.reshape(n, -1)
    dt = np.result_type(*[u.dtype for _, _, u, _, _ in blocks], x.dtype) if blocks else x.dtype
    starts, slots, row_ptr, so = pack(blocks, n)
    s = np.zeros((so, x2.shape[1]), dt)
    for (a, b, u, v, same), (fwd, tr) in zip(blocks, starts):
        if v is None:
            s[fwd:fwd + len(a)] = u @ x2[b]
            if not same:
                s[tr:tr + len(b)] = u.T @ x2[a]
        else:
            s[fwd:fwd + len(a)] = u @ (v @ x2[b])
            if not same:
                s[tr:tr + len(b)] = v.T @ (u.T @ x2[a])
    y = np.zeros((n, x2.shape[1]), dt)
    for row in range(n):
        for p in range(row_ptr[row], row_ptr[row + 1]):
            y[row] += s[slots[p]]
    return y.reshape((n,) + x.shape[1:])


def dense(blocks, n):
    dt = np.result_type(*[u.dtype for _, _, u, _, _ in blocks]) if blocks else float
    A = np.zeros((n, n), dt)
    for a, b, u, v, same in blocks:
        blk = u if v is None else u @ v
        A[np.ix_(a, b)] += blk
        if not same:
            A[np.ix_(b, a)] += blk.T
    return A


def coverage(blocks, n):
    """How many times each entry (i, j) is stood for by some block or block transpose."""
    C = np.zeros((n, n), np.int64)
    for a, b, u, v, same in blocks:
        C[np.ix_(a, b)] += 1
        if not same:
            C[np.ix_(b, a)] += 1
    return C


def reference_kernel(n=512, seed=7, complex_=False):
    """The points and entry matrix of the reference's H-matrix tests (tests/test_compressed.py:11-12)."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3))
    pts[:, 2] *= .01
    A = (1 - .02j if complex_ else 1) / np.sqrt(((pts[:, None] - pts[None, :]) ** 2).sum(-1) + .005 ** 2)
    return rng, pts, A


# ---- entries: canonical pairs and the RWG inductance block ------------------------------------

def canonical_moments(g, p, q, near=3.0, outer=1, scalar_only=False):
    """Moments of the ordered pairs (p, q): each integrated as (min, max), the first moments
    swapped back for p > q and averaged for p == q (compressed_kernels.hpp:9-21)."""
    p, q = np.asarray(p, np.int64), np.asarray(q, np.int64)
    a, b = np.minimum(p, q), np.maximum(p, q)
    d = g.z[a] - g.z[b]
    m = F.cr_pairs(g, a, b, d, F.kinds(g, a, b, d, near, outer), scalar_only)
    if not scalar_only:
        flip = p > q
        m[flip, 1:5] = m[flip][:, [3, 4, 1, 2]]
        same = p == q
        x = (m[same, 1:3] + m[same, 3:5]) / 2
        m[same, 1:3] = x
        m[same, 3:5] = x
    return m


def rwg_arrays(g):
    """(support (ne, 2), vertices (ne, 2, 2), alpha (ne, 2)) of the interior edges of the centred
    geometry g (compressed.py:61-64): plus and minus triangles, their free vertices, +-1/(2 area)."""
    from .mqs import RWG
    r = RWG(g.nodes, g.tris)
    support = np.stack((r.tp, r.tm), 1).astype(np.int64)
    vertices = np.stack((r.vp, r.vm), 1)
    alpha = np.stack((1 / (2 * r.area[r.tp]), -1 / (2 * r.area[r.tm])), 1)
    return support, vertices, alpha


def edge_boxes(g, support):
    """Each edge's support box: the union of its triangles' (x, y, z) boxes (compressed.py:54-57,132)."""
    v = g.nodes[g.tris]
    lo, hi = np.c_[v.min(1), g.z], np.c_[v.max(1), g.z]
    return lo[support].min(1), hi[support].max(1)


def rwg_block(g, support, vertices, alpha, rows, cols, near=3.0, outer=1, scale=MU0 / (4 * np.pi)):
    """entry[i, j] = scale sum_{sa, sb} (w_a w_b) (m5 - v_q.m_src - v_p.m_obs + (v_p.v_q) m0), every
    product rounded, the four support terms summed in (sa, sb) order (compressed_kernels.hpp:23-39)."""
    rows, cols = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    la, lb = len(rows), len(cols)
    p = np.broadcast_to(support[rows][:, None, :, None], (la, lb, 2, 2)).ravel()
    q = np.broadcast_to(support[cols][None, :, None, :], (la, lb, 2, 2)).ravel()
    m = canonical_moments(g, p, q, near, outer).reshape(la, lb, 2, 2, 6)
    vp = vertices[rows][:, None, :, None, :]
    vq = vertices[cols][None, :, None, :, :]
    wa, wb = alpha[rows][:, None, :, None], alpha[cols][None, :, None, :]
    val = (m[..., 5] - (vq[..., 0] * m[..., 1] + vq[..., 1] * m[..., 2])
           - (vp[..., 0] * m[..., 3] + vp[..., 1] * m[..., 4])
           + (vp[..., 0] * vq[..., 0] + vp[..., 1] * vq[..., 1]) * m[..., 0])
    t = (wa * wb) * val
    total = ((t[..., 0, 0] + t[..., 0, 1]) + t[..., 1, 0]) + t[..., 1, 1]
    return scale * total


# ---- ACA -----------------------------------------------------------------------------------

MASK = (1 << 64) - 1


def splitmix64(state):
    """(next state, output): SplitMix64 (Steele, Lea & Flood, OOPSLA 2014)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def draw(state, m, k):
    """(state, k distinct indices of [0, m)): a partial Fisher-Yates shuffle of 0..m-1 whose swaps
    are kept in a map (hmatrix_build.hpp:151-160)."""
    moved, out = {}, []
    for i in range(k):
        state, r = splitmix64(state)
        j = i + r % (m - i)
        vi, vj = moved.get(i, i), moved.get(j, j)
        moved[i], moved[j] = vj, vi
        out.append(vj)
    return state, out


def block_seed(a0, b0, m, n):
    """The per-block stream's seed from the block's identity (hmatrix_build.hpp:172-173)."""
    return (0x5DEECE66D ^ ((int(a0) * 0x100000001B3) & MASK) ^ ((int(b0) << 21) & MASK)
            ^ ((int(m) << 42) & MASK) ^ int(n)) & MASK


def _seqsum(x):
    """A left-to-right sum (the compiled loops' order)."""
    x = np.ravel(x)
    return float(np.cumsum(x)[-1]) if len(x) else 0.0


def aca(entry, a, b, tol):
    """(U (m, r), V (r, n), entries evaluated) of the checked partial-pivot ACA of block (a, b), or
    (None, entries): hmatrix_build.hpp:167-270, the same pivots, draws and roundings. entry(rows,
    cols) gives the entries of global index arrays."""
    a, b = np.asarray(a, np.int64), np.asarray(b, np.int64)
    m, n = len(a), len(b)
    cap = min(128, int((0.7 * m) * n / (m + n)))
    if cap < 1:
        return None, 0
    state = block_seed(a[0], b[0], m, n)
    cache, count = {}, [0]

    def rows(idx):
        missing = [k for k in dict.fromkeys(int(k) for k in idx) if k not in cache]
        if missing:
            vals = np.asarray(entry(a[missing], b), float)
            count[0] += vals.size
            for k, r in zip(missing, vals):
                cache[k] = r

    U, V = np.zeros((cap, m)), np.zeros((cap, n))
    rank = 0

    def approx_row(i):
        s = np.zeros(n)
        for l in range(rank):
            s = s + U[l, i] * V[l]
        return s

    state, check = draw(state, m, min(m, 16))
    rows(check)
    residual = np.array([cache[k] for k in check])
    ref2 = _seqsum(residual * residual)
    used = np.zeros(m, bool)
    i = 0
    for _ in range(cap):
        rows([i])
        row = cache[i] - approx_row(i)
        j = int(np.argmax(np.abs(row)))
        pivot = row[j]
        if abs(pivot) > TINY:
            col = np.empty(m)
            unknown = [k for k in range(m) if k not in cache]
            for k in range(m):
                if k in cache:
                    col[k] = cache[k][j]
            if unknown:
                col[unknown] = np.asarray(entry(a[unknown], b[j:j + 1]), float)[:, 0]
                count[0] += len(unknown)
            s = np.zeros(m)
            for l in range(rank):
                s = s + U[l] * V[l, j]
            U[rank] = col - s
            V[rank] = row / pivot
            rank += 1
            residual = residual - U[rank - 1][check][:, None] * V[rank - 1][None, :]
        used[i] = True
        if np.sqrt(_seqsum(residual * residual)) <= tol * max(np.sqrt(ref2), TINY):
            state, extra = draw(state, m, min(m, 16))
            rows(extra)
            X = np.array([cache[k] for k in extra])
            E = X - np.array([approx_row(k) for k in extra])
            if np.sqrt(_seqsum(E * E)) <= tol * max(np.sqrt(_seqsum(X * X)), TINY):
                return (np.ascontiguousarray(U[:rank].T), np.ascontiguousarray(V[:rank])), count[0]
            check = check + extra
            residual = np.vstack((residual, E))
            ref2 += _seqsum(X * X)
        free = [c for c in range(len(check)) if not used[check[c]]]
        if free:
            s = np.cumsum(residual[free] * residual[free], axis=1)[:, -1]
            i = check[free[int(np.argmax(s))]]
        elif not used.all():
            idle = np.flatnonzero(~used)
            i = int(idle[np.argmax(np.abs(U[rank - 1, idle]))]) if rank else int(idle[0])
        else:
            break
    return None, count[0]


# ---- the level-by-level construction ---------------------------------------------------------

class Cluster:
    def __init__(self, ids, lo, hi):
        self.ids, self.lo, self.hi, self.children = ids, lo, hi, ()


def cluster_tree(lo, hi, leaf):
    """Bisection at the median along the widest spread of the box centres (hmatrix.py:88-93)."""
    def tree(ids):
        node = Cluster(ids, lo[ids].min(0), hi[ids].max(0))
        if len(ids) > leaf:
            centres = (lo[ids] + hi[ids]) / 2
            axis = np.argmax(np.ptp(centres, axis=0))
            order = ids[np.argsort(centres[:, axis], kind='stable')]
            mid = len(ids) // 2
            node.children = (tree(order[:mid]), tree(order[mid:]))
        return node
    return tree(np.arange(len(lo)))


def admissible(a, b, eta):
    gap = np.linalg.norm(np.maximum(0, np.maximum(a.lo - b.hi, b.lo - a.hi)))
    diameter = max(np.linalg.norm(a.hi - a.lo), np.linalg.norm(b.hi - b.lo))
    return a is not b and gap > eta * diameter


def children(a, b):
    if a is b:
        x, y = a.children
        return [(x, x), (x, y), (y, y)]
    if a.children and (not b.children or len(a.ids) >= len(b.ids)):
        return [(c, b) for c in a.children]
    return [(a, c) for c in b.children]


def build(entry, lo, hi, tol=1e-6, leaf=64, eta=1.0):
    """(blocks [(a, b, U, V or None, same)] in the construction's order, diagonal, entries):
    per tree level, the admissible pairs by aca (in order), then the dense leaf pairs (the level's
    leaves, then the leaf pairs whose ACA failed), diagonal blocks symmetrised (hmatrix.py:179-210)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    blocks, entries = [], 0
    diagonal = np.zeros(len(lo))
    pending = [(cluster_tree(lo, hi, leaf),) * 2] if len(lo) else []
    while pending:
        adm, dense_, following = [], [], []
        for a, b in pending:
            if admissible(a, b, eta):
                adm.append((a, b))
            elif not a.children and not b.children:
                dense_.append((a, b))
            else:
                following.extend(children(a, b))
        for a, b in adm:
            f, used = aca(entry, a.ids, b.ids, tol)
            entries += used
            if f is not None:
                blocks.append((a.ids, b.ids, f[0], f[1], False))
            elif not a.children and not b.children:
                dense_.append((a, b))
            else:
                following.extend(children(a, b))
        for a, b in dense_:
            block = np.asarray(entry(a.ids, b.ids), float)
            entries += block.size
            if a is b:
                block = (block + block.T) / 2
                diagonal[a.ids] = np.diag(block)
            blocks.append((a.ids, b.ids, block, None, a is b))
        pending = following
    return blocks, diagonal, entries
