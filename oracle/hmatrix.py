"""Oracle for the compressed (H-matrix) operator product (test infrastructure only; see
oracle/__init__.py).

Restates, in numpy:
  matmat       HMatrix._matmat, the per-block Python product (src/pcbqs/hmatrix.py:223-233)
  pack         HMatrix._pack's slot numbering and row order (src/pcbqs/hmatrix.py:235-270)
  two_phase    hmatrix_product's deterministic two-phase sum (src/pcbqs/core/hmatrix_kernels.hpp:20-101)
  dense        the operator the blocks stand for, as a dense array
  coverage     how often each entry is stood for by a block or a block's transpose

Pinned by tests/test_hmatrix_oracle.py: matmat equals dense(blocks) @ x (a different computation)
and the matrix the blocks were built from (the reference's own known-answer case: 512 points,
1/sqrt(r^2 + 0.005^2), leaf 16, tol 1e-7, tests/test_compressed.py:10-19); two_phase equals
matmat; a product that drops the transposed contributions fails these checks.
"""
import numpy as np


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
