"""The H-matrix product oracle pinned without the device: the per-block product equals the dense
operator the blocks stand for, the two-phase slot sum equals the per-block product, the partition
covers each entry exactly once, and the builder reproduces the reference's own known-answer test
(tests/test_compressed.py:10-19); the blocks come from the host builder pcbqs_amd.hmatrix.compress,
whose partition covers every entry exactly once."""
import numpy as np
import pytest

from oracle import hmatrix as H
from pcbqs_amd.hmatrix import compress


@pytest.mark.parametrize('complex_', [False, True])
def test_builder_meets_the_reference_kernel_test(complex_):
    rng, pts, A = H.reference_kernel(complex_=complex_)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-7, leaf=16,
                     dtype=complex if complex_ else float)
    x = rng.normal(size=(512, 3)) + 1j * rng.normal(size=(512, 3))
    np.testing.assert_allclose(H.matmat(blocks, 512, x), A @ x, rtol=1e-6, atol=1e-5)
    assert any(v is not None for *_, v, _ in blocks)
    storage = sum(u.nbytes + (v.nbytes if v is not None else 0) for *_, u, v, _ in blocks)
    assert storage < A.nbytes
    assert (H.coverage(blocks, 512) == 1).all()


@pytest.mark.parametrize('complex_', [False, True])
def test_products_agree(complex_):
    rng, pts, A = H.reference_kernel(n=200, seed=3, complex_=complex_)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-7, leaf=8,
                     dtype=complex if complex_ else float)
    D = H.dense(blocks, 200)
    for x in (rng.normal(size=200), rng.normal(size=(200, 3)), rng.normal(size=(200, 2)) + 1j * rng.normal(size=(200, 2))):
        y = H.matmat(blocks, 200, x)
        np.testing.assert_allclose(y, D @ x, rtol=1e-12, atol=1e-12 * np.abs(y).max())
        np.testing.assert_allclose(H.two_phase(blocks, 200, x), y, rtol=1e-12, atol=1e-12 * np.abs(y).max())


def test_a_dropped_transpose_is_caught():
    """The checks above fail for a product that forgets the off-diagonal blocks' transposes."""
    rng, pts, A = H.reference_kernel(n=100, seed=5)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-7, leaf=8)
    as_same = [(a, b, u, v, True) for a, b, u, v, _ in blocks]
    x = rng.normal(size=(100, 2))
    assert not np.allclose(H.matmat(as_same, 100, x), A @ x, rtol=1e-6, atol=1e-5)
    assert not (H.coverage(as_same, 100) == 1).all()


def test_randomised_factors_of_large_far_blocks_meet_the_tolerance():
    """Far blocks above 256 rows and columns take the randomised range path of the host builder:
    the factors meet the tolerance, and a block of too high a rank is refused."""
    from pcbqs_amd.hmatrix import _low_rank
    rng = np.random.default_rng(11)
    p, q = rng.random((400, 2)), rng.random((600, 2)) + [3.0, 0.0]
    B = 1 / np.sqrt(((p[:, None] - q[None, :]) ** 2).sum(-1) + .005 ** 2)
    u, v = _low_rank(B, 1e-6, rng)
    assert u.shape[1] < 40
    assert np.linalg.norm(u @ v - B) <= 1e-6 * np.linalg.norm(B)
    assert _low_rank(rng.standard_normal((400, 600)), 1e-6, rng) is None


def test_the_c_abi_rejects_inconsistent_layouts():
    """Layout errors are reported before any device work (tests/test_compressed.py:199-222)."""
    import ctypes
    from pcbqs_amd._lib import call, ptr
    meta = np.array([[0, 2, 2, 2, -1, 0]], np.int64)
    idx = np.array([0, 1, 2, 3], np.int32)
    f = np.ones(4)
    h = ctypes.c_void_p()
    for bad_meta, bad_idx, n_f, msg in [(np.array([[0, 2, 2, 3, -1, 0]], np.int64), idx, 4, 'index array'),
                                        (meta, np.array([0, 1, 2, 9], np.int32), 4, 'out of range'),
                                        (meta, idx, 3, 'factor array'),
                                        (np.array([[0, 2, 2, 2, 300, 0]], np.int64), idx, 4, 'rank')]:
        with pytest.raises(ValueError, match=msg):
            call('pq_hmatrix_create', 0, 4, 1, ptr(bad_meta), ptr(bad_idx), 4, ptr(f), n_f, 0, ctypes.byref(h))
