"""GPU parity of the compressed-operator product (pq_hmatrix_apply) against the oracle's restatement
of the reference's product: the reference's own case (512 points, leaf 16, tol 1e-7; real and
complex operators; a real vector, real and complex blocks of columns: tests/test_compressed.py:
170-196) at rtol 1e-13, bitwise repeatable; column groups across the 8-column boundary; large
blocks (more than 256 rows, rank up to 128); the adjoint; degenerate layouts."""
import numpy as np
import pytest

from oracle import hmatrix as H
from pcbqs_amd.hmatrix import DeviceHMatrix, compress

pytestmark = pytest.mark.gpu


def close(got, want):
    assert got.dtype == want.dtype and got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize('complex_', [False, True])
def test_reference_case_matches_the_oracle_and_repeats_bitwise(complex_):
    rng, pts, A = H.reference_kernel(complex_=complex_)
    dt = complex if complex_ else float
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-7, leaf=16, dtype=dt)
    D = DeviceHMatrix(blocks, 512, dt)
    print(f'hmatrix reference case complex={complex_}: {D.info}', flush=True)
    assert D.info['low_rank_blocks'] > 0
    inputs = [rng.normal(size=512), rng.normal(size=(512, 3)), rng.normal(size=(512, 2)) + 1j * rng.normal(size=(512, 2))]
    for x in inputs:
        want = H.matmat(blocks, 512, x)
        got = D @ x
        close(got, want)
        assert np.array_equal(D @ x, got)
        np.testing.assert_allclose(got, A @ x, rtol=1e-6, atol=1e-5)
    D.close()


@pytest.mark.parametrize('k', [1, 7, 8, 9, 17])
def test_column_groups(k):
    rng, pts, A = H.reference_kernel(n=300, seed=4)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-8, leaf=12)
    D = DeviceHMatrix(blocks, 300)
    x = rng.normal(size=(300, k))
    close(D @ x, H.two_phase(blocks, 300, x))
    xc = rng.normal(size=(300, k)) + 1j * rng.normal(size=(300, k))
    close(D @ xc, H.matmat(blocks, 300, xc))
    D.close()


@pytest.mark.parametrize('k', [1, 4, 5])
def test_complex_operator_column_groups(k):
    rng, pts, A = H.reference_kernel(n=300, seed=6, complex_=True)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-8, leaf=12, dtype=complex)
    D = DeviceHMatrix(blocks, 300, complex)
    x = rng.normal(size=(300, k)) + 1j * rng.normal(size=(300, k))
    close(D @ x, H.matmat(blocks, 300, x))
    D.close()


def test_large_blocks_and_the_adjoint():
    """Blocks of more than 256 rows (several output chunks per block) and ranks up to the cap."""
    rng = np.random.default_rng(2)
    n = 2500
    pts = np.c_[rng.random((n, 2)), np.zeros(n)]
    A = (1 - .05j) / np.sqrt(((pts[:, None] - pts[None, :]) ** 2).sum(-1) + .002 ** 2)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-9, leaf=300, dtype=complex)
    sizes = [len(a) for a, *_ in blocks]
    ranks = [u.shape[1] for _, _, u, v, _ in blocks if v is not None]
    print(f'large blocks: {len(blocks)} blocks, largest {max(sizes)} rows, ranks {min(ranks)}..{max(ranks)}',
          flush=True)
    assert max(sizes) > 256
    D = DeviceHMatrix(blocks, n, complex)
    x = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    close(D @ x, H.matmat(blocks, n, x))
    Dd = H.dense(blocks, n)
    np.testing.assert_allclose(D.H @ x, Dd.conj().T @ x, rtol=1e-12, atol=1e-12 * np.abs(Dd).max() * np.abs(x).max())
    D.close()


def test_degenerate_layouts():
    """No blocks; rows covered by no block; a rank-0 block; a `same` low-rank block; empty input."""
    rng = np.random.default_rng(9)
    D = DeviceHMatrix([], 10)
    assert np.array_equal(D @ np.ones((10, 2)), np.zeros((10, 2)))
    D.close()
    a, b, c = np.array([0, 2, 4]), np.array([5, 7]), np.array([1, 3])
    blocks = [(a, b, rng.normal(size=(3, 2)), None, False),
              (a, c, np.zeros((3, 0)), np.zeros((0, 2)), False),
              (c, c, rng.normal(size=(2, 2)), rng.normal(size=(2, 2)), True),
              (b, a, rng.normal(size=(2, 1)), rng.normal(size=(1, 3)), False)]
    D = DeviceHMatrix(blocks, 9)
    x = rng.normal(size=(9, 3))
    close(D @ x, H.matmat(blocks, 9, x))
    assert (D @ x)[[6, 8]].tolist() == [[0.0] * 3] * 2
    assert (D @ np.zeros((9, 0))).shape == (9, 0)
    D.close()


def test_from_a_reference_style_operator():
    """DeviceHMatrix.from_hmatrix takes any object with the reference HMatrix's `blocks`, `shape`
    and `dtype` (src/pcbqs/hmatrix.py:59-83), and its products keep the operator's result types."""
    from types import SimpleNamespace
    rng, pts, A = H.reference_kernel(n=200, seed=8, complex_=True)
    blocks = compress(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-8, leaf=16, dtype=complex)
    ref = SimpleNamespace(blocks=blocks, shape=(200, 200), dtype=np.dtype(complex))
    D = DeviceHMatrix.from_hmatrix(ref)
    x = rng.normal(size=200)
    y = D @ x
    assert y.dtype == np.complex128 and y.shape == (200,)
    close(y, H.matmat(blocks, 200, x.astype(complex)))
    np.testing.assert_allclose(D.matvec(x), y, rtol=1e-15, atol=0)
    D.close()
