"""The oracle's construction of the compressed RWG inductance pinned without the device:
- its entry blocks equal the reference core's rwg_inductance bit for bit (near 3 and infinity;
  tests/test_compressed.py:228-263 holds the reference's own entries to the same bar) and the dense
  RWG inductance of the oracle's fill and combine (a different computation) to 1e-12;
- its ACA equals the reference core's aca_blocks bit for bit, factors and failures alike, over
  every admissible block of a two-layer board;
- ACA recovers a random rank-5 matrix at rank 5, gives up on a full-rank one, and its draws are
  distinct indices;
- the level-by-level build covers every entry once and its product is within the tolerance of the
  dense matrix (the reference's kernel case, tests/test_compressed.py:10-19)."""
import numpy as np
import pytest

from oracle import fill as F
from oracle import hmatrix as H
from oracle import mqs as M
from tests.meshes import two_layer

CORE = F.ref_core()
needs_ref = pytest.mark.skipif(CORE is None, reason='reference core (oracle/_ref) not built')


@pytest.fixture(scope='module')
def board():
    g = F.prepare(*two_layer(20, 12, seed=3), 1)
    return g, *H.rwg_arrays(g)


def _common(g, near):
    return (F.ref_geometry_dict(g), g.z, g.cent, g.radius, float(near), 1)


@needs_ref
@pytest.mark.parametrize('near', [3.0, np.inf])
def test_entry_blocks_are_the_reference_core_bitwise(board, near):
    g, sup, vx, al = board
    rng = np.random.default_rng(5)
    rows = rng.choice(len(sup), 40, replace=False)
    cols = np.r_[rows[:5], rng.choice(len(sup), 60, replace=False)]
    ref = CORE.rwg_inductance(*_common(g, near), sup, np.ascontiguousarray(vx), np.ascontiguousarray(al),
                              rows.astype(np.int64), cols.astype(np.int64), H.MU0 / (4 * np.pi), 2)
    assert np.array_equal(H.rwg_block(g, sup, vx, al, rows, cols, near=near), ref)


def test_entry_blocks_match_the_dense_combine(board):
    g, sup, vx, al = board
    mats, _, _ = F.cr_fill(g, 3.0, 1, False)
    M0, Mrx, Mry, Mrr = mats
    r = M.RWG(g.nodes, g.tris)
    L = M.combine(r, Mrr, Mrx, Mry, M0) * (H.MU0 / (4 * np.pi)) / np.outer(r.l, r.l)
    rng = np.random.default_rng(2)
    rows, cols = rng.choice(len(sup), 50, replace=False), rng.choice(len(sup), 70, replace=False)
    got = H.rwg_block(g, sup, vx, al, rows, cols)
    np.testing.assert_allclose(got, L[np.ix_(rows, cols)], rtol=0, atol=1e-12 * np.abs(L).max())
    # the diagonal and a reciprocal pair
    d = H.rwg_block(g, sup, vx, al, rows[:8], rows[:8])
    np.testing.assert_allclose(np.diag(d), np.diag(L)[rows[:8]], rtol=1e-12)
    np.testing.assert_allclose(d, d.T, rtol=1e-13, atol=1e-15 * np.abs(d).max())


def _admissible_blocks(g, sup, leaf):
    lo, hi = H.edge_boxes(g, sup)
    pending, out = [(H.cluster_tree(lo, hi, leaf),) * 2], []
    while pending:
        nxt = []
        for a, b in pending:
            if H.admissible(a, b, 1.0):
                out.append((a, b))
            elif a.children or b.children:
                nxt.extend(H.children(a, b))
        pending = nxt
    return out


@needs_ref
def test_aca_is_the_reference_core_bitwise():
    g = F.prepare(*two_layer(32, 20, seed=3), 1)
    sup, vx, al = H.rwg_arrays(g)
    blocks = _admissible_blocks(g, sup, 32)
    provider = CORE.RwgEntries(*_common(g, 3.0), sup, np.ascontiguousarray(vx), np.ascontiguousarray(al),
                               H.MU0 / (4 * np.pi))
    factors, _, _ = CORE.aca_blocks(provider, [a.ids.astype(np.int64) for a, _ in blocks],
                                    [b.ids.astype(np.int64) for _, b in blocks], 1e-6, 4)
    entry = lambda r, c: H.rwg_block(g, sup, vx, al, r, c)
    converged = 0
    for (a, b), ref in zip(blocks, factors):
        ours, _ = H.aca(entry, a.ids, b.ids, 1e-6)
        assert (ours is None) == (ref is None)
        if ref is not None:
            converged += 1
            assert np.array_equal(ours[0], ref[0]) and np.array_equal(ours[1], ref[1])
    assert converged >= 50, converged


def test_aca_recovers_a_low_rank_matrix_and_gives_up_on_a_full_rank_one():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(120, 5)) @ rng.normal(size=(5, 90))
    f, used = H.aca(lambda r, c: A[np.ix_(r, c)], np.arange(120), np.arange(90), 1e-10)
    assert f is not None and f[0].shape == (120, 5) and f[1].shape == (5, 90)
    assert np.linalg.norm(f[0] @ f[1] - A) <= 1e-12 * np.linalg.norm(A)
    assert used < A.size                            # partial pivoting: not every entry is read
    B = rng.normal(size=(40, 40))
    f, _ = H.aca(lambda r, c: B[np.ix_(r, c)], np.arange(40), np.arange(40), 1e-6)
    assert f is None
    # a zero block is rank 0
    f, _ = H.aca(lambda r, c: np.zeros((len(r), len(c))), np.arange(30), np.arange(30), 1e-6)
    assert f is not None and f[0].shape == (30, 0)


def test_draws_are_distinct_and_follow_the_stream():
    state = H.block_seed(7, 11, 100, 64)
    s2, first = H.draw(state, 100, 16)
    assert len(set(first)) == 16 and all(0 <= k < 100 for k in first)
    _, again = H.draw(state, 100, 16)
    assert again == first
    _, second = H.draw(s2, 100, 16)
    assert second != first
    _, everything = H.draw(state, 10, 10)           # a full shuffle is a permutation
    assert sorted(everything) == list(range(10))


def test_build_covers_every_entry_once_within_tolerance():
    """The reference's kernel (tests/test_compressed.py:10-19) on 1,500 points, leaf 32, tol 1e-6:
    at its 512 points and leaf 16 no admissible block of the checked ACA converges within the rank
    cap (the largest, 32 x 64, needs rank 12 of a cap of 14 even for a truncated SVD)."""
    rng, pts, A = H.reference_kernel(n=1500)
    blocks, diagonal, _ = H.build(lambda a, b: A[np.ix_(a, b)], pts, pts, tol=1e-6, leaf=32)
    assert (H.coverage(blocks, 1500) == 1).all()
    assert sum(v is not None for *_, v, _ in blocks) > 100
    x = rng.normal(size=(1500, 3))
    y = A @ x
    assert np.linalg.norm(H.matmat(blocks, 1500, x) - y) <= 50 * 1e-6 * np.linalg.norm(y)
    np.testing.assert_allclose(diagonal, np.diag(A))
