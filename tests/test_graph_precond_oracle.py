"""The graph preconditioner's oracle pinned by its defining property: one application inverts
C_f^T D C_f (formed densely) for a diagonal D and for D with coupled branch pairs; and the
product's host set-up (incidence, kept nodes, cotree branches) equals the oracle's restatement of
the reference's."""
import numpy as np
import pytest

from oracle import precond as O
from pcbqs_amd import board as boards
from pcbqs_amd.loops import LoopModel
from pcbqs_amd.precond import _inverse_impedance, graph_setup


def model_and_setup(make):
    b = make()
    m = LoopModel(b)
    p = O.setup(b.n_nodes, m.ends, m.is_src, m.basis.C, m.basis.tree, np.arange(m.n_free))
    return m, p


def random_d(m, rng):
    d = rng.uniform(1, 2, m.nb) + 1j * rng.uniform(0.5, 3, m.nb)
    return d


@pytest.mark.parametrize('make', [lambda: boards.board_c1(target=400), lambda: boards.board_batch(0, (800, 1200))])
@pytest.mark.parametrize('coupled', [False, True])
def test_one_application_inverts_the_loop_system(make, coupled):
    rng = np.random.default_rng(1)
    m, p = model_and_setup(make)
    d = random_d(m, rng)
    zpair = np.zeros(0, complex)
    if coupled:
        # coupled pairs among branches that carry loop current (a branch inside one merged terminal
        # node carries none, and a coupling to it would make D's inverse differ from C_f^T D C_f's)
        carrying = np.flatnonzero(np.asarray(abs(m.basis.C.tocsr()[:, :m.n_free]).sum(1)).ravel() > 0)
        cand = rng.permutation(np.intersect1d(carrying, p['passive']))[:40]
        p = dict(p, pairs=np.stack([cand[:20], cand[20:]]))
        zpair = 0.3 * (d[cand[:20]] + d[cand[20:]]) / 2
    M = O.graph(p, d, zpair)
    Dm = O.dinv(p, d, zpair)
    # D itself (its passive part), from D^-1 block by block
    D = np.diag(d).astype(complex)
    if coupled:
        e, g = p['pairs']
        D[e, g] = D[g, e] = zpair
    Cf = m.basis.C.tocsc()[:, :m.n_free].toarray()
    K = Cf.T @ D @ Cf
    y = rng.normal(size=(m.n_free, 3)) + 1j * rng.normal(size=(m.n_free, 3))
    np.testing.assert_allclose(M @ (K @ y), y, rtol=1e-9, atol=1e-9 * np.abs(y).max())
    assert Dm.shape == (m.nb, m.nb)


@pytest.mark.parametrize('make', [lambda: boards.board_c1(target=400), lambda: boards.board_batch(0, (800, 1200))])
def test_host_setup_equals_the_oracle(make):
    m, p = model_and_setup(make)
    q = graph_setup(m)
    assert np.array_equal(q['cotree'], p['cotree'])
    assert np.array_equal(q['keep'], p['keep'])
    assert np.array_equal(q['passive'], p['passive'])
    assert abs(q['A'] - p['A']).max() == 0
    rng = np.random.default_rng(2)
    d = random_d(m, rng)
    cand = rng.permutation(p['passive'])[:10]
    pp, qq = dict(p, pairs=np.stack([cand[:5], cand[5:]])), dict(q, pairs=np.stack([cand[:5], cand[5:]]))
    z = rng.normal(size=5) + 0j
    assert abs(_inverse_impedance(qq, d, z) - O.dinv(pp, d, z)).max() <= 1e-15 * abs(O.dinv(pp, d, z)).max()


def test_laplacian_beyond_the_dense_limit_is_refused_before_any_device_call():
    from scipy import sparse
    from pcbqs_amd.precond import MAX_LAPLACIAN_NODES, _DeviceLaplacian
    lap = sparse.identity(MAX_LAPLACIAN_NODES + 1, dtype=complex, format='csr')
    with pytest.raises(ValueError, match='kept nodes exceed'):
        _DeviceLaplacian(lap, 0)
