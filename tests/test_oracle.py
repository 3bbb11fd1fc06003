"""The oracle pinned: the C restatement equals the reference core bit for bit; the closed-form
inner integral equals brute-force quadrature; the far rule converges to the exact 4-D integral;
the kinds rule is the explicit formula; the combine equals the sparse Galerkin products."""
import numpy as np
import pytest
from scipy import sparse

from oracle import fill as F
from tests.meshes import two_layer

CORE = F.ref_core()
needs_ref = pytest.mark.skipif(CORE is None, reason='reference core (oracle/_ref) not built')


@needs_ref
@pytest.mark.parametrize('level', [0, 1, 2])
@pytest.mark.parametrize('scalar_only', [False, True])
@pytest.mark.parametrize('near', [0.0, 0.5, 3.0, np.inf])
def test_c_restatement_is_the_reference_core_bitwise(mesh, level, scalar_only, near):
    g = F.prepare(*mesh, level)
    ours, pairs, analytic = F.cr_fill(g, near, level, scalar_only, workers=2)
    ref, rpairs, ranalytic = F.ref_fill(CORE, g, near, level, scalar_only, workers=2)
    assert (pairs, analytic) == (rpairs, ranalytic)
    assert all(np.array_equal(a, b) for a, b in zip(ours, ref))


@needs_ref
def test_c_restatement_on_a_two_layer_mesh_bitwise():
    g = F.prepare(*two_layer(16, 10, seed=3), 1)
    ours, pairs, analytic = F.cr_fill(g, 3.0, 1, False, workers=4)
    ref, rpairs, ranalytic = F.ref_fill(CORE, g, 3.0, 1, False, workers=4)
    assert (pairs, analytic) == (rpairs, ranalytic) and 0 < analytic < pairs
    assert all(np.array_equal(a, b) for a, b in zip(ours, ref))
    n = len(g.tris)
    p, q = np.meshgrid(np.arange(0, n, 7), np.arange(0, n, 5), indexing='ij')
    p, q = p.ravel().astype(np.int32), q.ravel().astype(np.int32)
    d = g.z[p] - g.z[q]
    kind = F.kinds(g, p, q, d, 3.0, 1)
    out = F.cr_pairs(g, p, q, d, kind, False, workers=2)
    assert np.array_equal(out, CORE.static_pairs(F.ref_geometry_dict(g), p, q, d, kind, False, 2))


@pytest.mark.parametrize('level', [0, 1, 2])
@pytest.mark.parametrize('scalar_only', [False, True])
def test_c_restatement_matches_numpy_restatement(mesh, level, scalar_only):
    g = F.prepare(*mesh, level)
    n = len(g.tris)
    p, q = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    p, q = p.ravel(), q.ravel()
    d = g.z[p] - g.z[q]
    kind = F.kinds(g, p, q, d, 3.0, level)
    a = F.cr_pairs(g, p, q, d, kind, scalar_only)
    b = F.numpy_pairs(g, p, q, d, kind, scalar_only)
    assert a == pytest.approx(b, rel=2e-11, abs=1e-27)


def _duffy_inner(P, V, d, n=48):
    """Brute force of int_T 1/R and int_T r'/R over triangle V (3,2) at point P (2,), offset d:
    three Duffy-mapped subtriangles from P (signed areas), Gauss-Legendre n x n each."""
    x, w = np.polynomial.legendre.leggauss(n)
    u, wu = (x + 1) / 2, w / 2
    U, Vv = np.meshgrid(u, u, indexing='ij')
    W = np.outer(wu, wu)
    I0, I1 = 0.0, np.zeros(2)
    for a, b in ((0, 1), (1, 2), (2, 0)):
        A, B = V[a] - P, V[b] - V[a]
        cross = A[0] * B[1] - A[1] * B[0]
        rel = U[..., None] * (A + Vv[..., None] * B)            # r' - P
        R = np.sqrt((rel ** 2).sum(-1) + d * d)
        jac = U * cross
        I0 += (W * jac / R).sum()
        I1 += ((W * jac / R)[..., None] * (rel + P)).sum((0, 1))
    return I0, I1


@pytest.mark.parametrize('d', [0.0, 3e-4, -1e-3])
def test_closed_form_inner_integral_equals_brute_force(d):
    rng = np.random.default_rng(7)
    nodes = np.array([[0., 0.], [2e-3, 0.3e-3], [0.6e-3, 1.7e-3], [3e-3, 2e-3]])
    tris = np.array([[0, 1, 2], [1, 3, 2]])
    g = F.prepare(nodes, tris, np.zeros(2), 0)
    V, L, tau, un, area, Pts, W = g.levels[0]
    # single observation points: replace the level-0 rule by a delta at chosen points
    for point in [g.nodes[tris[0]].mean(0), g.nodes[tris[0]].mean(0) + [4e-4, -2e-4],
                  np.array([5e-3, -1e-3]), g.nodes[tris[0][0]] * 0.5 + g.nodes[tris[0][1]] * 0.5]:
        geo = F.prepare(nodes, tris, np.zeros(2), 0)
        lev = list(geo.levels[0])
        lev[5] = np.broadcast_to(point, (2, 6, 2)).copy()
        lev[6] = np.array([1.0, 0, 0, 0, 0, 0])
        lev[4] = np.ones(2)
        geo.levels[0] = tuple(lev)
        m = F.analytic_pairs(geo, np.array([0]), np.array([0]), np.array([d]), 0)[0]
        i0, i1 = _duffy_inner(point, geo.nodes[tris[0]], d)
        assert m[0] == pytest.approx(i0, rel=1e-9)
        assert m[3:5] == pytest.approx(i1, rel=1e-8, abs=1e-9 * abs(i1).max())
        assert m[1:3] == pytest.approx(point * i0, rel=1e-9)


def test_far_rule_converges_to_the_exact_integral():
    # two triangles at growing distance: the 3x3 rule error falls at least as (size/dist)^2
    nodes0 = np.array([[0., 0.], [1e-3, 0.], [0.2e-3, 0.9e-3]])
    errors = []
    for dist in (5e-3, 10e-3, 20e-3):
        nodes = np.r_[nodes0, nodes0 + [dist, 0.3e-3]]
        tris = np.array([[0, 1, 2], [3, 4, 5]])
        g = F.prepare(nodes, tris, np.zeros(2), 2)
        far = F.far_pairs(g, np.array([0]), np.array([1]), np.array([0.0]))[0]
        exact = F.analytic_pairs(g, np.array([0]), np.array([1]), np.array([0.0]), 2)[0]
        errors.append(np.abs(far - exact).max() / np.abs(exact).max())
    assert errors[0] < 2e-3 and errors[1] < errors[0] / 3.5 and errors[2] < errors[1] / 3.5


def test_refined_self_integral_converges():
    nodes = np.array([[0., 0.], [1., 0.], [0., 1.]])
    tris = np.array([[0, 1, 2]])
    vals = []
    for k in (0, 1, 2):
        g = F.prepare(nodes, tris, np.zeros(1), k)
        vals.append(F.analytic_pairs(g, np.array([0]), np.array([0]), np.array([0.0]), k)[0, 0])
    assert abs(vals[1] - vals[2]) < abs(vals[0] - vals[2]) / 3
    assert vals[1] == pytest.approx(vals[2], rel=.01)


@pytest.mark.parametrize('near', [0., .5, 3., np.inf])
@pytest.mark.parametrize('level', [0, 1, 2])
def test_kinds_rule_is_the_explicit_formula(mesh, near, level):
    g = F.prepare(*mesh, level)
    n = len(mesh[1])
    p, q = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    p, q = p.ravel(), q.ravel()
    d = g.z[p] - g.z[q]
    distance = np.sqrt(((g.cent[p] - g.cent[q]) ** 2).sum(-1) + d * d)
    size = g.radius[p] + g.radius[q]
    close = np.ones(len(p), bool) if np.isinf(near) else distance <= near * size
    expected = np.where(close, np.where(distance <= size, level, 0), -1)
    assert np.array_equal(F.kinds(g, p, q, d, near, level), expected)
    _, pairs, analytic = F.cr_fill(g, near, level, False)
    assert analytic == np.count_nonzero(expected[p <= q] >= 0)


def test_combine_equals_sparse_galerkin_products():
    from oracle import mqs as M
    nodes, tris, z = two_layer(8, 6, seed=1)
    g = F.prepare(nodes, tris, z, 1)
    (m0, mx, my, rr), _, _ = F.cr_fill(g, 3.0, 1, False)
    r = M.RWG(g.nodes, tris)
    A, Bx, By = r.Aal, r.Bx, r.By
    qf = lambda S1, Mm, S2: (S2.T @ (S1.T @ Mm).T).T
    ref = qf(A, rr, A) - qf(A, mx, Bx) - qf(A, my, By) - qf(Bx, mx.T, A) - qf(By, my.T, A) + qf(Bx, m0, Bx) + qf(By, m0, By)
    scale = 1 / r.l
    expected = 1e-7 * 0.5 * (ref + ref.T) * np.outer(scale, scale)
    L = F.cr_combine(r, rr, mx, my, m0, 1e-7, scale)
    assert L == pytest.approx(expected, rel=1e-12, abs=1e-12 * np.abs(expected).max())
    assert np.array_equal(L, L.T)
    assert np.linalg.eigvalsh(L).min() > 0
