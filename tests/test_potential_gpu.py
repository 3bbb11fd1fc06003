"""Potential coefficients P = W^T M0 W / (4 pi eps0 er) on the device (src/pcbqs/mqs.py:532-556)
against the oracle's scalar fill and the explicit sparse product, with merged charge nodes; and the
parallel-plate bracket eps A / h < C < 1.5 eps A / h (the reference's tests/test_peec.py:31-33) of the
capacitance that a charge solve on P gives."""
import numpy as np
import pytest

from oracle import fill as F
from tests.meshes import two_layer

pytestmark = pytest.mark.gpu


def oracle_potential(nodes, tris, z, node_of, er=1.0):
    from pcbqs_amd.geometry import EPS0
    g = F.prepare(nodes, tris, z, 1)
    planes, _, _ = F.cr_fill(g, 3.0, 1, True)
    M0 = planes[0]
    n = node_of.max() + 1
    area = np.bincount(node_of, weights=g.area, minlength=n)
    W = np.zeros((len(tris), n))
    W[np.arange(len(tris)), node_of] = 1 / area[node_of]
    return W.T @ M0 @ W / (4 * np.pi * EPS0 * er)


@pytest.mark.parametrize('precision,tol', [(64, 1e-10), (32, 1e-5)])
def test_potential_matches_the_oracle_with_merged_nodes(precision, tol):
    from pcbqs_amd.static import potential
    nodes, tris, z = two_layer(10, 6)
    nt = len(tris)
    node_of = np.arange(nt)
    node_of[:7] = 0                           # a merged pad of seven triangles
    node_of[nt - 5:] = node_of[nt - 6]        # and one at the end
    node_of = np.unique(node_of, return_inverse=True)[1]
    want = oracle_potential(nodes, tris, z, node_of, er=4.4)
    got = potential(nodes, tris, z, node_of, er=4.4, precision=precision)
    assert got.shape == want.shape
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
    assert np.array_equal(got, got.T)


def test_parallel_plate_capacitance_bracket():
    from pcbqs_amd.geometry import EPS0
    from pcbqs_amd.static import potential
    nodes, tris, z = two_layer(12, 8)
    P = potential(nodes, tris, z, np.arange(len(tris)))
    top = z == z.max()
    q = np.linalg.solve(P, np.where(top, 0.5, -0.5))
    C = q[top].sum()
    g = F.prepare(nodes, tris, z, 1)
    A = min(g.area[top].sum(), g.area[~top].sum())
    h = z.max() - z.min()
    assert EPS0 * A / h < C < 1.5 * EPS0 * A / h
