"""The oracle's whole chain against the reference's frozen copper-reference.npz (the parity test of
tests/test_standalone.py:18-31): stored mesh -> fill -> combine -> R -> vias -> loops -> port Z ->
branch currents -> H, at the reference's own tolerances; and Biot-Savart against the
magnetic-reference.npz spectral field (tests/test_mqs.py:141-151)."""
import os

import numpy as np
import pytest

from oracle import biot as B
from oracle import mqs as M

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.fixture(scope='module')
def copper():
    with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
        expected = {k: data[k] for k in data.files}
    q = M.copper_reference_model(expected['nodes'], expected['tris'])
    q.assemble()
    return q, expected


def test_oracle_reproduces_the_frozen_port_impedance_and_currents(copper):
    q, expected = copper
    Z, currents = q.port_response([0, 1e5, 2e5])
    np.testing.assert_allclose(Z, expected['impedance'], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(currents, expected['currents'], rtol=1e-10, atol=1e-11)
    R, L = M.rl_equivalent(np.array([0, 1e5, 2e5]), Z)
    assert R[0, 0] > 0 and L[0, 0] > 0


def test_oracle_reproduces_the_frozen_field(copper):
    q, expected = copper
    _, currents = q.port_response([0, 1e5, 2e5])
    i = np.einsum('fbp,fp->fb', currents, np.ones((3, 1)))[[1]].T
    r = M.RWG(q.nodes, q.tris)
    a, b = B.rwg_currents(r, i[:q.ne] / r.l[:, None])
    sheets = [(q.nodes, q.tris[idx], a[idx], b[idx], z) for idx, z in q.sheets()]
    pts = np.array([[.015, .010, .002]])
    v = np.array(q.vseg)
    H = B.h_sheets(sheets, pts) + B.h_segments(np.c_[v[:, 0], v[:, 1], v[:, 3]], np.c_[v[:, 0], v[:, 1], v[:, 4]],
                                               i[q.ne:q.ne + q.nv], pts)
    np.testing.assert_allclose(H[:, :, 0], expected['H'], rtol=1e-10)


def test_port_impedance_does_not_depend_on_the_loop_basis(copper):
    q, expected = copper
    m = q.mats
    # another loop basis: a random unimodular recombination of the free loops keeps Z
    rng = np.random.default_rng(0)
    C = m['C'].toarray()
    offset = len(q.br)
    sc = m['src'][offset]
    free = [c for c in range(C.shape[1]) if c != sc]
    T = np.eye(C.shape[1])
    for _ in range(200):
        i, j = rng.choice(free, 2, replace=False)
        T[:, i] += T[:, j] * rng.choice([-1, 1])
    C2 = C @ T
    Rl = C2.T @ m['R'] @ C2
    Ll = C2[:q.ne].T @ m['Le'] @ C2[:q.ne] + C2[q.ne:q.ne + q.nv].T @ m['Lv'] @ C2[q.ne:q.ne + q.nv]
    from scipy import sparse
    Z2, _ = M.direct_ports(sparse.csc_matrix(C2), Rl, Ll, np.array([sc]), [1e5])
    np.testing.assert_allclose(Z2[0], expected['impedance'][1], rtol=1e-9)


def test_biot_savart_matches_the_spectral_reference():
    with np.load(os.path.join(GOLDEN, 'magnetic-reference.npz')) as data:
        nodes, tris, I, pts, H = [data[k] for k in ['nodes', 'tris', 'I', 'pts', 'H']]
    r = M.RWG(nodes, tris)
    a, b = B.rwg_currents(r, I)
    Hb = B.h_sheets([(nodes, tris, a, b, .5e-3), (nodes, tris, -a, -b, -.5e-3)], pts)
    for h, hb in zip(H, Hb):
        assert np.abs(h - hb).max() < .03 * np.abs(hb).max() + 1e-3 * np.abs(Hb).max()


def test_via_partial_inductance_is_rosas():
    h, a = 1.6e-3, 0.15e-3
    L = M.filament_mutual(0, h, 0, h, a)
    rosa = M.MU0 / (2 * np.pi) * (h * np.log((h + np.sqrt(h * h + a * a)) / a) - np.sqrt(h * h + a * a) + a)
    assert L == pytest.approx(rosa, rel=1e-9)
