"""Biot-Savart on the device against the oracle and the reference's frozen fields: the copper
fixture's H (rtol 1e-10, FP64), the spectral magnetic reference (3 %), and random sheets, filaments
and several excitations (FP64 to 1e-11, FP32 to 2e-5 of the largest field)."""
import os

import numpy as np
import pytest

from oracle import biot as B
from oracle import mqs as M

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.fixture(scope='module')
def copper():
    with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
        expected = {k: data[k] for k in data.files}
    q = M.copper_reference_model(expected['nodes'], expected['tris'])
    q.assemble()
    return q, expected


def test_copper_fixture_field(copper):
    from pcbqs_amd import biot as D
    from pcbqs_amd.geometry import RWG
    q, expected = copper
    _, currents = q.port_response([0, 1e5, 2e5])
    i = np.einsum('fbp,fp->fb', currents, np.ones((3, 1)))[[1]].T
    r = RWG(q.nodes, q.tris)
    a, b = D.rwg_currents(r, i[:q.ne] / r.l[:, None])
    sheets = [(q.nodes, q.tris[idx], a[idx], b[idx], z) for idx, z in q.sheets()]
    pts = np.array([[.015, .010, .002]])
    v = np.array(q.vseg)
    H = D.h_field(sheets, pts, (np.c_[v[:, 0], v[:, 1], v[:, 3]], np.c_[v[:, 0], v[:, 1], v[:, 4]],
                                i[q.ne:q.ne + q.nv]))
    np.testing.assert_allclose(H[:, :, 0], expected['H'], rtol=1e-10)


def test_spectral_magnetic_reference():
    from pcbqs_amd import biot as D
    from pcbqs_amd.geometry import RWG
    with np.load(os.path.join(GOLDEN, 'magnetic-reference.npz')) as data:
        nodes, tris, I, pts, H = [data[k] for k in ['nodes', 'tris', 'I', 'pts', 'H']]
    r = RWG(nodes, tris)
    a, b = D.rwg_currents(r, I)
    Hb = D.h_sheets([(nodes, tris, a, b, .5e-3), (nodes, tris, -a, -b, -.5e-3)], pts)
    for h, hb in zip(H, Hb):
        assert np.abs(h - hb).max() < .03 * np.abs(hb).max() + 1e-3 * np.abs(Hb).max()


@pytest.mark.parametrize('k', [1, 3, 8])
def test_random_sheets_and_filaments_match_the_oracle(k):
    from pcbqs_amd import biot as D
    from pcbqs_amd.geometry import RWG
    from tests.meshes import two_layer
    nodes, tris, z_of = two_layer(8, 6)
    rng = np.random.default_rng(k)
    r = RWG(nodes, tris)
    I = rng.normal(size=(r.Ne, k)) + 1j * rng.normal(size=(r.Ne, k))
    a, b = D.rwg_currents(r, I)
    ao, bo = B.rwg_currents(r, I)
    np.testing.assert_allclose(a, ao, rtol=1e-13, atol=1e-13 * np.abs(ao).max())
    sheets = [(nodes, tris[z_of == z], a[z_of == z], b[z_of == z], z) for z in np.unique(z_of)]
    lo, hi = nodes.min(0), nodes.max(0)
    pts = np.c_[rng.uniform(lo[0], hi[0], 300), rng.uniform(lo[1], hi[1], 300), rng.uniform(-1e-3, 2e-3, 300)]
    pts[:20, 2] = z_of.max() + 1e-5                                # near points: refined triangles
    p0 = np.c_[rng.uniform(lo[0], hi[0], 5), rng.uniform(lo[1], hi[1], 5), np.zeros(5)]
    p1 = p0 + np.c_[np.zeros(5), np.zeros(5), np.full(5, 2e-4)]
    Iv = rng.normal(size=(5, k)) + 1j * rng.normal(size=(5, k))
    want = B.h_sheets(sheets, pts) + B.h_segments(p0, p1, Iv, pts)
    got = D.h_field(sheets, pts, (p0, p1, Iv))
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-11 * scale
    got32 = D.h_field(sheets, pts, (p0, p1, Iv), precision=32)
    assert np.abs(got32 - want).max() <= 2e-5 * scale
    # segments alone
    np.testing.assert_allclose(D.h_segments(p0, p1, Iv, pts), B.h_segments(p0, p1, Iv, pts), rtol=1e-10,
                               atol=1e-12 * np.abs(B.h_segments(p0, p1, Iv, pts)).max())
