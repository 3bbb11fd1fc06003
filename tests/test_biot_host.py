"""The host half of the device Biot-Savart: currents and the refined source list.

The 6-point rule integrates the linear J = a r + b exactly, so the weighted sum of the sources'
currents equals sum over triangles of area x J(centroid) whether or not triangles are refined."""
import numpy as np

from oracle import biot as B


def test_source_currents_integrate_j_exactly_with_and_without_refinement():
    from pcbqs_amd import biot as D
    from pcbqs_amd.geometry import RWG
    from tests.meshes import two_layer
    nodes, tris, z_of = two_layer(6, 5)
    rng = np.random.default_rng(0)
    r = RWG(nodes, tris)
    I = rng.normal(size=(r.Ne, 2)) + 1j * rng.normal(size=(r.Ne, 2))
    a, b = D.rwg_currents(r, I)
    ao, bo = B.rwg_currents(r, I)
    assert np.allclose(a, ao) and np.allclose(b, bo)
    z = z_of.max()
    sel = z_of == z
    sheets = [(nodes, tris[sel], a[sel], b[sel], z)]
    V = nodes[tris[sel]]
    area = 0.5 * np.abs((V[:, 1, 0] - V[:, 0, 0]) * (V[:, 2, 1] - V[:, 0, 1]) -
                        (V[:, 1, 1] - V[:, 0, 1]) * (V[:, 2, 0] - V[:, 0, 0]))
    cen = V.mean(1)
    want_x = (area[:, None] * (a[sel] * cen[:, 0:1] + b[sel][:, 0])).sum(0)
    want_y = (area[:, None] * (a[sel] * cen[:, 1:2] + b[sel][:, 1])).sum(0)
    far_pts = np.array([[0.0, 0.0, 1.0]])                       # nothing refined
    near_pts = np.c_[cen[:5], np.full(5, z)]                     # some triangles refined twice
    K = 2
    for pts in (far_pts, near_pts):
        S = D.sheet_sources(sheets, pts, K)
        jx = S[:, 3] + 1j * S[:, 4], S[:, 7] + 1j * S[:, 8]
        jy = S[:, 5] + 1j * S[:, 6], S[:, 9] + 1j * S[:, 10]
        for c in range(2):
            assert abs(jx[c].sum() - want_x[c]) <= 1e-12 * np.abs(want_x).max()
            assert abs(jy[c].sum() - want_y[c]) <= 1e-12 * np.abs(want_y).max()
    assert len(D.sheet_sources(sheets, near_pts, K)) > len(D.sheet_sources(sheets, far_pts, K))
