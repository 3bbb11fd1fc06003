"""The FFT near-field grid's oracle pinned on the CPU: the deposit keeps each sheet's total current
and its moment (both exact integrals of the linear current), the FFT convolution equals the
direct sum over the grid's point currents, and the grid field agrees with direct Biot-Savart to
1 % on the copper fixture's solution (the reference's own check, tests/test_mqs.py:122-133)."""
import os

import numpy as np
import pytest

from oracle import biot as B
from oracle.fill import BAR6, W6
from tests.meshes import two_layer

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def random_sheets(seed=0, k=2):
    nodes, tris, z_of = two_layer(8, 6, seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    for z in np.unique(z_of):
        t = tris[z_of == z]
        a = rng.normal(size=(len(t), k)) + 1j * rng.normal(size=(len(t), k))
        b = (rng.normal(size=(len(t), 2, k)) + 1j * rng.normal(size=(len(t), 2, k))) * 1e-2
        out.append((nodes, t, a, b, z))
    return out


def exact_moments(sheet):
    """sum over triangles of the integrals of J and of J x, J y (the 6-point rule is exact here)."""
    nodes, tris, a, b, _ = sheet
    V = nodes[tris]
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Q = np.einsum('qk,tkc->tqc', BAR6, V)
    w = area[:, None] * W6[None]
    J = a[:, None, None, :] * Q[..., None] + b[:, None, :, :]             # (t, 6, 2, k)
    total = np.einsum('tq,tqck->ck', w, J)
    first = np.einsum('tq,tqd,tqck->dck', w, Q, J)
    return total, first


def test_deposit_keeps_total_current_and_moment():
    for sheet in random_sheets():
        gx, gy = np.linspace(-0.003, 0.025, 57), np.linspace(-0.003, 0.016, 39)
        J = B.current_deposit(sheet, gx, gy, per_cell=1.0)               # splits the larger triangles
        dA = (gx[1] - gx[0]) * (gy[1] - gy[0])
        total, first = exact_moments(sheet)
        np.testing.assert_allclose(J.sum((1, 2)) * dA, total, rtol=1e-12, atol=1e-14 * np.abs(total).max())
        X, Y = np.meshgrid(gx, gy, indexing='ij')
        mom = np.stack([np.einsum('ij,cijk->ck', X, J), np.einsum('ij,cijk->ck', Y, J)]) * dA
        np.testing.assert_allclose(mom, first, rtol=1e-11, atol=1e-13 * np.abs(first).max())


def test_fft_convolution_equals_the_direct_sum():
    sheets = random_sheets(1, k=1)
    x, y, heights = np.linspace(0.004, 0.012, 5), np.linspace(0.002, 0.008, 4), [0.0015, -0.001]
    H = B.h_grid(sheets, x, y, heights)[..., 0]                           # one excitation, as columns
    (gx, gy), _, (dx, dy) = B.grid_setup(sheets, x, y)
    ref = np.zeros_like(H)
    X, Y = np.meshgrid(gx, gy, indexing='ij')
    for sh in sheets:
        J = B.current_deposit(sh, gx, gy)[..., 0]                         # (2, nx, ny)
        for h, z in enumerate(heights):
            for i, xo in enumerate(x):
                for j, yo in enumerate(y):
                    R = np.stack([xo - X, yo - Y, np.full_like(X, z - sh[4])])
                    r3 = (R ** 2).sum(0) ** 1.5
                    c = dx * dy / (4 * np.pi * r3)
                    ref[i, j, h] += [np.sum(c * J[1] * R[2]), -np.sum(c * J[0] * R[2]),
                                     np.sum(c * (J[0] * R[1] - J[1] * R[0]))]
    np.testing.assert_allclose(H, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def copper_solution():
    from oracle import mqs as M
    from pcbqs_amd.geometry import RWG
    with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
        expected = {k: data[k] for k in data.files}
    q = M.copper_reference_model(expected['nodes'], expected['tris'])
    q.assemble()
    _, currents = q.port_response([0, 1e5])
    i = currents[:, :, 0].T                                                 # (branches, 2 frequencies)
    r = RWG(q.nodes, q.tris)
    a, b = B.rwg_currents(r, i[:q.ne] / r.l[:, None])
    sheets = [(q.nodes, q.tris[idx], a[idx], b[idx], z) for idx, z in q.sheets()]
    v = np.array(q.vseg)
    segs = (np.c_[v[:, 0], v[:, 1], v[:, 3]], np.c_[v[:, 0], v[:, 1], v[:, 4]], i[q.ne:q.ne + q.nv])
    return sheets, segs


@pytest.mark.parametrize('refine', [1, 2])
def test_grid_field_matches_biot_savart_on_the_copper_fixture(refine):
    sheets, segs = copper_solution()
    x, y, heights = np.linspace(.012, .018, 25), np.linspace(.007, .013, 25), [.0025, -.002]
    H = B.h_grid(sheets, x, y, heights, segs, refine=refine)
    X, Y, Z = np.meshgrid(x[::6], y[::6], heights, indexing='ij')
    pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    ref = (B.h_sheets(sheets, pts) + B.h_segments(*segs, pts)).reshape(5, 5, 2, 3, 2)
    err = np.abs(H[::6, ::6] - ref).max() / np.abs(ref).max()
    print(f'refine {refine}: max error {err:.2e} of the largest field')
    assert err < 0.01
