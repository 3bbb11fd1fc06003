"""Oracle for Biot-Savart field reconstruction (test infrastructure only; see oracle/__init__.py).

Restates, in numpy/scipy:
  rwg_currents   src/pcbqs/_numerics/currents.py:5-18  (J = a r + b per triangle)
  h_tris         src/pcbqs/biot.py:24-51               (6-point rule per triangle)
  h_sheets       src/pcbqs/biot.py:54-79               (near triangles split 4x, twice)
  h_segments     src/pcbqs/biot.py:82-105              (finite straight filaments)

Pinned by: the copper-reference.npz field (rtol 1e-10) and the magnetic-reference.npz spectral
field (3 %), tests/test_copper_reference.py.
"""
import numpy as np
from scipy.spatial import cKDTree

from .fill import BAR6, W6, split4


def rwg_currents(rwg, I):
    I = np.asarray(I)
    multi = I.ndim > 1
    I2 = I if multi else I[:, None]
    k = I2.shape[1]
    a = np.zeros((len(rwg.tris), k), complex)
    b = np.zeros((len(rwg.tris), 2, k), complex)
    for t, alpha, v in ((rwg.tp, rwg.l / (2 * rwg.area[rwg.tp]), rwg.vp),
                        (rwg.tm, -rwg.l / (2 * rwg.area[rwg.tm]), rwg.vm)):
        ia = I2 * alpha[:, None]
        np.add.at(a, t, ia)
        np.add.at(b, t, -ia[:, None, :] * v[:, :, None])
    return (a, b) if multi else (a[:, 0], b[:, :, 0])


def h_tris(V, a, b, z, pts):
    multi = np.ndim(a) > 1
    a2 = a if multi else a[:, None]
    b2 = b if multi else b[:, :, None]
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Q = np.einsum('qk,tkc->tqc', BAR6, V)
    w = (area[:, None] * W6[None]).reshape(-1)
    Jx = (a2[:, None, :] * Q[..., 0:1] + b2[:, None, 0, :]).reshape(-1, a2.shape[1]) * w[:, None]
    Jy = (a2[:, None, :] * Q[..., 1:2] + b2[:, None, 1, :]).reshape(-1, a2.shape[1]) * w[:, None]
    sx, sy, sz = Q[..., 0].reshape(-1), Q[..., 1].reshape(-1), np.repeat(z, 6)
    H = np.zeros((len(pts), 3, a2.shape[1]), complex)
    step = max(1, 4_000_000 // max(len(sx), 1))
    for p0 in range(0, len(pts), step):
        P = pts[p0:p0 + step]
        Rx, Ry, Rz = P[:, 0:1] - sx[None], P[:, 1:2] - sy[None], P[:, 2:3] - sz[None]
        r3 = (Rx * Rx + Ry * Ry + Rz * Rz) ** 1.5
        inv = np.where(r3 > 1e-30, 1.0 / np.where(r3 > 1e-30, r3, 1.0), 0.0)
        Az, Ay, Ax = Rz * inv, Ry * inv, Rx * inv
        H[p0:p0 + step, 0] = Az @ Jy
        H[p0:p0 + step, 1] = -(Az @ Jx)
        H[p0:p0 + step, 2] = Ay @ Jx - Ax @ Jy
    H /= 4 * np.pi
    return H if multi else H[:, :, 0]


def far_mask(V, z, pts, near=1.5):
    """Triangles whose nearest point is beyond near x 2 x size (biot.py:64-68)."""
    size = np.sqrt(np.abs((V[:, 1, 0] - V[:, 0, 0]) * (V[:, 2, 1] - V[:, 0, 1]) -
                          (V[:, 1, 1] - V[:, 0, 1]) * (V[:, 2, 0] - V[:, 0, 0])) / 2)
    cen = V.mean(1)
    tree = cKDTree(np.c_[pts[:, :2], (pts[:, 2] - z)])
    dist, _ = tree.query(np.c_[cen, np.zeros(len(cen))])
    return dist > near * 2 * size


def h_sheets(sheets, pts, near=1.5, levels=2):
    pts = np.asarray(pts, float)
    H = None
    for nodes, tris, a, b, z in sheets:
        V = nodes[tris]
        far = far_mask(V, z, pts, near)
        h = h_tris(V[far], a[far], b[far], np.full(far.sum(), z), pts)
        Vn, an, bn = V[~far], a[~far], b[~far]
        for _ in range(levels):
            if not len(Vn):
                break
            Vn = split4(Vn)
            an, bn = np.concatenate([an] * 4), np.concatenate([bn] * 4)
        if len(Vn):
            h = h + h_tris(Vn, an, bn, np.full(len(Vn), z), pts)
        H = h if H is None else H + h
    return H


def h_segments(p0, p1, I, pts):
    p0, p1, pts = np.asarray(p0, float), np.asarray(p1, float), np.asarray(pts, float)
    I = np.asarray(I)
    multi = I.ndim > 1
    I2 = I if multi else I[:, None]
    H = np.zeros((len(pts), 3, I2.shape[1] if len(p0) else 1), complex)
    if len(p0):
        L = p1 - p0
        ln = np.linalg.norm(L, axis=1)
        u = L / ln[:, None]
        for k in range(len(p0)):
            r0 = pts - p0[k]
            s_ = r0 @ u[k]
            perp = r0 - s_[:, None] * u[k]
            d = np.linalg.norm(perp, axis=1)
            ok = d > 1e-9
            c1 = s_ / np.sqrt(s_ ** 2 + d ** 2)
            c2 = (s_ - ln[k]) / np.sqrt((s_ - ln[k]) ** 2 + d ** 2)
            g = np.where(ok, (c1 - c2) / (4 * np.pi * np.where(ok, d, 1)), 0)
            dirn = np.cross(u[k], np.where(ok[:, None], perp / np.where(ok, d, 1)[:, None], 0))
            H += (dirn * g[:, None])[:, :, None] * I2[k][None, None, :]
    return H if multi else H[:, :, 0]
