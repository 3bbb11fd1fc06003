"""Oracle for Biot-Savart field reconstruction (test infrastructure only; see oracle/__init__.py).

Restates, in numpy/scipy:
  rwg_currents   src/pcbqs/_numerics/currents.py:5-18  (J = a r + b per triangle)
  h_tris         src/pcbqs/biot.py:24-51               (6-point rule per triangle)
  h_sheets       src/pcbqs/biot.py:54-79               (near triangles split 4x, twice)
  h_segments     src/pcbqs/biot.py:82-105              (finite straight filaments)
  current_deposit src/pcbqs/_numerics/currents.py:28-61 (split, 6-point moments, cloud in cell)
  h_grid         src/pcbqs/biot.py:108-181             (linear FFT convolution with the kernel)

current_deposit and h_grid are pinned by tests/test_biot_grid_oracle.py: the deposit keeps each
sheet's total current and current moment exactly (the rule is exact for linear currents, bilinear
weights keep first moments); the FFT convolution equals the direct convolution sum; and the grid
field agrees with h_sheets (direct Biot-Savart) to 1 % on the copper fixture's solution, the
reference's own check (tests/test_mqs.py:122-133).

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


def current_deposit(sheet, gx, gy, per_cell=2.0):
    nodes, tris, a, b, _ = sheet
    multi = np.ndim(a) > 1
    a = np.asarray(a).reshape(len(tris), -1)
    b = np.asarray(b).reshape(len(tris), 2, -1)
    V = np.asarray(nodes, float)[np.asarray(tris)][:, :, :2]
    dx, dy = gx[1] - gx[0], gy[1] - gy[0]
    for _ in range(6):
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        size = np.sqrt(np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) / 2)
        big = size > per_cell * min(dx, dy)
        if not big.any():
            break
        V = np.concatenate([V[~big], split4(V[big])])
        a = np.concatenate([a[~big], np.tile(a[big], (4, 1))])
        b = np.concatenate([b[~big], np.tile(b[big], (4, 1, 1))])
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Q = np.einsum('qk,tkc->tqc', BAR6, V)
    mom = (a[:, None, None, :] * Q[..., None] + b[:, None, :, :]) * (area[:, None] * W6[None])[..., None, None]
    fx, fy = (Q[..., 0].ravel() - gx[0]) / dx, (Q[..., 1].ravel() - gy[0]) / dy
    i0, j0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
    tx, ty = fx - i0, fy - j0
    nx, ny = len(gx), len(gy)
    out = np.zeros((2, nx, ny, a.shape[1]), complex)
    for di, dj, wgt in ((0, 0, (1 - tx) * (1 - ty)), (1, 0, tx * (1 - ty)), (0, 1, (1 - tx) * ty), (1, 1, tx * ty)):
        ii, jj = i0 + di, j0 + dj
        ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny)
        for c in range(2):
            np.add.at(out[c], (ii[ok], jj[ok]), mom[:, :, c, :].reshape(-1, a.shape[1])[ok] * wgt[ok, None])
    out /= dx * dy
    return out if multi else out[..., 0]


def grid_setup(sheets, x, y, refine=1):
    dx, dy = (x[1] - x[0]) / refine, (y[1] - y[0]) / refine
    pts = np.concatenate([np.asarray(sh[0], float)[np.unique(np.asarray(sh[1]))][:, :2] for sh in sheets])
    grids, offsets = [], []
    for dim, axis, step in ((0, x, dx), (1, y, dy)):
        lo = min(0, int(np.floor((pts[:, dim].min() - axis[0]) / step)) - 2)
        hi = max(refine * (len(axis) - 1), int(np.ceil((pts[:, dim].max() - axis[0]) / step)) + 2)
        grids.append(axis[0] + np.arange(lo, hi + 1) * step)
        offsets.append(-lo)
    return grids, offsets, (dx, dy)


def h_grid(sheets, x, y, heights, segments=None, refine=1, per_cell=2.0):
    from scipy.fft import fft2, ifft2, next_fast_len
    x, y, heights = [np.asarray(v, float) for v in (x, y, np.atleast_1d(heights))]
    multi = any(np.ndim(sh[2]) > 1 for sh in sheets)
    (gx, gy), offsets, (dx, dy) = grid_setup(sheets, x, y, refine)
    nx, ny = len(gx), len(gy)
    shape = (next_fast_len(3 * nx - 2), next_fast_len(3 * ny - 2))
    rx = np.arange(1 - nx, nx)[:, None] * dx
    ry = np.arange(1 - ny, ny)[None, :] * dy
    crop = (slice(nx - 1 + offsets[0], nx - 1 + offsets[0] + refine * len(x), refine),
            slice(ny - 1 + offsets[1], ny - 1 + offsets[1] + refine * len(y), refine))
    k = max(np.asarray(sh[2]).reshape(len(sh[1]), -1).shape[1] for sh in sheets)
    out = np.zeros((len(x), len(y), len(heights), 3, k), complex)
    for sh in sheets:
        J = current_deposit(sh, gx, gy, per_cell)
        J = J.reshape(2, nx, ny, -1)
        sjx, sjy = fft2(J[0], s=shape, axes=(0, 1)), fft2(J[1], s=shape, axes=(0, 1))
        for h, height in enumerate(heights):
            dz = height - float(sh[4])
            kern = dx * dy / (4 * np.pi * (rx * rx + ry * ry + dz * dz) ** 1.5)
            kx, ky, kz = [fft2(v * kern, s=shape)[..., None] for v in (rx, ry, dz + 0 * rx)]
            for c, spec in enumerate((sjy * kz, -sjx * kz, sjx * ky - sjy * kx)):
                out[:, :, h, c] += ifft2(spec, axes=(0, 1))[crop]
    if segments is not None and len(segments[0]):
        X, Y, Z = np.meshgrid(x, y, heights, indexing='ij')
        pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
        out += np.asarray(h_segments(*segments, pts)).reshape(len(x), len(y), len(heights), 3, -1)
    return out if multi else out[..., 0]
