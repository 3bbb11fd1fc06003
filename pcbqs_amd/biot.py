"""Biot-Savart field reconstruction on the GPU (kernels in csrc/biot.hip).

Mirrors the reference's interface:
  rwg_currents(rwg, I)            src/pcbqs/_numerics/currents.py:5-18  J = a r + b per triangle
  h_sheets(sheets, pts, ...)      src/pcbqs/biot.py:54-79               6-point rule; triangles whose
                                                                       centroid lies within
                                                                       near x 2 x size of some point
                                                                       are split 4x, levels times
  h_segments(p0, p1, I, pts)      src/pcbqs/biot.py:82-105              finite straight filaments
  h_field(sheets, pts, segments)  both in one launch sequence
The host builds the source list (quadrature points with their weighted currents); the device sums
every (point, source) pair. precision=64 is the parity mode, precision=32 the fast one.
"""
import ctypes

import numpy as np
from scipy.spatial import cKDTree

from ._lib import call, ptr
from .geometry import BAR6, W6, _split4


def rwg_currents(rwg, I):
    """Per-triangle (a (nt[, k]), b (nt, 2[, k])) with J = a r + b from RWG coefficients I (Ne[, k])."""
    I = np.asarray(I)
    multi = I.ndim > 1
    I2 = I.reshape(len(I), -1)
    nt = len(rwg.tris)
    a = np.zeros((nt, I2.shape[1]), complex)
    b = np.zeros((nt, 2, I2.shape[1]), complex)
    for tri, alpha, v in ((rwg.tp, rwg.l / (2 * rwg.area[rwg.tp]), rwg.vp),
                          (rwg.tm, -rwg.l / (2 * rwg.area[rwg.tm]), rwg.vm)):
        ia = I2 * alpha[:, None]
        np.add.at(a, tri, ia)
        np.add.at(b, tri, -ia[:, None, :] * v[:, :, None])
    return (a, b) if multi else (a[:, 0], b[:, :, 0])


def _padded(K):
    for k in (1, 2, 4, 8):
        if K <= k:
            return k
    raise ValueError('at most 8 excitations per call')


def sheet_sources(sheets, pts, K, near=1.5, levels=2):
    """(ns, 3 + 4 K) sources: position, then (Jx re, Jx im, Jy re, Jy im) per excitation, each
    current times its quadrature weight."""
    pts = np.asarray(pts, float)
    tree = cKDTree(pts) if len(pts) else None
    out = []
    for nodes, tris, a, b, z in sheets:
        V = np.asarray(nodes, float)[np.asarray(tris)]
        a = np.asarray(a).reshape(len(V), -1)
        b = np.asarray(b).reshape(len(V), 2, -1)
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        size = np.sqrt(np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) / 2)
        cen = V.mean(1)
        if tree is not None:
            dist, _ = tree.query(np.c_[cen, np.full(len(cen), z)])
        else:
            dist = np.full(len(cen), np.inf)
        nearm = ~(dist > near * 2 * size)
        groups = [(V[~nearm], a[~nearm], b[~nearm])]
        Vn, an, bn = V[nearm], a[nearm], b[nearm]
        for _ in range(levels if len(Vn) else 0):
            Vn = _split4(Vn)
            an, bn = np.tile(an, (4, 1)), np.tile(bn, (4, 1, 1))
        groups.append((Vn, an, bn))
        for Vg, ag, bg in groups:
            if not len(Vg):
                continue
            ar = 0.5 * np.abs((Vg[:, 1, 0] - Vg[:, 0, 0]) * (Vg[:, 2, 1] - Vg[:, 0, 1]) -
                              (Vg[:, 1, 1] - Vg[:, 0, 1]) * (Vg[:, 2, 0] - Vg[:, 0, 0]))
            Q = np.einsum('qk,tkc->tqc', BAR6, Vg)                       # (t, 6, 2)
            w = ar[:, None] * W6[None]                                    # (t, 6)
            k = ag.shape[1]
            Jx = (ag[:, None, :] * Q[..., 0:1] + bg[:, None, 0, :]) * w[..., None]
            Jy = (ag[:, None, :] * Q[..., 1:2] + bg[:, None, 1, :]) * w[..., None]
            S = np.zeros((len(Vg) * 6, 3 + 4 * K))
            S[:, 0], S[:, 1], S[:, 2] = Q[..., 0].ravel(), Q[..., 1].ravel(), z
            Jx, Jy = Jx.reshape(-1, k), Jy.reshape(-1, k)
            S[:, 3:3 + 4 * k:4], S[:, 4:4 + 4 * k:4] = Jx.real, Jx.imag
            S[:, 5:5 + 4 * k:4], S[:, 6:6 + 4 * k:4] = Jy.real, Jy.imag
            out.append(S)
    return np.concatenate(out) if out else np.zeros((0, 3 + 4 * K))


def segment_sources(p0, p1, I, K):
    p0, p1 = np.asarray(p0, float).reshape(-1, 3), np.asarray(p1, float).reshape(-1, 3)
    I = np.asarray(I).reshape(len(p0), -1)
    S = np.zeros((len(p0), 6 + 2 * K))
    S[:, 0:3], S[:, 3:6] = p0, p1
    S[:, 6:6 + 2 * I.shape[1]:2], S[:, 7:7 + 2 * I.shape[1]:2] = I.real, I.imag
    return S


def h_field(sheets, pts, segments=None, near=1.5, levels=2, precision=64, device=0, timing=None):
    """H (npts, 3[, k]) complex of sheets [(nodes, tris, a, b, z)] and segments (p0, p1, I)."""
    pts = np.ascontiguousarray(np.asarray(pts, float).reshape(-1, 3))
    ks = [np.asarray(a).reshape(len(np.asarray(tris)), -1).shape[1] for _, tris, a, _, _ in sheets]
    multi = any(np.ndim(a) > 1 for _, _, a, _, _ in sheets)
    if segments is not None:
        Iseg = np.asarray(segments[2])
        ks.append(Iseg.reshape(len(Iseg), -1).shape[1] if len(Iseg) else 1)
        multi = multi or Iseg.ndim > 1
    k = max(ks) if ks else 1
    K = _padded(k)
    src = np.ascontiguousarray(sheet_sources(sheets, pts, K, near, levels))
    seg = np.ascontiguousarray(segment_sources(*segments, K) if segments is not None and len(segments[0])
                               else np.zeros((0, 6 + 2 * K)))
    H = np.zeros((len(pts), 3, K, 2))
    ms = ctypes.c_double()
    call('pq_biot_field', device, precision, K, len(src), ptr(src) if len(src) else None, len(seg),
         ptr(seg) if len(seg) else None, len(pts), ptr(pts), ptr(H), ctypes.byref(ms))
    if timing is not None:
        timing.update(kernel_ms=ms.value, sources=len(src), segments=len(seg), points=len(pts))
    Hc = (H[..., 0] + 1j * H[..., 1])[:, :, :k]
    return Hc if multi else Hc[:, :, 0]


def h_sheets(sheets, pts, near=1.5, levels=2, precision=64, device=0):
    """The reference's h_sheets on the device."""
    return h_field(sheets, pts, None, near, levels, precision, device)


def h_segments(p0, p1, I, pts, precision=64, device=0):
    """The reference's h_segments on the device."""
    return h_field([], pts, (p0, p1, I), precision=precision, device=device)


def current_deposit(sheet, gx, gy, K, per_cell=2.0):
    """A sheet's current as densities Jx, Jy (nx, ny, K) complex (A/m) on the uniform grid gx, gy:
    each triangle split in four while its size exceeds per_cell cells (at most six times), its
    current a r + b integrated by the 6-point rule, each point's moment spread onto the four
    surrounding grid points by bilinear (cloud-in-cell) weights, which keeps the total current and
    its centre (src/pcbqs/_numerics/currents.py:28-61)."""
    nodes, tris, a, b, _ = sheet
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
        V = np.concatenate([V[~big], _split4(V[big])])
        a = np.concatenate([a[~big], np.tile(a[big], (4, 1))])
        b = np.concatenate([b[~big], np.tile(b[big], (4, 1, 1))])
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Q = np.einsum('qk,tkc->tqc', BAR6, V)                               # (M, 6, 2)
    w = (area[:, None] * W6[None]).reshape(-1)
    k = a.shape[1]
    mx = ((a[:, None, :] * Q[..., 0:1] + b[:, None, 0, :]).reshape(-1, k)) * w[:, None]
    my = ((a[:, None, :] * Q[..., 1:2] + b[:, None, 1, :]).reshape(-1, k)) * w[:, None]
    fx, fy = (Q[..., 0].ravel() - gx[0]) / dx, (Q[..., 1].ravel() - gy[0]) / dy
    i0, j0 = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    tx, ty = fx - i0, fy - j0
    nx, ny = len(gx), len(gy)
    J = np.zeros((2, nx * ny, K), complex)
    for di, dj, wgt in ((0, 0, (1 - tx) * (1 - ty)), (1, 0, tx * (1 - ty)), (0, 1, (1 - tx) * ty), (1, 1, tx * ty)):
        ii, jj = i0 + di, j0 + dj
        ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny)
        cell = ii[ok] * ny + jj[ok]
        np.add.at(J[0, :, :k], cell, mx[ok] * wgt[ok, None])
        np.add.at(J[1, :, :k], cell, my[ok] * wgt[ok, None])
    return J.reshape(2, nx, ny, K) / (dx * dy)


def h_grid(sheets, x, y, heights, segments=None, refine=1, per_cell=2.0, device=0, timing=None):
    """Open-boundary H (len(x), len(y), len(heights), 3[, k]) complex on uniform x, y grids at
    absolute heights: every sheet's current deposited on a grid of spacing (x step / refine) that
    covers the observation window and all the copper, convolved with the Biot-Savart kernel by FFT
    on the device (pq_biot_grid: linear, not periodic, convolution), plus the segments' exact field
    (src/pcbqs/biot.py:108-181)."""
    from scipy.fft import next_fast_len
    x, y, heights = [np.asarray(v, float) for v in (x, y, np.atleast_1d(heights))]
    for axis in (x, y):
        if (axis.ndim != 1 or len(axis) < 2 or not np.isfinite(axis).all() or np.diff(axis).min() <= 0 or
                not np.allclose(np.diff(axis), axis[1] - axis[0], rtol=1e-8, atol=1e-15)):
            raise ValueError('x and y must be increasing uniform grids (m)')
    if heights.ndim != 1 or not np.isfinite(heights).all():
        raise ValueError('heights must be finite absolute z coordinates (m)')
    zs = np.array([float(sh[4]) for sh in sheets])
    if len(zs) and np.any(np.isclose(heights[:, None], zs[None], rtol=0, atol=1e-15)):
        raise ValueError('the FFT field must be taken away from copper planes')
    if not isinstance(refine, (int, np.integer)) or refine < 1:
        raise ValueError('refine must be a positive integer')
    multi = any(np.ndim(sh[2]) > 1 for sh in sheets) or (segments is not None and np.ndim(segments[2]) > 1)
    ks = [np.asarray(sh[2]).reshape(len(sh[1]), -1).shape[1] for sh in sheets]
    if segments is not None:
        Iseg = np.asarray(segments[2])
        ks.append(Iseg.reshape(len(Iseg), -1).shape[1] if len(Iseg) else 1)
    k = max(ks) if ks else 1
    out = np.zeros((len(x), len(y), len(heights), 3, k), complex)
    if sheets:
        step = (x[1] - x[0]) / refine, (y[1] - y[0]) / refine
        allnodes = np.concatenate([np.asarray(sh[0], float)[np.unique(np.asarray(sh[1]))][:, :2] for sh in sheets])
        grids, offsets = [], []
        for dim, axis, d in ((0, x, step[0]), (1, y, step[1])):
            lo = min(0, int(np.floor((allnodes[:, dim].min() - axis[0]) / d)) - 2)
            hi = max(refine * (len(axis) - 1), int(np.ceil((allnodes[:, dim].max() - axis[0]) / d)) + 2)
            grids.append(axis[0] + np.arange(lo, hi + 1) * d)
            offsets.append(-lo)
        gx, gy = grids
        sx, sy = len(gx), len(gy)
        fx, fy = next_fast_len(3 * sx - 2), next_fast_len(3 * sy - 2)
        J = np.ascontiguousarray(np.stack([current_deposit(sh, gx, gy, k, per_cell) for sh in sheets]))
        H = np.zeros((len(x), len(y), len(heights), 3, k), complex)
        ms = ctypes.c_double()
        call('pq_biot_grid', device, len(sheets), ptr(np.ascontiguousarray(zs)), k, sx, sy, ptr(J), step[0], step[1],
             len(heights), ptr(np.ascontiguousarray(heights)), fx, fy, sx - 1 + offsets[0], sy - 1 + offsets[1],
             refine, len(x), len(y), ptr(H), ctypes.byref(ms))
        out += H
        if timing is not None:
            timing.update(kernel_ms=ms.value, source_grid=(sx, sy), fft=(fx, fy))
    if segments is not None and len(segments[0]):
        X, Y, Z = np.meshgrid(x, y, heights, indexing='ij')
        pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
        Hs = h_field([], pts, segments, device=device)
        out += np.asarray(Hs).reshape(len(x), len(y), len(heights), 3, -1)
    return out if multi else out[..., 0]


def main(argv=None):
    """Time the field of random RWG currents on the seeded plane pair at a grid of points."""
    import argparse
    import json
    import time
    from . import board as boards
    from .geometry import RWG
    ap = argparse.ArgumentParser(description='Biot-Savart timing on the seeded 2-layer plane pair')
    ap.add_argument('--grid', type=int, default=1000, help='points per side of the observation grid')
    ap.add_argument('--excitations', type=int, default=8)
    ap.add_argument('--precision', type=int, default=32)
    ap.add_argument('--height', type=float, default=2e-3)
    ap.add_argument('--seed', type=int, default=0)
    args = ap.parse_args(argv)
    b = boards.board_c2(args.seed)
    r = RWG(b.nodes, b.tris)
    rng = np.random.default_rng(args.seed)
    K = args.excitations
    I = rng.normal(size=(r.Ne, K)) + 1j * rng.normal(size=(r.Ne, K))
    a, bb = rwg_currents(r, I)
    sheets = [(b.nodes, b.tris[b.layer_of == lay], a[b.layer_of == lay], bb[b.layer_of == lay],
               float(b.z_of[b.layer_of == lay][0])) for lay in np.unique(b.layer_of)]
    lo, hi = b.nodes.min(0), b.nodes.max(0)
    X, Y = np.meshgrid(np.linspace(lo[0], hi[0], args.grid), np.linspace(lo[1], hi[1], args.grid))
    pts = np.c_[X.ravel(), Y.ravel(), np.full(X.size, args.height)]
    timing = {}
    t = time.perf_counter()
    h_field(sheets, pts, precision=args.precision, timing=timing)
    total = time.perf_counter() - t
    inter = timing['sources'] * timing['points']
    kk = 1 if K <= 1 else 2 if K <= 2 else 4 if K <= 4 else 8
    print(json.dumps(dict(points=timing['points'], sources=timing['sources'], excitations=K, precision=args.precision,
                          kernel_ms=timing['kernel_ms'], total_s=total,
                          interactions_per_s=inter / (timing['kernel_ms'] * 1e-3),
                          tflops=inter * (20 + 16 * kk) / (timing['kernel_ms'] * 1e-3) / 1e12)))


if __name__ == '__main__':
    main()
