"""Oracle for the triangle-pair fill (test infrastructure only; see oracle/__init__.py).

Restates, in numpy:
  rules          src/pcbqs/_numerics/base.py:64-71 (_BAR3/_W3/_BAR6/_W6; Dunavant degree-4 6-point rule)
  edge_frames    src/pcbqs/_numerics/base.py:82-88
  split4         src/pcbqs/_numerics/currents.py:21-26
  level_geometry src/pcbqs/_numerics/sing.py:56-69 (_offplane_geometry)
  prepare        src/pcbqs/static_backends.py:105-130 (FillGeometry.prepare)
  kinds          src/pcbqs/static_backends.py:288-298 (FillEngine.kinds)
  far_pairs      src/pcbqs/static_backends.py:137-147 (_far)
  analytic_pairs src/pcbqs/_numerics/sing.py:5-53 (offplane_pairs)
and wraps libcpu_ref.so (oracle/cpu_ref.c) and, where built, the reference core in oracle/_ref.

Pinned by: bitwise equality of cpu_ref.c with the reference core (tests/test_oracle.py), brute-force
adaptive quadrature of the 1/R moments on tiny meshes, the closed-form Wilton-Rao inner integral
against a brute-force 2-D integral, and the copper-reference.npz chain (port Z at 1e-11).
"""
import ctypes
import glob
import importlib.util
import os
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_a1, _a2 = 0.445948490915965, 0.091576213509771
BAR6 = np.array([[1 - 2 * _a1, _a1, _a1], [_a1, 1 - 2 * _a1, _a1], [_a1, _a1, 1 - 2 * _a1],
                 [1 - 2 * _a2, _a2, _a2], [_a2, 1 - 2 * _a2, _a2], [_a2, _a2, 1 - 2 * _a2]])
W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
BAR3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
W3 = np.ones(3) / 3


def edge_frames(V):
    e = np.roll(V, -1, axis=1) - V
    L = np.linalg.norm(e, axis=2)
    tau = e / L[..., None]
    return e, L, tau, np.stack([tau[..., 1], -tau[..., 0]], -1)


def split4(V):
    """(M,3,c) -> (4M,3,c): corner triangles (v0,m01,m20), (m01,v1,m12), (m20,m12,v2), then the middle one."""
    m01, m12, m20 = (V[:, 0] + V[:, 1]) / 2, (V[:, 1] + V[:, 2]) / 2, (V[:, 2] + V[:, 0]) / 2
    return np.concatenate([np.stack(t, 1) for t in ((V[:, 0], m01, m20), (m01, V[:, 1], m12),
                                                      (m20, m12, V[:, 2]), (m01, m12, m20))])


def level_geometry(nodes, tris, level):
    V = nodes[tris]
    e, L, tau, un = edge_frames(V)
    area = 0.5 * np.abs(e[:, 0, 0] * (-e[:, 2, 1]) - e[:, 0, 1] * (-e[:, 2, 0]))
    bar, weights = BAR6, W6
    if level:
        ref = np.eye(3)[None]
        for _ in range(level):
            ref = split4(ref)
        bar = np.einsum('qk,tkc->tqc', BAR6, ref).reshape(-1, 3)
        weights = np.tile(W6, len(ref)) / len(ref)
    P = np.einsum('qk,tkc->tqc', bar, V)
    return V, L, tau, un, area, P, np.array(weights, copy=True)


def prepare(nodes, tris, z, outer_levels=1):
    """FillGeometry.prepare: validated, centred geometry (a namespace of read-only arrays)."""
    nodes, tris, z = np.asarray(nodes, float), np.asarray(tris), np.asarray(z, float)
    if (nodes.ndim != 2 or nodes.shape[1] != 2 or not len(nodes) or tris.ndim != 2 or tris.shape[1] != 3
            or not len(tris) or tris.dtype.kind not in 'iu' or z.shape != (len(tris),)
            or not np.isfinite(nodes).all() or not np.isfinite(z).all()
            or tris.min() < 0 or tris.max() >= len(nodes)):
        raise ValueError('fill requires finite planar nodes, integer triangles and one z per triangle')
    nodes = np.ascontiguousarray(nodes - nodes.mean(0))
    tris = np.array(tris, dtype=np.int32, order='C', copy=True)
    z = np.array(z, order='C', copy=True)
    V = nodes[tris]
    edge = V[:, 1:] - V[:, :1]
    if (edge[:, 0, 0] * edge[:, 1, 1] - edge[:, 0, 1] * edge[:, 1, 0] <= 0).any():
        raise ValueError('fill triangles must be nondegenerate and counter-clockwise')
    cent = V.mean(1)
    radius = np.linalg.norm(V - cent[:, None], axis=2).max(1)
    points = np.einsum('qk,tkc->tqc', BAR3, V)
    levels = {level: level_geometry(nodes, tris, level) for level in {0, outer_levels}}
    return SimpleNamespace(nodes=nodes, tris=tris, z=z, cent=cent, radius=radius, points=points,
                           levels=levels, outer_levels=outer_levels, area=levels[0][4])


def kinds(g, p, q, d, near=3.0, outer_levels=1):
    distance = np.sqrt(((g.cent[p] - g.cent[q]) ** 2).sum(-1) + d * d)
    size = g.radius[p] + g.radius[q]
    refine = distance <= size
    if np.isinf(near):
        return np.where(refine, np.int32(outer_levels), np.int32(0)).astype(np.int32)
    close = distance <= near * size
    kind = np.where(close, np.int32(0), np.int32(-1)).astype(np.int32)
    kind[close & refine] = outer_levels
    return kind


def far_pairs(g, p, q, d, scalar_only=False):
    pa, pb = g.points[p], g.points[q]
    a, b = pa[:, :, None], pb[:, None, :]
    w = (g.area[p] * g.area[q] / 9)[:, None, None] / np.sqrt(((a - b) ** 2).sum(-1) + d[:, None, None] ** 2)
    m0 = w.sum((1, 2))
    if scalar_only:
        return m0[:, None]
    ro = np.einsum('mq,mqc->mc', w.sum(2), pa)
    rs = np.einsum('mq,mqc->mc', w.sum(1), pb)
    rr = np.einsum('mab,mac,mbc->m', w, pa, pb)
    return np.column_stack((m0, ro, rs, rr))


def analytic_pairs(g, p, q, d, level, scalar_only=False):
    """Closed-form inner integral of 1/R and r'/R over the source triangle q (Wilton, Rao et al.,
    IEEE TAP 1984), summed at the observation points of p for the given level."""
    V, L, tau, un, area, P, weights = g.levels[level]
    dz = np.asarray(d, float)[:, None, None]
    ad = np.abs(dz)
    r = P[p]
    dv = V[q][:, None] - r[:, :, None]
    sm = (dv * tau[q][:, None]).sum(-1)
    sp = sm + L[q][:, None]
    t0 = (dv * un[q][:, None]).sum(-1)
    r02 = t0 ** 2 + dz * dz
    r0 = np.sqrt(r02)
    rm, rp = np.sqrt(r02 + sm ** 2), np.sqrt(r02 + sp ** 2)
    ok = r0 > 1e-14 * L[q][:, None]
    safe = np.where(ok, r0, 1.0)
    lg = np.where(ok, np.arcsinh(sp / safe) - np.arcsinh(sm / safe), 0.0)
    beta = np.arctan2(t0 * sp, r02 + ad * rp) - np.arctan2(t0 * sm, r02 + ad * rm)
    inner = (t0 * lg - ad * beta).sum(-1)
    w = weights[None] * area[p][:, None]
    m0 = (w * inner).sum(1)
    if scalar_only:
        return m0[:, None]
    er = 0.5 * (r02 * lg + sp * rp - sm * rm)
    ir = (er[..., None] * un[q][:, None]).sum(-2)
    src = ir + r * inner[..., None]
    ro = (w[..., None] * r * inner[..., None]).sum(1)
    rs = (w[..., None] * src).sum(1)
    rr = (w * (r * src).sum(-1)).sum(1)
    return np.column_stack((m0, ro, rs, rr))


def numpy_pairs(g, p, q, d, kind, scalar_only=False):
    """The serial numpy backend (NumpyBackend.evaluate, static_backends.py:158-169)."""
    out = np.empty((len(p), 1 if scalar_only else 6))
    for level in np.unique(kind):
        mask = kind == level
        if level == -1:
            out[mask] = far_pairs(g, p[mask], q[mask], d[mask], scalar_only)
        else:
            out[mask] = analytic_pairs(g, p[mask], q[mask], d[mask], int(level), scalar_only)
    return out


# ---- libcpu_ref.so (oracle/cpu_ref.c) -----------------------------------------------------------

class _Geometry(ctypes.Structure):
    _fields_ = [('nt', ctypes.c_int32), ('V', ctypes.c_void_p), ('L', ctypes.c_void_p),
                ('tau', ctypes.c_void_p), ('un', ctypes.c_void_p), ('area', ctypes.c_void_p),
                ('P3', ctypes.c_void_p), ('qp', ctypes.c_void_p * 3), ('qw', ctypes.c_void_p * 3),
                ('nq', ctypes.c_int32 * 3), ('z', ctypes.c_void_p), ('cent', ctypes.c_void_p),
                ('radius', ctypes.c_void_p)]


_LIB = None


def cpu_ref():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, 'libcpu_ref.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/libcpu_ref.so is not built: run make -C oracle')
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.cr_fill.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P]
        lib.cr_pairs.argtypes = [P, P, P, P, P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, P]
        lib.cr_combine.argtypes = [P, P, P, P, ctypes.c_int64, P, P, P, P, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_int, P]
        lib.cr_project.argtypes = [P, ctypes.c_int64, ctypes.c_int64, P, P, P, ctypes.c_int64, ctypes.c_int, P, P]
        lib.cr_max_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c_geometry(g):
    keep = []

    def c(a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype)
        keep.append(a)
        return _ptr(a)
    s = _Geometry()
    V, L, tau, un, area, _, _ = g.levels[0]
    s.nt = len(g.tris)
    s.V, s.L, s.tau, s.un, s.area, s.P3 = c(V), c(L), c(tau), c(un), c(area), c(g.points)
    for k in range(3):
        lev = g.levels.get(k, g.levels[0])
        s.qp[k], s.qw[k] = c(lev[5]).value, c(lev[6]).value
        s.nq[k] = len(lev[6])
    s.z, s.cent, s.radius = c(g.z), c(g.cent), c(g.radius)
    return s, keep


def cr_fill(g, near=3.0, outer_levels=1, scalar_only=False, workers=None):
    """(M0, Mrx, Mry, Mrr) (or (M0,)) float64 (nt, nt), pairs, analytic."""
    lib = cpu_ref()
    workers = workers or lib.cr_max_threads()
    n = len(g.tris)
    mats = [np.empty((n, n)) for _ in range(1 if scalar_only else 4)]
    arr = (ctypes.c_void_p * 4)(*[m.ctypes.data for m in mats], *([None] * (4 - len(mats))))
    s, keep = _c_geometry(g)
    pairs, analytic = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.cr_fill(ctypes.byref(s), float(near), int(outer_levels), int(scalar_only), int(workers), arr,
                     ctypes.byref(pairs), ctypes.byref(analytic))
    if rc < 0:
        raise ValueError('cr_fill: invalid input')
    if rc == 1:
        raise FloatingPointError('non-finite triangle-pair moments from fill backend')
    return mats, pairs.value, analytic.value


def cr_pairs(g, p, q, d, kind, scalar_only=False, workers=None):
    lib = cpu_ref()
    workers = workers or lib.cr_max_threads()
    p, q = np.ascontiguousarray(p, np.int32), np.ascontiguousarray(q, np.int32)
    d, kind = np.ascontiguousarray(d, np.float64), np.ascontiguousarray(kind, np.int32)
    out = np.empty((len(p), 1 if scalar_only else 6))
    s, keep = _c_geometry(g)
    if lib.cr_pairs(ctypes.byref(s), _ptr(p), _ptr(q), _ptr(d), _ptr(kind), len(p), int(scalar_only),
                    int(workers), _ptr(out)):
        raise ValueError('cr_pairs: invalid triangle index, kind or offset')
    return out


def combine_arrays(rwg):
    """(tri (2,ne) int32, alpha (2,ne), vertex (2,ne,2)) as static.edge_inductance builds them (static.py:116-118)."""
    tri = np.ascontiguousarray(np.stack([rwg.tp, rwg.tm]), np.int32)
    alpha = np.stack([rwg.l / (2 * rwg.area[rwg.tp]), -rwg.l / (2 * rwg.area[rwg.tm])])
    vertex = np.ascontiguousarray(np.stack([rwg.vp, rwg.vm]), float)
    return tri, np.ascontiguousarray(alpha), vertex


def cr_combine(rwg, Mrr, Mrx, Mry, M0, factor=1.0, scale=None, workers=None):
    lib = cpu_ref()
    workers = workers or lib.cr_max_threads()
    tri, alpha, vertex = combine_arrays(rwg)
    ne = tri.shape[1]
    scale = np.ones(ne) if scale is None else np.ascontiguousarray(scale, float)
    out = np.empty((ne, ne))
    mats = [np.ascontiguousarray(m, np.float64) for m in (Mrr, Mrx, Mry, M0)]
    if lib.cr_combine(*[_ptr(m) for m in mats], M0.shape[0], _ptr(tri), _ptr(alpha), _ptr(vertex), _ptr(scale),
                      ne, float(factor), int(workers), _ptr(out)):
        raise ValueError('cr_combine: invalid triangle index')
    return out


def cr_project(C, L, workers=None):
    """C^T L C (dense, nl x nl) for a sparse C (n x nl) and a symmetric dense L (n x n)."""
    lib = cpu_ref()
    workers = workers or lib.cr_max_threads()
    C = C.tocsc()
    n, nl = C.shape
    L = np.ascontiguousarray(L, np.float64)
    if L.shape != (n, n):
        raise ValueError('cr_project: L must be n x n for a basis of n rows')
    colptr, rows = np.ascontiguousarray(C.indptr, np.int32), np.ascontiguousarray(C.indices, np.int32)
    sign = np.ascontiguousarray(C.data, np.float64)
    Ut, out = np.empty((nl, n)), np.empty((nl, nl))
    if lib.cr_project(_ptr(L), n, n, _ptr(colptr), _ptr(rows), _ptr(sign), nl, int(workers), _ptr(Ut), _ptr(out)):
        raise ValueError('cr_project: branch index out of range')
    return out


# ---- the reference's own compiled core (oracle/_ref), when it was built --------------------------

def ref_core():
    """The reference core module built by oracle/Makefile, or None where it is not built."""
    found = glob.glob(os.path.join(HERE, '_ref', '_core*.so'))
    if not found:
        return None
    spec = importlib.util.spec_from_file_location('_core', found[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def ref_geometry_dict(g):
    """The geometry dict the reference core's static_fill / static_pairs take (static_backends.py:190-194)."""
    V, L, tau, un, area, _, _ = g.levels[0]
    data = dict(V=V, L=L, tau=tau, un=un, area=area, P3=g.points)
    for level in range(3):
        lev = g.levels.get(level, g.levels[0])
        data[f'P{level}'], data[f'W{level}'] = lev[5], lev[6]
    return data


def ref_fill(core, g, near=3.0, outer_levels=1, scalar_only=False, workers=1):
    n = len(g.tris)
    mats = [np.empty((n, n)) for _ in range(1 if scalar_only else 4)]
    count, analytic = core.static_fill(ref_geometry_dict(g), g.cent, g.radius, g.z, float(near),
                                       int(outer_levels), bool(scalar_only), int(workers), mats)
    return mats, int(count), int(analytic)
