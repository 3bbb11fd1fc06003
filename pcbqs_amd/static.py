"""The fill interface of pcbqs, with `mi355x` as its device backend.

Mirrors, for users switching from the reference (same names, arguments, results, errors):
  FillEngine, register_fill_backend, resolve_workers  src/pcbqs/static_backends.py:57-67,234-336
  static_moments                                      src/pcbqs/static.py:11-95
  edge_inductance                                     src/pcbqs/static.py:103-128
  potential                                           src/pcbqs/mqs.py:532-570 (_set_potential
                                                      on the scalar fill, _ensure_potential)
The only built-in backend is 'mi355x' ('auto' selects it; PCBQS_FILL_BACKEND is honoured);
other names come from register_fill_backend or the pcbqs.static_backends entry points, as in
the reference. There is no CPU fallback: without a device, the backend raises RuntimeError.
"""
from importlib import metadata
import os
import time

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr
from .backend import Mi355xBackend, combine_device, padded
from .geometry import EPS0, FillGeometry

_FACTORIES = {'mi355x': Mi355xBackend}


def _positive_integer(value, name):
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)) or value < 1:
        raise ValueError(f'{name} must be a positive integer')
    return int(value)


def resolve_workers(workers=None):
    """Explicit count, else PCBQS_FILL_THREADS, PCBQS_OMP_THREADS, OMP_NUM_THREADS, else 1."""
    if workers is None:
        key = next((k for k in ('PCBQS_FILL_THREADS', 'PCBQS_OMP_THREADS', 'OMP_NUM_THREADS')
                    if k in os.environ), None)
        env = os.environ.get(key) if key else None
        try:
            workers = int(env.split(',')[0] if key == 'OMP_NUM_THREADS' else env) if env is not None else 1
        except ValueError as exc:
            raise ValueError(f'{key} must be a positive integer') from exc
    return _positive_integer(workers, 'workers')


def register_fill_backend(name, factory):
    if not isinstance(name, str) or not name or name == 'auto' or name in _FACTORIES:
        raise ValueError(f'invalid or already registered fill backend {name!r}')
    if not callable(factory):
        raise TypeError('fill backend factory must be callable')
    _FACTORIES[name] = factory


class FillEngine:
    def __init__(self, nodes, tris, z, near=3., outer_levels=1, backend='auto', workers=None,
                 chunk=16384, device=None):
        if np.ndim(near) != 0 or near < 0 or np.isnan(near):
            raise ValueError('near must be nonnegative (or infinity)')
        if not isinstance(outer_levels, (int, np.integer)) or outer_levels not in (0, 1, 2):
            raise ValueError('outer_levels must be 0, 1 or 2')
        requested = backend or os.environ.get('PCBQS_FILL_BACKEND', 'auto')
        if requested == 'auto':
            requested = os.environ.get('PCBQS_FILL_BACKEND', 'auto')
        name = 'mi355x' if requested == 'auto' else requested
        workers = resolve_workers(workers)
        self.chunk = _positive_integer(chunk, 'chunk')
        self.near, self.outer_levels = near, outer_levels
        if name not in _FACTORIES:
            providers = [ep for ep in metadata.entry_points(group='pcbqs.static_backends') if ep.name == name]
            if len(providers) != 1:
                raise ValueError(f'unknown or ambiguous fill backend {name!r}')
            register_fill_backend(name, providers[0].load())
        self.geometry = FillGeometry.prepare(nodes, tris, z, outer_levels)
        self.backend = _FACTORIES[name](self.geometry, workers, self.chunk, device)
        self.info = dict(backend=name, workers=workers, chunk=self.chunk,
                         device=getattr(self.backend, 'device_id', device), precision='float64',
                         triangles=len(self.geometry.tris), outer_levels=int(outer_levels),
                         near='all' if np.isinf(near) else float(near))
        if hasattr(self.backend, 'routing'):
            self.info['routing'] = self.backend.routing

    @property
    def pair_batch(self):
        return max(self.chunk, getattr(self.backend, 'min_batch_pairs', 1))

    def kinds(self, p, q, d):
        g = self.geometry
        distance = np.sqrt(((g.cent[p] - g.cent[q]) ** 2).sum(-1) + d * d)
        size = g.radius[p] + g.radius[q]
        refine = distance <= size
        if np.isinf(self.near):
            return np.where(refine, np.int32(self.outer_levels), np.int32(0)).astype(np.int32)
        close = distance <= self.near * size
        kind = np.where(close, np.int32(0), np.int32(-1)).astype(np.int32)
        kind[close & refine] = self.outer_levels
        return kind

    def evaluate(self, p, q, d, kind=None, scalar_only=False):
        p, q = np.asarray(p), np.asarray(q)
        n = len(self.geometry.tris)
        if (p.ndim != 1 or q.shape != p.shape or p.dtype.kind not in 'iu' or q.dtype.kind not in 'iu' or
                (len(p) and (p.min() < 0 or q.min() < 0 or p.max() >= n or q.max() >= n))):
            raise ValueError('pair indices must be matching integer vectors within the mesh')
        p, q = np.ascontiguousarray(p, np.int32), np.ascontiguousarray(q, np.int32)
        d = np.asarray(d, float)
        if d.ndim == 0:
            if not np.isfinite(d):
                raise ValueError('pair offsets must be finite')
            d = np.full(p.shape, float(d))
        else:
            d = np.ascontiguousarray(np.broadcast_to(d, p.shape))
            if not np.isfinite(d).all():
                raise ValueError('pair offsets must be finite')
        if kind is None:
            kind = self.kinds(p, q, d)
        else:
            kind = np.asarray(kind)
            if kind.shape != p.shape or kind.dtype.kind not in 'iu':
                raise ValueError('pair kinds must select far (-1) or a prepared analytic level')
            if len(kind):
                lo, hi = int(kind.min()), int(kind.max())
                levels = self.geometry.levels
                if (lo < -1 or hi > max(levels) or
                        any(level not in levels and (kind == level).any() for level in range(max(0, lo), hi + 1))):
                    raise ValueError('pair kinds must select far (-1) or a prepared analytic level')
        kind = np.ascontiguousarray(kind, np.int32)
        out = self.backend.evaluate(p, q, d, kind, scalar_only)
        if not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.shape != (len(p), 1 if scalar_only else 6):
            raise TypeError('fill backend must return a host float64 array of pair moments')
        if not np.isfinite(out).all():
            raise FloatingPointError('non-finite triangle-pair moments from fill backend')
        return out


def static_moments(nodes, tris, z, near=3.0, outer_levels=1, chunk=16384, scalar_only=False,
                   backend='auto', workers=None, device=None, return_info=False):
    """(local nodes, M0, Mrx, Mry, Mrr) or (local nodes, M0), SI units, as the reference's."""
    started = time.perf_counter()
    engine = FillEngine(nodes, tris, z, near, outer_levels, backend, workers, chunk, device)
    g = engine.geometry
    n = len(g.tris)
    matrices = [np.empty((n, n)) for _ in range(1 if scalar_only else 4)]
    if hasattr(engine.backend, 'fill_symmetric'):
        analytic = engine.backend.fill_symmetric(matrices, engine.near, engine.outer_levels, scalar_only)
        scheduler = 'device'
        count = n * (n + 1) // 2
    else:
        # host tiles through evaluate() for registered pair-only backends (static.py:52-91)
        side = max(1, int(np.sqrt(engine.chunk)))
        count = analytic = 0
        for p0 in range(0, n, side):
            for q0 in range(p0, n, side):
                p1, q1 = min(n, p0 + side), min(n, q0 + side)
                p, q = np.meshgrid(np.arange(p0, p1, dtype=np.int32), np.arange(q0, q1, dtype=np.int32),
                                   indexing='ij')
                p, q = p.ravel(), q.ravel()
                if p0 == q0:
                    upper = p <= q
                    p, q = p[upper], q[upper]
                d = g.z[p] - g.z[q]
                kind = engine.kinds(p, q, d)
                m = engine.evaluate(p, q, d, kind, scalar_only)
                count += len(p)
                analytic += int(np.count_nonzero(kind >= 0))
                matrices[0][p, q] = matrices[0][q, p] = m[:, 0]
                if not scalar_only:
                    mx, my, rr = matrices[1:]
                    mx[p, q], my[p, q] = m[:, 1], m[:, 2]
                    mx[q, p], my[q, p] = m[:, 3], m[:, 4]
                    rr[p, q] = rr[q, p] = m[:, 5]
                    diagonal = p == q
                    mx[p[diagonal], q[diagonal]] = (m[diagonal, 1] + m[diagonal, 3]) / 2
                    my[p[diagonal], q[diagonal]] = (m[diagonal, 2] + m[diagonal, 4]) / 2
        scheduler = 'host'
    info = dict(engine.info, scheduler=scheduler, pairs=int(count), analytic_pairs=int(analytic),
                scalar_only=bool(scalar_only), seconds=time.perf_counter() - started)
    result = (g.nodes, *matrices)
    return (result, info) if return_info else result


def edge_inductance(rwg, Mrr, Mrx, Mry, M0, factor=1.0, scale=None, workers=None, device=0):
    """factor * diag(scale) combine(rwg, Mrr, Mrx, Mry, M0) diag(scale), symmetric: (L, info).

    The host moment matrices are uploaded and combined on the device (static_combine's formula)."""
    from ._lib import DeviceArray
    started = time.perf_counter()
    nt = M0.shape[0]
    planes = DeviceArray((4, nt, nt), np.float64, device)
    for k, m in enumerate((M0, Mrx, Mry, Mrr)):
        m = np.ascontiguousarray(m, np.float64)
        if m.shape != (nt, nt):
            raise ValueError('static_combine: moments must be C-contiguous float64 n x n')
        from ._lib import call, ptr
        call('pq_memcpy_h2d', planes.offset(k * nt * nt), ptr(m), m.nbytes)
    cent = rwg.nodes[rwg.tris].mean(1)
    L = combine_device(planes, 64, rwg, cent, factor, scale, 64, device)
    out = np.ascontiguousarray(L.host()[:, :L.shape[0]])
    L.close()
    planes.close()
    return out, dict(backend='mi355x', workers=1, seconds=time.perf_counter() - started)


def potential(nodes, tris, z, node_of, er=1.0, near=3.0, outer_levels=1, precision=64, device=0, keep=False):
    """Potential coefficients P (n_nodes, n_nodes) FP64 host of the charge nodes node_of (nt,):
    P = W^T M0 W / (4 pi eps0 er), W[t, node(t)] = 1 / area(node) (src/pcbqs/mqs.py:532-556; the
    reference's side-charge panels excluded). M0 is the device's scalar fill (precision 64: the
    reference layout; 32: FP32 far pairs, FP64 near pairs), which stays in HBM. keep=True returns
    the device array (the caller closes it) instead of the host copy."""
    nodes, tris = np.asarray(nodes, float), np.asarray(tris)
    node_of = np.asarray(node_of, np.int64)
    if node_of.shape != (len(tris),) or node_of.min() < 0:
        raise ValueError('node_of must give one nonnegative node per triangle')
    g = FillGeometry.prepare(nodes, tris, z, outer_levels)
    backend = Mi355xBackend(g, 1, 16384, device)
    try:
        planes, _ = backend.fill_device(near, outer_levels, True, precision)
        n = int(node_of.max()) + 1
        area = np.bincount(node_of, weights=g.area, minlength=n)
        if (area <= 0).any():
            raise ValueError('every charge node needs at least one triangle')
        order = np.argsort(node_of, kind='stable').astype(np.int32)
        ptr_ = np.r_[0, np.cumsum(np.bincount(node_of, minlength=n))].astype(np.int32)
        d_ptr, d_tris, d_area = (DeviceArray.from_host(a, device) for a in (ptr_, order, area))
        P = DeviceArray((n, n), np.float64, device)
        call('pq_potential_device', precision, planes.ptr, len(tris), padded(len(tris)), n, d_ptr.ptr, d_tris.ptr,
             d_area.ptr, 1.0 / (4 * np.pi * EPS0 * er), 64, P.ptr, n, None)
        call('pq_synchronize', device)
        for a in (planes, d_ptr, d_tris, d_area):
            a.close()
        if keep:
            return P
        out = P.host()
        P.close()
        return out
    finally:
        backend.close()


def capacitance_matrix(nodes, tris, z, node_of, labels, er=1.0, near=3.0, outer_levels=1, precision=64, device=0,
                       tol=1e-12, max_steps=12, info=None):
    """Maxwell capacitance C (n_conductors, n_conductors) in F of the conductors labels (n_nodes,)
    over the charge nodes node_of (nt,): C = B^T P^-1 B with B[node, labels[node]] = 1 and P the
    potential matrix of `potential` (src/pcbqs/peec.py:95-123: potentials relative to infinity, so
    row sums need not vanish; symmetrised as the reference does). P stays on the device, where it
    is factored as L D L^T in FP32 and the solve is refined in FP64 to a relative residual of tol.
    info (a dict), when given, receives corrections, relative_residual and factor_ms."""
    import ctypes
    labels = np.asarray(labels, np.int64)
    n = int(np.asarray(node_of).max()) + 1
    if labels.shape != (n,) or labels.min() < 0:
        raise ValueError('labels must give one conductor per charge node')
    nc = int(labels.max()) + 1
    B = np.zeros((n, nc))
    B[np.arange(n), labels] = 1.0
    P = potential(nodes, tris, z, node_of, er, near, outer_levels, precision, device, keep=True)
    try:
        X = np.empty((n, nc))
        steps, res, fms = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        call('pq_spd_solve', device, n, P.ptr, n, nc, ptr(np.ascontiguousarray(B)), float(tol), int(max_steps),
             ptr(X), ctypes.byref(steps), ctypes.byref(res), ctypes.byref(fms))
    finally:
        P.close()
    if info is not None:
        info.update(corrections=int(steps.value), relative_residual=float(res.value), factor_ms=float(fms.value))
    C = B.T @ X
    return (C + C.T) / 2


def board_capacitance(board, er=1.0, precision=64, device=0, info=None):
    """(C, labels) for a board: the Maxwell capacitance matrix of its conductors, the connected
    components of its copper and vias (ports and lumped elements do not join them), labels the
    conductor of each charge node; src/pcbqs/peec.py:95-123."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from .loops import LoopModel
    model = LoopModel(board)
    node_of = np.asarray(board.node_of, np.int64)
    n = int(node_of.max()) + 1
    size = max(n, int(np.max(model.br)) + 1)
    a, b = np.asarray(model.br, np.int64).T
    _, comp = connected_components(coo_matrix((np.ones(len(a)), (a, b)), shape=(size, size)), directed=False)
    labels = np.unique(comp[:n], return_inverse=True)[1]
    C = capacitance_matrix(board.nodes, board.tris, board.z_of, node_of, labels, er, precision=precision,
                           device=device, info=info)
    return C, labels
