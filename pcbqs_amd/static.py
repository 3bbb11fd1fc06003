"""Host entry points around the `mi355x` fill backend.

  static_moments    the whole-matrix fill (the role of src/pcbqs/static.py:11-95) through the
                    plugin object's fill_symmetric: host FP64 planes, device tiling
  edge_inductance   src/pcbqs/static.py:103-128 (static_combine's formula, on the device)
  potential         src/pcbqs/mqs.py:532-570 (P on the scalar fill)
  capacitance_matrix, board_capacitance   src/pcbqs/peec.py:95-123

Backend selection by name (`--fill-backend`, PCBQS_FILL_BACKEND, the pcbqs.static_backends
entry-point group) is the reference's FillEngine's job; this package only provides the plugin
(`pcbqs_amd.backend:make_backend`, registered in pyproject.toml). There is no CPU fallback:
without a device the backend raises RuntimeError.
"""
import time

import numpy as np

from ._lib import DeviceArray, call, ptr
from .backend import Mi355xBackend, combine_device, padded
from .geometry import EPS0, FillGeometry


def _check_rule(near, outer_levels):
    if np.ndim(near) != 0 or not near >= 0:
        raise ValueError('near must be a nonnegative number or infinity')
    if isinstance(outer_levels, bool) or not isinstance(outer_levels, (int, np.integer)) or \
            int(outer_levels) not in (0, 1, 2):
        raise ValueError('outer_levels must be 0, 1 or 2')
    return float(near), int(outer_levels)


def static_moments(nodes, tris, z, near=3.0, outer_levels=1, scalar_only=False, device=None, return_info=False):
    """(local nodes, M0, Mrx, Mry, Mrr), or (local nodes, M0) when scalar_only: every unordered
    triangle pair's moments in the reference's layout (SI units, centred frame), filled on the
    device by the plugin object's fill_symmetric and copied into host FP64 (n, n) matrices.
    info: pairs n(n+1)/2, analytic pairs, seconds."""
    near, outer_levels = _check_rule(near, outer_levels)
    t0 = time.perf_counter()
    g = FillGeometry.prepare(nodes, tris, z, outer_levels)
    backend = Mi355xBackend(g, 1, 16384, device)
    try:
        n = len(g.tris)
        planes = [np.empty((n, n)) for _ in range(1 if scalar_only else 4)]
        analytic = backend.fill_symmetric(planes, near, outer_levels, bool(scalar_only))
    finally:
        backend.close()
    info = dict(backend='mi355x', device=backend.device_id, precision='float64', triangles=n,
                outer_levels=outer_levels, near='all' if np.isinf(near) else near, scalar_only=bool(scalar_only),
                pairs=n * (n + 1) // 2, analytic_pairs=int(analytic), seconds=time.perf_counter() - t0)
    out = (g.nodes, *planes)
    return (out, info) if return_info else out


def edge_inductance(rwg, Mrr, Mrx, Mry, M0, factor=1.0, scale=None, workers=None, device=0):
    """factor * diag(scale) combine(rwg, Mrr, Mrx, Mry, M0) diag(scale), symmetric: (L, info).

    The host moment matrices are uploaded and combined on the device (static_combine's formula)."""
    started = time.perf_counter()
    nt = M0.shape[0]
    planes = DeviceArray((4, nt, nt), np.float64, device)
    for k, m in enumerate((M0, Mrx, Mry, Mrr)):
        m = np.ascontiguousarray(m, np.float64)
        if m.shape != (nt, nt):
            raise ValueError('static_combine: moments must be C-contiguous float64 n x n')
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
