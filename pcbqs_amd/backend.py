"""The `mi355x` fill backend: the pcbqs.static_backends plugin object.

Protocol (src/pcbqs/static_backends.py:1-10, docs/static-backends.md:157-193): a factory
make_backend(geometry, workers, chunk, device) returns an object with
  parallel_tiles = False               device backend: evaluate() is never called concurrently
  evaluate(p, q, d, kind, scalar_only) host float64 (n, 6) / (n, 1) pair moments
  fill_symmetric(matrices, near, outer_levels, scalar_only) -> analytic pair count
and the optional attributes device_id and min_batch_pairs. The geometry (a FillGeometry
from the reference's FillEngine or from pcbqs_amd.geometry) is uploaded once, here.

Beyond the protocol, fill_device() keeps the moment planes in HBM (FP64 reference layout,
or FP32 local-frame planes) for the device-resident solve pipeline.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr


class Mi355xBackend:
    parallel_tiles = False
    # Explicit pair-list callers batch at least this many pairs per call (the same role as
    # CupyBackend.min_batch_pairs, static_cupy.py:69-71): a device call costs a launch and a sync.
    min_batch_pairs = 1 << 20

    def __init__(self, geometry, workers=1, chunk=16384, device=None):
        if device is not None and (isinstance(device, bool) or not isinstance(device, (int, np.integer))):
            raise ValueError('device must be an integer ordinal or None')
        try:
            count = _lib.device_count()
        except RuntimeError as exc:
            raise RuntimeError(f'MI355X fill unavailable: {exc}') from exc
        if count < 1:
            raise RuntimeError('MI355X fill unavailable: no HIP device is visible')
        self.device_id = 0 if device is None else int(device)
        self.chunk = chunk
        self.geometry = geometry
        self.nt = len(geometry.tris)
        levels = geometry.levels
        V, L, tau, un, area, _, _ = levels[0]
        keep = []

        def c(a):
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return ptr(a).value
        g = _lib.Geometry()
        g.nt = self.nt
        g.V, g.L, g.tau, g.un, g.area, g.P3 = c(V), c(L), c(tau), c(un), c(area), c(geometry.points)
        for k in range(3):
            lev = levels.get(k, levels[0])
            g.qp[k], g.qw[k] = c(lev[5]), c(lev[6])
            g.nq[k] = len(lev[6])
        g.z, g.cent, g.radius = c(geometry.z), c(geometry.cent), c(geometry.radius)
        handle = ctypes.c_void_p()
        try:
            call('pq_fill_create', ctypes.byref(g), self.device_id, ctypes.byref(handle))
        except (RuntimeError, MemoryError) as exc:
            raise RuntimeError(f'MI355X fill could not initialize device {device}: {exc}') from exc
        self._handle = handle
        self.routing = dict(device_batches=0, device_pairs=0)

    # -- the reference protocol ---------------------------------------------------------------
    def evaluate(self, p, q, d, kind, scalar_only):
        p, q = np.ascontiguousarray(p, np.int32), np.ascontiguousarray(q, np.int32)
        d, kind = np.ascontiguousarray(d, np.float64), np.ascontiguousarray(kind, np.int32)
        n = len(p)
        out = np.empty((n, 1 if scalar_only else 6))
        self.routing['device_batches'] += 1
        self.routing['device_pairs'] += n
        if n:
            call('pq_pairs', self._handle, ptr(p), ptr(q), ptr(d), ptr(kind), n, int(bool(scalar_only)), ptr(out))
        return out

    def fill_symmetric(self, matrices, near, outer_levels, scalar_only):
        n = self.nt
        width = 1 if scalar_only else 4
        if len(matrices) != width:
            raise ValueError('fill_symmetric: 1 (scalar_only) or 4 output matrices')
        for m in matrices:
            if (not isinstance(m, np.ndarray) or m.dtype != np.float64 or m.shape != (n, n)
                    or not m.flags.c_contiguous or not m.flags.writeable):
                raise ValueError('fill_symmetric: outputs must be writable C-contiguous float64 n x n')
        planes = (ctypes.c_void_p * 4)(*[m.ctypes.data for m in matrices], *([None] * (4 - width)))
        analytic = ctypes.c_int64()
        call('pq_fill_symmetric', self._handle, float(near), int(outer_levels), int(bool(scalar_only)), planes,
             ctypes.byref(analytic))
        return int(analytic.value)

    # -- device-resident fill -----------------------------------------------------------------
    def fill_device(self, near=3.0, outer_levels=1, scalar_only=False, precision=64, out=None, stream=None):
        """Fill the moment planes into HBM: a DeviceArray (1 or 4, nt, ld) of float64 (reference
        layout) or float32 (local frame), ld = nt rounded up to 64 (whole cache lines per row
        segment); columns nt..ld-1 are padding. Returns (planes, analytic pairs)."""
        n = self.nt
        width = 1 if scalar_only else 4
        dtype = np.float64 if precision == 64 else np.float32
        ld = padded(n)
        if out is None:
            out = DeviceArray((width, n, ld), dtype, self.device_id)
        elif out.shape != (width, n, ld) or out.dtype != dtype:
            raise ValueError('fill_device: output planes of the wrong shape or dtype')
        analytic = ctypes.c_int64()
        call('pq_fill_device', self._handle, float(near), int(outer_levels), int(bool(scalar_only)),
             int(precision), out.ptr, ld, ctypes.byref(analytic), stream)
        return out, int(analytic.value)

    def last_timing(self):
        """(far kernel ms, near kernels ms, near pairs per list) of the last fill_device()."""
        far, near = ctypes.c_double(), ctypes.c_double()
        pairs = (ctypes.c_int64 * 2)()
        call('pq_fill_last_timing', self._handle, ctypes.byref(far), ctypes.byref(near), pairs)
        return far.value, near.value, (pairs[0], pairs[1])

    def close(self):
        if getattr(self, '_handle', None) is not None and _lib._lib is not None:
            _lib._lib.pq_fill_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass


def padded(n, multiple=64):
    """Leading dimension of device matrices: n rounded up to a whole number of 64-element rows."""
    return (int(n) + multiple - 1) // multiple * multiple


def make_backend(geometry, workers, chunk, device):
    """Entry point `mi355x` of the pcbqs.static_backends group (pyproject.toml)."""
    return Mi355xBackend(geometry, workers, chunk, device)


def combine_device(planes, precision, rwg, cent, factor=1.0, scale=None, out_precision=64, device=0,
                   stream=None):
    """L (Ne, ldl) on the device (ldl = Ne rounded up to 64; padding columns unused) from device
    moment planes (1 or 4, nt, ld): static_combine's formula."""
    nt, ld = planes.shape[-2], planes.shape[-1]
    tri, alpha, vertex = rwg.combine_arrays()
    ne = tri.shape[1]
    scale = np.ones(ne) if scale is None else np.ascontiguousarray(scale, float)
    dt = DeviceArray.from_host(tri, device)
    da = DeviceArray.from_host(alpha, device)
    dv = DeviceArray.from_host(vertex, device)
    ds = DeviceArray.from_host(scale, device)
    dc = DeviceArray.from_host(np.ascontiguousarray(cent, np.float64), device)
    ldl = padded(ne)
    L = DeviceArray((ne, ldl), np.float64 if out_precision == 64 else np.float32, device)
    call('pq_combine_device', int(precision), planes.ptr, nt, ld, dc.ptr, dt.ptr, da.ptr, dv.ptr, ds.ptr, ne,
         float(factor), int(out_precision), L.ptr, ldl, stream)
    call('pq_synchronize', device)
    for a in (dt, da, dv, ds, dc):
        a.close()
    return L
