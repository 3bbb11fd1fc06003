"""ctypes binding of libpcbqs_hip.so (declared in include/pcbqs_amd.h).

Loading fails loudly: there is no CPU fallback anywhere in this package. A missing
library raises RuntimeError naming the build step; a HIP runtime with no device raises
RuntimeError from the first call that needs one.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, 'libpcbqs_hip.so')

PQ_OK, PQ_EINVAL, PQ_ENONFINITE, PQ_EDEVICE, PQ_ENOMEM, PQ_ECONVERGE = range(6)


class ConvergenceError(RuntimeError):
    """An iterative solve did not reach its tolerance (peec_iterative.ConvergenceError's role)."""

    def __init__(self, message, relative_residual=float('nan'), iterations=0):
        super().__init__(message)
        self.relative_residual, self.iterations = relative_residual, iterations


class Geometry(ctypes.Structure):
    """pq_geometry: pointers into the read-only float64 arrays of a prepared FillGeometry."""
    _fields_ = [('nt', ctypes.c_int32), ('V', ctypes.c_void_p), ('L', ctypes.c_void_p),
                ('tau', ctypes.c_void_p), ('un', ctypes.c_void_p), ('area', ctypes.c_void_p),
                ('P3', ctypes.c_void_p), ('qp', ctypes.c_void_p * 3), ('qw', ctypes.c_void_p * 3),
                ('nq', ctypes.c_int32 * 3), ('z', ctypes.c_void_p), ('cent', ctypes.c_void_p),
                ('radius', ctypes.c_void_p)]


class LoopTree(ctypes.Structure):
    """pq_loop_tree: the spanning tree of a fundamental loop basis (int32/float32 host arrays)."""
    _fields_ = [('n_nodes', ctypes.c_int32), ('parent', ctypes.c_void_p), ('up_branch', ctypes.c_void_p),
                ('up_sign', ctypes.c_void_p), ('n_levels', ctypes.c_int32), ('level_ptr', ctypes.c_void_p),
                ('level_nodes', ctypes.c_void_p), ('closing', ctypes.c_void_p), ('end_a', ctypes.c_void_p),
                ('end_b', ctypes.c_void_p)]


_lib = None
P = ctypes.c_void_p
I32, I64, DBL = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
SIZE = ctypes.c_size_t

_SIGNATURES = {
    'pq_last_error': ([], ctypes.c_char_p),
    'pq_device_count': ([P], ctypes.c_int),
    'pq_malloc': ([ctypes.c_int, SIZE, P], ctypes.c_int),
    'pq_free': ([P], ctypes.c_int),
    'pq_memcpy_h2d': ([P, P, SIZE], ctypes.c_int),
    'pq_memcpy_d2h': ([P, P, SIZE], ctypes.c_int),
    'pq_memcpy_d2d': ([P, P, SIZE], ctypes.c_int),
    'pq_memcpy_d2h_strided': ([P, P, SIZE, SIZE, SIZE], ctypes.c_int),
    'pq_memset': ([P, ctypes.c_int, SIZE], ctypes.c_int),
    'pq_synchronize': ([ctypes.c_int], ctypes.c_int),
    'pq_stream_create': ([ctypes.c_int, P], ctypes.c_int),
    'pq_stream_destroy': ([P], ctypes.c_int),
    'pq_stream_synchronize': ([P], ctypes.c_int),
    'pq_store_probe': ([P, I64, I64, ctypes.c_int, P], ctypes.c_int),
    'pq_fill_create': ([P, ctypes.c_int, P], ctypes.c_int),
    'pq_fill_destroy': ([P], ctypes.c_int),
    'pq_fill_symmetric': ([P, DBL, ctypes.c_int, ctypes.c_int, P, P], ctypes.c_int),
    'pq_pairs': ([P, P, P, P, P, I64, ctypes.c_int, P], ctypes.c_int),
    'pq_pairs_noalloc': ([P, P, P, P, P, I64, ctypes.c_int, P], ctypes.c_int),
    'pq_fill_device': ([P, DBL, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, I64, P, P], ctypes.c_int),
    'pq_fill_last_timing': ([P, P, P, P], ctypes.c_int),
    'pq_combine_device': ([ctypes.c_int, P, I64, I64, P, P, P, P, P, I64, DBL, ctypes.c_int, P, I64, P], ctypes.c_int),
    'pq_potential_device': ([ctypes.c_int, P, I64, I64, I64, P, P, P, DBL, ctypes.c_int, P, I64, P], ctypes.c_int),
    'pq_loops_create': ([ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, I64, P, ctypes.c_int, I64, ctypes.c_int,
                         P, P, P, P], ctypes.c_int),
    'pq_loops_create_tree': ([ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int, I64, ctypes.c_int, P, P,
                              P, P],
                             ctypes.c_int),
    'pq_loops_destroy': ([P], ctypes.c_int),
    'pq_csym_create': ([ctypes.c_int, ctypes.c_int, I64, P, P, P, P, P, P], ctypes.c_int),
    'pq_csym_solve': ([P, ctypes.c_int, P, DBL, ctypes.c_int, P, P, P], ctypes.c_int),
    'pq_csym_destroy': ([P], ctypes.c_int),
    'pq_hmatrix_create': ([ctypes.c_int, I64, I64, P, P, I64, P, I64, ctypes.c_int, P], ctypes.c_int),
    'pq_hmatrix_destroy': ([P], ctypes.c_int),
    'pq_hmatrix_info': ([P, P], ctypes.c_int),
    'pq_hmatrix_apply': ([P, I64, P, P], ctypes.c_int),
    'pq_hmatrix_apply_device': ([P, I64, P, P, P], ctypes.c_int),
    'pq_hmatrix_time': ([P, ctypes.c_int, P], ctypes.c_int),
    'pq_rwg_entries_create': ([P, I64, P, P, P, DBL, DBL, ctypes.c_int, P], ctypes.c_int),
    'pq_entries_destroy': ([P], ctypes.c_int),
    'pq_entries_block': ([P, P, I64, P, I64, P], ctypes.c_int),
    'pq_aca_blocks': ([P, I64, P, P, P, P, DBL, P, P, P], ctypes.c_int),
    'pq_dense_blocks': ([P, I64, P, P, P, P, P, P], ctypes.c_int),
    'pq_entries_factors': ([P, I64, P, P], ctypes.c_int),
    'pq_entries_hmatrix': ([P, I64, P, P], ctypes.c_int),
    'pq_entries_stats': ([P, P, P, P, P], ctypes.c_int),
    'pq_loops_matrix': ([P, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_last_fallbacks': ([P, P], ctypes.c_int),
    'pq_peec_electro': ([ctypes.c_int, P, I64, ctypes.c_int, ctypes.c_int, P, P, DBL, P, I64], ctypes.c_int),
    'pq_peec_solve': ([P, ctypes.c_int, P, I64, DBL, DBL, ctypes.c_int, P, P, P], ctypes.c_int),
    'pq_loops_solve_direct_multi': ([ctypes.c_int, P, P, P, ctypes.c_int, P, DBL, ctypes.c_int, P, P, P, P],
                                    ctypes.c_int),
    'pq_loops_solve_ports': ([P, DBL, ctypes.c_int, P, ctypes.c_int, DBL, ctypes.c_int, ctypes.c_int, P, P, P, P, P],
                             ctypes.c_int),
    'pq_loops_factor_batch': ([P, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_solve_direct_batch': ([P, ctypes.c_int, P, ctypes.c_int, P, DBL, ctypes.c_int, P, P, P], ctypes.c_int),
    'pq_loops_solve_direct': ([P, DBL, ctypes.c_int, P, DBL, ctypes.c_int, P, P, P, P], ctypes.c_int),
    'pq_loops_factor_time': ([P, P], ctypes.c_int),
    'pq_spd_solve': ([ctypes.c_int, ctypes.c_int, P, I64, ctypes.c_int, P, DBL, ctypes.c_int, P, P, P, P],
                     ctypes.c_int),
    'pq_loops_set_coarse': ([P, ctypes.c_int, P, ctypes.c_int], ctypes.c_int),
    'pq_loops_coarse_matrices': ([P, P, P], ctypes.c_int),
    'pq_biot_field': ([ctypes.c_int, ctypes.c_int, ctypes.c_int, I64, P, I64, P, I64, P, P, P], ctypes.c_int),
    'pq_biot_grid': ([ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, DBL, DBL,
                      ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, P, P], ctypes.c_int),
    'pq_mpnn_gemm': ([ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, P], ctypes.c_int),
    'pq_mpnn_workspace': ([ctypes.c_int, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
    'pq_mpnn_gemm_tn': ([ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, SIZE, P, P, P], ctypes.c_int),
    'pq_mpnn_aggregate': ([ctypes.c_int, P, P, P, P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P], ctypes.c_int),
    'pq_mpnn_relu_mask': ([SIZE, P, P, P, P], ctypes.c_int),
    'pq_loops_rhs': ([P, DBL, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_operator': ([P, DBL, ctypes.c_int, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_coarse_solution': ([P, DBL, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_set_coarse_diagonal': ([P, ctypes.c_int, P, DBL, P], ctypes.c_int),
    'pq_loops_set_coarse_csc': ([P, ctypes.c_int, P, P, P, DBL, P, P], ctypes.c_int),
    'pq_loops_prepare_block_jacobi': ([P, P], ctypes.c_int),
    'pq_loops_last_solve': ([P, P, P, P, P], ctypes.c_int),
    'pq_loops_tuning': ([P, ctypes.c_int, P, P], ctypes.c_int),
    'pq_loops_use_graphs': ([P, ctypes.c_int], ctypes.c_int),
    'pq_loop_basis_create': ([I64, I64, P, P, I64, I64, P, ctypes.c_int, P], ctypes.c_int),
    'pq_loop_basis_sizes': ([P, P], ctypes.c_int),
    'pq_loop_basis_copy': ([P, P, P, P, P, P, P, P, P, P, P, P], ctypes.c_int),
    'pq_loop_basis_destroy': ([P], ctypes.c_int),
    'pq_sheet_resistance': ([I64, P, P, P, P, P, I64, P, P, P, DBL, P], ctypes.c_int),
    'pq_loop_resistance': ([P, P, I64, P, P], ctypes.c_int),
    'pq_csr_sizes': ([P, P], ctypes.c_int),
    'pq_csr_copy': ([P, P, P, P], ctypes.c_int),
    'pq_csr_destroy': ([P], ctypes.c_int),
}


def exported_symbols():
    """Every entry point this binding declares (include/pcbqs_amd.h)."""
    return sorted(_SIGNATURES)


def lib():
    """The loaded library (RuntimeError when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f'{LIBRARY} is not built: run __graft_entry__.build() '
                               '(make -C pcbqs_amd/csrc); there is no CPU fallback')
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7). Whichever is
        # loaded first serves the whole process, so torch goes first: the learned correction's
        # tensors and this library then share one runtime (and torch gets the one it was built for).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIBRARY)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes, fn.restype = args, res
        _lib = handle
    return _lib


def check(status, what=''):
    """Raise the Python exception of a nonzero pq_status."""
    if status == PQ_OK:
        return
    message = lib().pq_last_error().decode(errors='replace') or what
    if status == PQ_EINVAL:
        raise ValueError(message)
    if status == PQ_ENONFINITE:
        raise FloatingPointError(message)
    if status == PQ_ENOMEM:
        raise MemoryError(message)
    if status == PQ_ECONVERGE:
        raise ConvergenceError(message)
    raise RuntimeError(message)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def device_count():
    n = ctypes.c_int(0)
    call('pq_device_count', ctypes.byref(n))
    return n.value


def ptr(a):
    """Address of a C-contiguous numpy array."""
    return ctypes.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device allocation with a dtype and shape, owned by this object (freed on close/GC)."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.device = int(device)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        call('pq_malloc', self.device, max(self.nbytes, 1), ctypes.byref(p))
        self._ptr = p

    @classmethod
    def from_host(cls, array, device=0, dtype=None):
        a = np.ascontiguousarray(array, dtype=dtype)
        d = cls(a.shape, a.dtype, device)
        if a.nbytes:
            call('pq_memcpy_h2d', d._ptr, ptr(a), a.nbytes)
        return d

    @property
    def ptr(self):
        if self._ptr is None:
            raise RuntimeError('device array already freed')
        return self._ptr

    def offset(self, elements):
        return ctypes.c_void_p(self.ptr.value + int(elements) * self.dtype.itemsize)

    def host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call('pq_memcpy_d2h', ptr(out), self.ptr, self.nbytes)
        return out

    def zero(self):
        if self.nbytes:
            call('pq_memset', self.ptr, 0, self.nbytes)
        return self

    def close(self):
        if self._ptr is not None and _lib is not None:
            _lib.pq_free(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                   # noqa: BLE001 (interpreter shutdown)
            pass
