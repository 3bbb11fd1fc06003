"""GPU parity of the loop operator product W = (Rl + j w Ll) X, the GMRES matvec, on synthetic loop
systems of chosen sizes (identity loop basis, so Ll is exactly the random symmetric FP32 branch
matrix): the symmetric upper-triangle kernel and the whole-matrix panel kernel against an FP64
numpy product of the same matrices. The sizes cover every ragged edge of the symmetric kernel's
256 x 256 tiles and 64-column panels (a single partial tile, a last column block narrower than one
panel, a partial last panel, exact multiples)."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from pcbqs_amd._lib import DeviceArray, call, lib, ptr

pytestmark = pytest.mark.gpu


def synthetic_loops(n, ports, seed, lb_precision=32):
    """A loop handle with n free loops and `ports` port loops: C = I, Lb random symmetric (FP32, or
    FP64 with lb_precision 64), Rl random sparse symmetric FP64; returns (handle, Ll (nl, nl) FP64,
    Rl CSR)."""
    rng = np.random.default_rng(seed)
    nl = n + ports
    dtype = np.float64 if lb_precision == 64 else np.float32
    A = rng.standard_normal((nl, nl)).astype(dtype)
    Lb = ((A + A.T) * dtype(0.5)).astype(dtype)
    ldb = (nl + 63) // 64 * 64
    host = np.zeros((nl, ldb), dtype)
    host[:, :nl] = Lb
    dev = DeviceArray.from_host(host)
    colptr = np.arange(nl + 1, dtype=np.int32)
    rows = np.arange(nl, dtype=np.int32)
    sign = np.ones(nl, np.float32)
    R = sp.random(nl, nl, density=min(1.0, 24.0 / nl), random_state=seed, format='csr')
    R = (R + R.T + sp.identity(nl)).tocsr()
    R.sort_indices()
    rptr, rcol = np.ascontiguousarray(R.indptr, np.int32), np.ascontiguousarray(R.indices, np.int32)
    rval = np.ascontiguousarray(R.data, np.float64)
    handle = ctypes.c_void_p()
    call('pq_loops_create', 0, nl, n, ptr(colptr), ptr(rows), ptr(sign), nl, dev.ptr, lb_precision, ldb, nl, ptr(rptr),
         ptr(rcol), ptr(rval), ctypes.byref(handle))
    dev.close()
    return handle, Lb.astype(np.float64), R


def apply(handle, omega, X, precision=32):
    n, P = X.shape
    Xs = np.ascontiguousarray(np.hstack([X.real, X.imag]))
    W = np.empty_like(Xs)
    call('pq_loops_operator', handle, omega, P, precision, ptr(Xs), ptr(W))
    return W[:, :P] + 1j * W[:, P:]


def tuning(handle, NR):
    chunk, us = ctypes.c_int(), ctypes.c_double()
    call('pq_loops_tuning', handle, NR, ctypes.byref(chunk), ctypes.byref(us))
    return chunk.value


@pytest.mark.parametrize('mode', ['symmetric', 'panel'])
@pytest.mark.parametrize('n', [40, 200, 256, 300, 600, 1024, 1500, 2100])
def test_operator_matches_the_fp64_product(n, mode, monkeypatch):
    monkeypatch.setenv('PCBQS_AMD_OPERATOR', mode)
    print(f'operator n={n} mode={mode}', flush=True)
    handle, Ll, R = synthetic_loops(n, 2, n)
    try:
        rng = np.random.default_rng(n + 1)
        omega = 2 * np.pi * 1e6
        Rf, Lf = R[:n, :n], Ll[:n, :n]
        for P in (1, 2, 4, 8, 16):
            X = rng.standard_normal((n, P)) + 1j * rng.standard_normal((n, P))
            W = apply(handle, omega, X)
            ref = Rf @ X + 1j * omega * (Lf @ X)
            err = np.abs(W - ref).max() / np.abs(ref).max()
            assert err < 1e-13, (n, mode, P, err)
            chunk = tuning(handle, 2 * P)
            if mode == 'symmetric' and P <= 8:
                assert chunk == 0                          # the symmetric kernel ran
            if mode == 'panel':
                assert chunk > 0
    finally:
        lib().pq_loops_destroy(handle)


def test_autotuner_result_is_deterministic_and_agrees_across_kernels(monkeypatch):
    """Both kernels on one 2,100-loop system: repeated products are bitwise equal, and the two
    kernels (different summation orders) agree to FP64 rounding."""
    n = 2100
    out = {}
    for mode in ('symmetric', 'panel'):
        monkeypatch.setenv('PCBQS_AMD_OPERATOR', mode)
        handle, _, _ = synthetic_loops(n, 3, 7)
        try:
            rng = np.random.default_rng(3)
            X = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
            a, b = apply(handle, 1e7, X), apply(handle, 1e7, X)
            assert np.array_equal(a, b)
            out[mode] = a
        finally:
            lib().pq_loops_destroy(handle)
    assert np.abs(out['symmetric'] - out['panel']).max() / np.abs(out['panel']).max() < 1e-13


@pytest.mark.parametrize('n', [63, 300, 1500])
def test_fp64_product_reads_the_fp64_matrix(n):
    """precision 64 (the residual product) uses the FP64 Ll: with an FP64 branch matrix that FP32
    cannot represent it agrees with numpy to FP64 rounding, while the FP32 copy's product (the
    Krylov matvec) differs by FP32 rounding; the matrix read back is Lb itself."""
    print(f'operator fp64 n={n}', flush=True)
    handle, Ll, R = synthetic_loops(n, 3, n + 5, lb_precision=64)
    try:
        rng = np.random.default_rng(n)
        omega = 2 * np.pi * 1e6
        X = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
        ref = R[:n, :n] @ X + 1j * omega * (Ll[:n, :n] @ X)
        err64 = np.abs(apply(handle, omega, X, 64) - ref).max() / np.abs(ref).max()
        err32 = np.abs(apply(handle, omega, X, 32) - ref).max() / np.abs(ref).max()
        assert err64 < 1e-13, err64
        assert 1e-10 < err32 < 1e-5, err32
        dev, ld = ctypes.c_void_p(), ctypes.c_int64()
        nl = Ll.shape[0]
        call('pq_loops_matrix', handle, 64, ctypes.byref(dev), ctypes.byref(ld))
        back = np.empty((nl, ld.value))
        call('pq_memcpy_d2h', ptr(back), dev, back.nbytes)
        assert np.array_equal(back[:, :nl], Ll)
        call('pq_loops_matrix', handle, 32, ctypes.byref(dev), ctypes.byref(ld))
        back32 = np.empty((nl, ld.value), np.float32)
        call('pq_memcpy_d2h', ptr(back32), dev, back32.nbytes)
        assert np.array_equal(back32[:, :nl], Ll.astype(np.float32))
    finally:
        lib().pq_loops_destroy(handle)
