"""GPU parity of the direct port solve: Z_ff = Rl + j w Ll formed in FP32 on the device, factored
as L D L^T (64-wide panels, MFMA trailing updates) and refined in FP64 to a relative residual of
1e-12 (the reference's MixedLU bar, src/pcbqs/kernels.py:125-181).

- synthetic loop systems with symmetric positive definite Rl and Ll at sizes that cover every
  ragged panel shape (one partial panel, exact multiples of 64, a last panel of one row, several
  trailing tiles with a partial last tile), P = 1 .. 16: against numpy's FP64 solve of the same
  (FP32-stored Ll) matrices;
- boards: port R and L against the oracle's FP64 dense solve within 1e-6, and against GMRES;
- the copper fixture at DC (Z_ff = Rl alone) and 100 / 200 kHz;
- a system whose FP32 factor cannot reach 1e-12 (condition number about 5e7) goes through the FP64
  fallback factor and matches numpy's FP64 solve, alone and inside a batch whose other frequencies
  keep their refined solutions; an exactly singular system raises instead of returning garbage;
- above 4,096 unknowns, the one-launch sweeps (per-block flags) against the per-block launches."""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp

from pcbqs_amd._lib import DeviceArray, call, lib, ptr

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def spd_loops(n, ports, seed, singular_loop=None):
    """A loop handle with n free and `ports` port loops: C = I, Lb = G G^T / nl + I/2 (symmetric
    positive definite, FP32) and Rl diagonally dominant sparse symmetric (FP64). singular_loop: a
    free loop whose row and column of both matrices are zero. Returns (handle, Ll FP64, Rl dense)."""
    rng = np.random.default_rng(seed)
    nl = n + ports
    G = rng.standard_normal((nl, nl))
    Lb = (G @ G.T / nl + 0.5 * np.eye(nl)) * 1e-9
    R = sp.random(nl, nl, density=min(1.0, 8.0 / nl), random_state=seed).toarray() * 1e-4
    R = R + R.T
    R += np.diag(np.abs(R).sum(1) + 1e-3)
    if singular_loop is not None:
        for M in (Lb, R):
            M[singular_loop, :] = 0
            M[:, singular_loop] = 0
    Lb = Lb.astype(np.float32)
    ldb = (nl + 63) // 64 * 64
    host = np.zeros((nl, ldb), np.float32)
    host[:, :nl] = Lb
    dev = DeviceArray.from_host(host)
    Rs = sp.csr_matrix(R)
    Rs.sort_indices()
    colptr = np.arange(nl + 1, dtype=np.int32)
    rows = np.arange(nl, dtype=np.int32)
    sign = np.ones(nl, np.float32)
    handle = ctypes.c_void_p()
    call('pq_loops_create', 0, nl, n, ptr(colptr), ptr(rows), ptr(sign), nl, dev.ptr, 32, ldb, nl,
         ptr(np.ascontiguousarray(Rs.indptr, np.int32)), ptr(np.ascontiguousarray(Rs.indices, np.int32)),
         ptr(np.ascontiguousarray(Rs.data, np.float64)), ctypes.byref(handle))
    dev.close()
    return handle, Lb.astype(np.float64), R


def direct(handle, omega, n, P, cols, tol=1e-12, steps=12):
    zp = np.zeros((P, P, 2))
    xh = np.zeros((n, 2 * P))
    its, res = ctypes.c_int64(), ctypes.c_double()
    call('pq_loops_solve_direct', handle, omega, P, ptr(np.ascontiguousarray(cols, np.int32)), tol, steps, ptr(zp),
         ptr(xh), ctypes.byref(its), ctypes.byref(res))
    return zp[..., 0] + 1j * zp[..., 1], xh[:, :P] + 1j * xh[:, P:], its.value, res.value


def expected(Ll, R, omega, n, cols):
    Z = R + 1j * omega * Ll
    Zff, Zfs = Z[:n, :n], Z[:n][:, cols]
    X = np.linalg.solve(Zff, -Zfs)
    return Z[np.ix_(cols, cols)] + Z[cols, :n] @ X, X


@pytest.mark.parametrize('n', [1, 5, 63, 64, 65, 130, 200, 321, 700])
def test_direct_solve_matches_the_fp64_solve(n):
    ports = 16
    print(f'direct n={n}', flush=True)
    handle, Ll, R = spd_loops(n, ports, n)
    try:
        for P in (1, 3, 4, 8, 16):
            cols = np.arange(n, n + P)
            for omega in (0.0, 2 * np.pi * 1e6, 2 * np.pi * 1e9):
                Zp, X, its, res = direct(handle, omega, n, P, cols)
                Ze, Xe = expected(Ll, R, omega, n, cols)
                assert res <= 1e-12, (n, P, omega, res)
                assert np.abs(X - Xe).max() <= 1e-9 * np.abs(Xe).max(), (n, P, omega)
                assert np.abs(Zp - Ze).max() <= 1e-10 * np.abs(Ze).max(), (n, P, omega)
    finally:
        lib().pq_loops_destroy(handle)


def test_factor_is_reused_at_the_same_frequency_and_rebuilt_at_another():
    n = 300
    handle, Ll, R = spd_loops(n, 4, 11)
    try:
        cols = np.arange(n, n + 4)
        w = 2 * np.pi * 1e6
        a = direct(handle, w, n, 4, cols)
        b = direct(handle, w, n, 4, cols)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        c = direct(handle, 2 * w, n, 4, cols)
        Ze, _ = expected(Ll, R, 2 * w, n, cols)
        assert np.abs(c[0] - Ze).max() <= 1e-10 * np.abs(Ze).max()
        d = direct(handle, w, n, 2, cols[:2])
        Ze, _ = expected(Ll, R, w, n, cols[:2])
        assert np.abs(d[0] - Ze).max() <= 1e-10 * np.abs(Ze).max()
    finally:
        lib().pq_loops_destroy(handle)


def test_singular_system_raises():
    """A zero row and column: the FP32 factor meets a zero pivot, the FP64 fallback finds the
    matrix singular, and the solve raises."""
    from pcbqs_amd._lib import ConvergenceError
    n = 100
    handle, _, _ = spd_loops(n, 2, 5, singular_loop=70)
    try:
        with pytest.raises(ConvergenceError, match='singular'):
            direct(handle, 2 * np.pi * 1e6, n, 2, np.arange(n, n + 2))
    finally:
        lib().pq_loops_destroy(handle)


def ill_conditioned_loops(n, ports, seed):
    """C = I; Lb = Q diag(logspace(0, -8)) Q^T 1e-9 (FP32-stored) and Rl = 1e-10 I: at 1 MHz the
    free block's condition number is about 5e7, beyond what an FP32 factor can refine."""
    rng = np.random.default_rng(seed)
    nl = n + ports
    Q, _ = np.linalg.qr(rng.standard_normal((nl, nl)))
    Lb = ((Q * np.logspace(0, -8, nl)) @ Q.T * 1e-9)
    Lb = ((Lb + Lb.T) / 2).astype(np.float32)
    ldb = (nl + 63) // 64 * 64
    host = np.zeros((nl, ldb), np.float32)
    host[:, :nl] = Lb
    dev = DeviceArray.from_host(host)
    Rs = sp.identity(nl, format='csr') * 1e-10
    Rs.sort_indices()
    colptr = np.arange(nl + 1, dtype=np.int32)
    rows = np.arange(nl, dtype=np.int32)
    sign = np.ones(nl, np.float32)
    handle = ctypes.c_void_p()
    call('pq_loops_create', 0, nl, n, ptr(colptr), ptr(rows), ptr(sign), nl, dev.ptr, 32, ldb, nl,
         ptr(np.ascontiguousarray(Rs.indptr, np.int32)), ptr(np.ascontiguousarray(Rs.indices, np.int32)),
         ptr(np.ascontiguousarray(Rs.data, np.float64)), ctypes.byref(handle))
    dev.close()
    return handle, Lb.astype(np.float64), Rs.toarray()


def fallbacks(handle):
    count = ctypes.c_int64()
    call('pq_loops_last_fallbacks', handle, ctypes.byref(count))
    return count.value


def test_fp64_fallback_factor_solves_what_fp32_cannot():
    n, P = 300, 4
    print(f'direct fallback n={n}', flush=True)
    handle, Ll, R = ill_conditioned_loops(n, P, 3)
    try:
        cols = np.arange(n, n + P)
        omega = 2 * np.pi * 1e6
        Zff = (R + 1j * omega * Ll)[:n, :n]
        assert np.linalg.cond(Zff) > 1e7
        Zp, X, its, res = direct(handle, omega, n, P, cols)
        assert fallbacks(handle) == 1
        Ze, Xe = expected(Ll, R, omega, n, cols)
        assert res <= 1e-12, res
        assert np.abs(X - Xe).max() <= 1e-6 * np.abs(Xe).max()
        assert np.abs(Zp - Ze).max() <= 1e-8 * np.abs(Ze).max()
        # a well-conditioned frequency of the same handle needs no fallback
        Zp, X, its, res = direct(handle, 0.0, n, P, cols)
        assert fallbacks(handle) == 0 and res <= 1e-12
    finally:
        lib().pq_loops_destroy(handle)


def test_one_launch_sweeps_equal_the_per_block_sweeps(monkeypatch):
    """Above 4,096 unknowns each triangular sweep is one launch whose workgroups wait on per-block
    flags; it gives the same values as one launch per 64-row block (PCBQS_AMD_LDL_BLOCK_SWEEPS set),
    alone (the factor kept, so only the sweeps differ) and batched over frequencies, and the refined
    solution matches numpy's FP64 solve. 4,500 unknowns: 70 full blocks and a partial one."""
    n = 4500
    print(f'one-launch sweeps n={n}', flush=True)
    handle, Ll, R = spd_loops(n, 16, 21)
    try:
        w = 2 * np.pi * 1e6
        for P in (1, 8, 16):
            cols = np.arange(n, n + P)
            monkeypatch.delenv('PCBQS_AMD_LDL_BLOCK_SWEEPS', raising=False)
            a = direct(handle, w, n, P, cols)
            monkeypatch.setenv('PCBQS_AMD_LDL_BLOCK_SWEEPS', '1')
            b = direct(handle, w, n, P, cols)
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0]), P
            assert a[2] == b[2] and a[3] <= 1e-12, (P, a[2], b[2], a[3])
            if P == 8:
                Ze, Xe = expected(Ll, R, w, n, cols)
                assert np.abs(a[1] - Xe).max() <= 1e-9 * np.abs(Xe).max()
                assert np.abs(a[0] - Ze).max() <= 1e-10 * np.abs(Ze).max()
        P = 4
        cols = np.ascontiguousarray(np.arange(n, n + P), np.int32)
        omegas = np.array([2 * np.pi * 1e3, 2 * np.pi * 1e7])
        out = []
        for env in (None, '1'):
            if env is None:
                monkeypatch.delenv('PCBQS_AMD_LDL_BLOCK_SWEEPS', raising=False)
            else:
                monkeypatch.setenv('PCBQS_AMD_LDL_BLOCK_SWEEPS', env)
            zp = np.zeros((2, P, P, 2))
            st, rs = np.zeros(2, np.int64), np.zeros(2)
            call('pq_loops_solve_direct_batch', handle, 2, ptr(omegas), P, ptr(cols), 1e-12, 12, ptr(zp), ptr(st),
                 ptr(rs))
            assert (rs <= 1e-12).all(), rs
            out.append(zp)
        assert np.array_equal(out[0], out[1])
    finally:
        lib().pq_loops_destroy(handle)


def test_batched_direct_solve_falls_back_per_frequency():
    """In one batched call, only the ill-conditioned frequency takes the FP64 factor; the others
    keep their FP32-factor solutions, and all match numpy."""
    n, P = 300, 4
    handle, Ll, R = ill_conditioned_loops(n, P, 4)
    try:
        cols = np.ascontiguousarray(np.arange(n, n + P), np.int32)
        omegas = np.array([0.0, 2 * np.pi * 1e3, 2 * np.pi * 1e6])
        zp = np.zeros((3, P, P, 2))
        st, rs = np.zeros(3, np.int64), np.zeros(3)
        call('pq_loops_solve_direct_batch', handle, 3, ptr(omegas), P, ptr(cols), 1e-12, 12, ptr(zp), ptr(st), ptr(rs))
        assert fallbacks(handle) == 1
        for b, w in enumerate(omegas):
            Ze, _ = expected(Ll, R, w, n, cols)
            Z = zp[b, ..., 0] + 1j * zp[b, ..., 1]
            assert np.abs(Z - Ze).max() <= 1e-8 * np.abs(Ze).max(), b
            assert rs[b] <= 1e-12, (b, rs[b])
    finally:
        lib().pq_loops_destroy(handle)


@pytest.mark.parametrize('name', ['c1', 'c2-small'])
def test_direct_port_impedance_matches_the_oracle_and_gmres(name):
    from oracle import mqs as M
    from pcbqs_amd import board as boards
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c1() if name == 'c1' else boards.board_c2(1, target=3000)
    freqs = [1e3, 1e6, 1e8]
    print(f'direct ports {name} nt={b.nt}', flush=True)
    s = PortSolver(b, device=0)
    Z, info = s.ports(freqs, solver='direct', currents=False)
    Zg, _ = s.ports(freqs, precond='block-jacobi', tol=1e-10)
    q = M.CopperModel(b.nodes, b.tris, b.z_of, b.t_of, b.node_of, b.terminals, b.vseg, b.via_nodes, b.ports)
    q.assemble()
    Zo = q.port_response(freqs)[0]
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    rel = lambda a, c: np.abs(a - c).max() / np.abs(c).max()
    assert rel(Z.real, Zo.real) < 1e-6
    assert rel(Z.imag / w, Zo.imag / w) < 1e-6
    assert rel(Z, Zg) < 1e-8
    assert all(i['relative_residual'] <= 1e-12 and i['iterations'] <= 12 for i in info)
    s.close()


def test_copper_reference_fixture_through_the_direct_solve():
    from oracle import mqs as M
    from pcbqs_amd.board import Board
    from pcbqs_amd.solve import PortSolver
    with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
        exp = {k: data[k] for k in data.files}
    q = M.copper_reference_model(exp['nodes'], exp['tris'])
    b = Board(q.nodes, q.tris, q.z_of, q.t_of, np.zeros(len(q.tris), int), q.node_of, q.terminals, q.vseg,
              q.via_nodes, q.ports)
    s = PortSolver(b, device=0, ideal_terminals=False)
    Z, info = s.ports([0, 1e5, 2e5], solver='direct')
    np.testing.assert_allclose(Z, exp['impedance'], rtol=1e-6)
    s.close()


def test_batched_factorisations_give_the_single_frequency_factors(monkeypatch):
    """The frequencies of one call are factored together (pq_loops_factor_batch: the L D L^T kernels
    take the frequency as a grid dimension); the port impedances equal those of one factorisation
    per frequency (PCBQS_AMD_DIRECT_BATCH=0) and match the oracle."""
    from pcbqs_amd import board as boards
    from pcbqs_amd.solve import PortSolver
    from tests.test_solve_gpu import oracle_ports, rel
    b = boards.board_batch(0)
    freqs = [0.0, 1e3, 1e6, 3e7, 1e8]
    out = {}
    for batch in ('0', '1'):
        monkeypatch.setenv('PCBQS_AMD_DIRECT_BATCH', batch)
        s = PortSolver(b, device=0)
        out[batch] = s.ports(freqs, solver='direct')
        s.close()
    (Z0, i0), (Z1, i1) = out['0'], out['1']
    print(f'batched vs single: max diff {np.abs(Z1 - Z0).max() / np.abs(Z0).max():.2e}, '
          f'factor ms {[round(r["factor_ms"], 2) for r in i1]} vs {[round(r["factor_ms"], 2) for r in i0]}', flush=True)
    assert np.array_equal(Z1, Z0)
    assert [r['iterations'] for r in i1] == [r['iterations'] for r in i0]
    assert all(r['relative_residual'] <= 1e-12 for r in i1)
    Zo = oracle_ports(b, freqs)
    w = 2 * np.pi * np.maximum(np.array(freqs), 1e-300)[:, None, None]
    assert rel(Z1.real, Zo.real) < 1e-6
    assert rel(Z1[1:].imag / w[1:], Zo[1:].imag / w[1:]) < 1e-6
