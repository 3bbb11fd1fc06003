"""GPU parity of the loop/port solve: the device pipeline (FP32-bulk fill, device combine, loop
projection, block-Jacobi GMRES to 1e-8) against the oracle's FP64 direct solve on the same boards:
port R and L within 1e-6 relative; and against the reference's frozen copper-reference.npz."""
import os

import numpy as np
import pytest

from oracle import mqs as M
from pcbqs_amd import board as boards
from pcbqs_amd.board import Board

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def oracle_ports(b, freqs, ideal_terminals=True):
    q = M.CopperModel(b.nodes, b.tris, b.z_of, b.t_of, b.node_of, b.terminals, b.vseg, b.via_nodes, b.ports,
                      ideal_terminals=ideal_terminals)
    q.assemble()
    return q.port_response(freqs)[0]


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize('name', ['c1', 'c2-small'])
@pytest.mark.parametrize('precond', ['block-jacobi', 'none', 'two-level'])
def test_port_impedance_matches_the_fp64_oracle(name, precond):
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c1() if name == 'c1' else boards.board_c2(1, target=3000)
    freqs = [1e3, 1e6, 1e8]
    print(f'ports {name} {precond} nt={b.nt}', flush=True)
    s = PortSolver(b, device=0)
    if precond == 'two-level':
        s.set_stream_coarse(grid=6)
    Z, info = s.ports(freqs, precond=precond, tol=1e-8)
    Zo = oracle_ports(b, freqs)
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    assert rel(Z.real, Zo.real) < 1e-6
    assert rel(Z.imag / w, Zo.imag / w) < 1e-6
    assert all(i['relative_residual'] <= 1e-8 for i in info)
    s.close()


def test_copper_reference_fixture_through_the_device_pipeline():
    from pcbqs_amd.solve import PortSolver
    with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
        expected = {k: data[k] for k in data.files}
    q = M.copper_reference_model(expected['nodes'], expected['tris'])
    b = Board(q.nodes, q.tris, q.z_of, q.t_of, np.zeros(len(q.tris), int), q.node_of, q.terminals, q.vseg,
              q.via_nodes, q.ports)
    s = PortSolver(b, device=0, ideal_terminals=False)
    Z, info = s.ports([0, 1e5, 2e5], tol=1e-10)
    np.testing.assert_allclose(Z, expected['impedance'], rtol=1e-6)
    s.close()


@pytest.mark.parametrize('precond', ['block-jacobi', 'two-level'])
def test_symmetric_operator_solve_matches_the_fp64_oracle(precond, monkeypatch):
    """The port solve with the operator product pinned to the symmetric upper-triangle kernel."""
    from pcbqs_amd.solve import PortSolver
    monkeypatch.setenv('PCBQS_AMD_OPERATOR', 'symmetric')
    b = boards.board_c2(1, target=3000)
    freqs = [1e3, 1e6, 1e8]
    s = PortSolver(b, device=0)
    if precond == 'two-level':
        s.set_stream_coarse(grid=6)
    Z, info = s.ports(freqs, precond=precond, tol=1e-8)
    assert s.tuning()[0] == 0
    Zo = oracle_ports(b, freqs)
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    assert rel(Z.real, Zo.real) < 1e-6
    assert rel(Z.imag / w, Zo.imag / w) < 1e-6
    s.close()


@pytest.mark.parametrize('operator', ['auto', 'symmetric'])
@pytest.mark.parametrize('precond', ['block-jacobi', 'two-level'])
def test_captured_steps_replay_the_eager_solve_bitwise(precond, operator, monkeypatch):
    """The Arnoldi steps captured as graphs and replayed (across restarts and frequencies, with the
    solver's eager work in between) give the eager launches' results bit for bit."""
    from pcbqs_amd.solve import PortSolver
    monkeypatch.setenv('PCBQS_AMD_OPERATOR', operator)
    b = boards.board_c2(2, target=3000)
    freqs = [1e3, 1e6, 1e8, 1e6]
    out = {}
    for graphs in ('0', '1'):
        monkeypatch.setenv('PCBQS_AMD_GRAPHS', graphs)
        s = PortSolver(b, device=0)
        s.use_graphs(None)                       # follow the environment
        if precond == 'two-level':
            s.set_stream_coarse(grid=6)
        out[graphs] = s.ports(freqs, precond=precond, tol=1e-9, restart=30)
        s.close()
    (Z0, i0), (Z1, i1) = out['0'], out['1']
    assert [r['iterations'] for r in i0] == [r['iterations'] for r in i1]
    assert np.array_equal(Z0, Z1)


def test_fp16_coarse_bases_give_the_fp32_bases_solution(monkeypatch):
    """The two-level solve reads FP16 copies of its coarse bases by default; with the FP32 bases
    (PCBQS_AMD_COARSE_FP16=0) the port Z agree to the solve tolerance and the iteration counts
    stay within two of each other, and both match the FP64 oracle."""
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c2(1, target=3000)
    freqs = [1e3, 1e6, 1e8]
    out = {}
    for half in ('0', '1'):
        monkeypatch.setenv('PCBQS_AMD_COARSE_FP16', half)
        s = PortSolver(b, device=0)
        s.set_stream_coarse(grid=8)
        out[half] = s.ports(freqs, precond='two-level', tol=1e-10)
        s.close()
    (Z0, i0), (Z1, i1) = out['0'], out['1']
    assert rel(Z1, Z0) < 1e-8
    its = [(a['iterations'], c['iterations']) for a, c in zip(i0, i1)]
    assert all(abs(a - c) <= 2 for a, c in its), its
    Zo = oracle_ports(b, freqs)
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    assert rel(Z1.real, Zo.real) < 1e-6
    assert rel(Z1.imag / w, Zo.imag / w) < 1e-6


@pytest.mark.parametrize('name', ['c1', 'c2-small', 'batch-11-ports', 'copper-reference'])
def test_tree_projection_matches_the_column_projection(name):
    """Ll = C^T Lb C by prefix sums down the loop basis's tree (the default) against the sparse
    column product: the same FP32 matrix to within FP32 rounding of the FP64 sums, bitwise
    symmetric, on boards with vias, many ports and non-ideal terminals."""
    from pcbqs_amd.loops import LoopModel
    from pcbqs_amd.solve import PortSolver
    ideal = True
    if name == 'copper-reference':
        with np.load(os.path.join(GOLDEN, 'copper-reference.npz')) as data:
            q = M.copper_reference_model(data['nodes'], data['tris'])
        b = Board(q.nodes, q.tris, q.z_of, q.t_of, np.zeros(len(q.tris), int), q.node_of, q.terminals, q.vseg,
                  q.via_nodes, q.ports)
        ideal = False
    else:
        b = {'c1': boards.board_c1, 'c2-small': lambda: boards.board_c2(1, target=3000),
             'batch-11-ports': lambda: boards.board_batch(0)}[name]()
    m = LoopModel(b, ideal_terminals=ideal)
    print(f'projection {name}: {m.ne + m.nv} branches, {m.n_loops} loops', flush=True)
    L = {}
    for how in ('columns', 'tree'):
        s = PortSolver(b, device=0, ideal_terminals=ideal, model=m, projection=how)
        L[how] = s.loop_matrix().astype(np.float64)
        print(how, s.setup_seconds, flush=True)
        s.close()
    assert np.array_equal(L['tree'], L['tree'].T)
    assert rel(L['tree'], L['columns']) < 1e-6


def test_failed_two_level_solve_falls_back_to_block_jacobi():
    """A two-level solve that fails its acceptance test (here: at most 2 iterations) is solved again
    by block-Jacobi GMRES from zero; the result is the plain solve's and info names the fallback."""
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c1()
    freqs = [1e3, 1e6]
    s = PortSolver(b, device=0)
    s.set_stream_coarse(grid=6)
    Zp, plain = s.ports(freqs, precond='block-jacobi', tol=1e-8)
    Z, info = s.ports(freqs, precond='two-level', tol=1e-8, maxiter=2)
    assert all(i['preconditioner'] == 'block-jacobi' and i['fallback_from'] == 'two-level' for i in info)
    assert all(i['relative_residual'] <= 1e-8 for i in info)
    assert np.array_equal(Z, Zp)
    from pcbqs_amd._lib import ConvergenceError
    with pytest.raises(ConvergenceError):
        s.ports(freqs, precond='block-jacobi', tol=1e-8, maxiter=2)
    s.close()


def test_device_diagonalised_coarse_space_equals_the_host_one():
    """The coarse space diagonalised on the device (rocSOLVER) and on the host (numpy eigh) span the
    same space: the coarse solution Phi (Phi^T Z Phi)^-1 Phi^T B does not depend on the basis, so the
    two agree to the FP32 bases' rounding; the two-level solves give the same port impedances."""
    from pcbqs_amd.coarse import loop_restrict, stream_patterns, via_loop_columns
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c2(1, target=3000)
    s = PortSolver(b, device=0)
    m = s.model
    phi = np.hstack([loop_restrict(m, stream_patterns(m, 10)), via_loop_columns(m)])
    out = {}
    for how in ('host', True):
        s.set_coarse(phi, diagonalize=how)
        k = s.coarse_k
        xc = {f: s.coarse_solution(f) for f in (1e3, 1e6, 1e8)}
        Z, info = s.ports([1e3, 1e6, 1e8], precond='two-level', tol=1e-10)
        out[how] = (k, xc, Z, [i['iterations'] for i in info])
    (k0, x0, Z0, it0), (k1, x1, Z1, it1) = out['host'], out[True]
    print(f'coarse columns host {k0} device {k1}; iterations {it0} vs {it1}', flush=True)
    assert k0 == k1
    for f in x0:
        assert np.abs(x1[f] - x0[f]).max() <= 1e-5 * np.abs(x0[f]).max(), f
    assert np.abs(Z1 - Z0).max() <= 1e-9 * np.abs(Z0).max()
    assert all(abs(a - c) <= max(3, 0.1 * a) for a, c in zip(it0, it1))
    s.close()


def test_sparse_coarse_columns_take_the_cholesky_route_unless_repeated():
    """Coarse columns sent as sparse columns (pq_loops_set_coarse_csc). Without repeated columns the
    Galerkin resistance's Cholesky factor passes its guard and one eigensolve diagonalises the pair;
    with a repeated column the guard fails and the eigenvector route drops the repeat. Either way
    the coarse solution equals the host-diagonalised one to the FP32 bases' rounding and the
    two-level solves give the same port impedances."""
    from scipy import sparse
    from pcbqs_amd.coarse import loop_restrict, stream_patterns, unique_columns, via_loop_columns
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c2(1, target=3000)
    s = PortSolver(b, device=0)
    m = s.model
    phi = unique_columns(sparse.hstack([loop_restrict(m, stream_patterns(m, 10), dense=False),
                                        via_loop_columns(m, dense=False)], format='csc'))
    freqs = [1e3, 1e6, 1e8]
    s.set_coarse(phi.toarray(), diagonalize='host')
    k0 = s.coarse_k
    x0 = {f: s.coarse_solution(f) for f in freqs}
    Z0, info0 = s.ports(freqs, precond='two-level', tol=1e-10)
    assert k0 == phi.shape[1]
    repeated = sparse.hstack([phi, -phi[:, [3]]], format='csc')
    for cols, path in ((phi, 'cholesky'), (repeated, 'eigenvectors')):
        s.set_coarse(cols)
        print(f'{cols.shape[1]} columns: {s.coarse_path}, kept {s.coarse_k}', flush=True)
        assert s.coarse_path == path and s.coarse_k == k0
        for f in freqs:
            x1 = s.coarse_solution(f)
            assert np.abs(x1 - x0[f]).max() <= 1e-5 * np.abs(x0[f]).max(), (path, f)
        Z1, info1 = s.ports(freqs, precond='two-level', tol=1e-10)
        assert np.abs(Z1 - Z0).max() <= 1e-9 * np.abs(Z0).max(), path
        assert all(abs(a['iterations'] - c['iterations']) <= max(3, 0.1 * a['iterations'])
                   for a, c in zip(info0, info1)), path
    s.close()


def test_block_jacobi_from_the_blocks_modes_equals_gauss_jordan(monkeypatch):
    """The block-Jacobi inverses formed from each block's generalised eigenvectors (one batched
    eigensolve per handle, then V diag(1 / (1 + j w lam)) V^T per frequency) and by the per-frequency
    Gauss-Jordan inversion (PCBQS_AMD_BJ_MODAL=0) are the same preconditioner: the solves take the
    same iterations to within two and give the same port impedances, at DC and up to 100 MHz."""
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c2(1, target=3000)
    freqs = [0.0, 1e3, 1e6, 1e8]
    out = {}
    for modal in ('1', '0'):
        monkeypatch.setenv('PCBQS_AMD_BJ_MODAL', modal)
        s = PortSolver(b, device=0)
        out[modal] = s.ports(freqs, precond='block-jacobi', tol=1e-10)
        s.close()
    (Z1, i1), (Z0, i0) = out['1'], out['0']
    print(f"iterations modal {[r['iterations'] for r in i1]} vs Gauss-Jordan {[r['iterations'] for r in i0]}",
          flush=True)
    assert all(abs(a['iterations'] - c['iterations']) <= 2 for a, c in zip(i1, i0))
    assert np.abs(Z1 - Z0).max() <= 1e-8 * np.abs(Z0).max()
