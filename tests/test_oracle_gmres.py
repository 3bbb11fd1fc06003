"""The CPU GMRES oracle and the C projection pinned: lockstep block-Jacobi GMRES reaches numpy's
dense solution of a seeded complex-symmetric loop system; with one block the preconditioner is the
exact inverse (one iteration per column); an iteration cap below convergence raises; cr_project
equals the reference's sparse products qform(C, L, C) and the dense C^T L C."""
import numpy as np
import pytest
from scipy import sparse

from oracle import fill as F
from oracle import gmres as G
from oracle import mqs as M


def system(n, k, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    L = (A @ A.T / n + 0.3 * np.eye(n)) * 1e-9
    R = sparse.random(n, n, density=6.0 / n, random_state=seed)
    R = R + R.T
    R = (R + sparse.diags(np.asarray(abs(R).sum(1)).ravel() + 1e-3)).tocsr() * 1e-3   # diagonally dominant
    B = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    return L, R, B


@pytest.mark.parametrize('omega', [0.0, 2 * np.pi * 1e6, 2 * np.pi * 1e8])
def test_lockstep_gmres_reaches_the_dense_solution(omega):
    L, R, B = system(600, 4, 1)
    X, its, res = G.plain_solve(L, R, B, omega, tol=1e-10, restart=30, block=64)
    Xe = np.linalg.solve(R.toarray() + 1j * omega * L, B)
    assert res <= 5e-10
    assert np.abs(X - Xe).max() <= 1e-7 * np.abs(Xe).max()
    assert its > 1


def test_one_block_is_the_exact_inverse():
    L, R, B = system(200, 3, 2)
    X, its, res = G.plain_solve(L, R, B, 2 * np.pi * 1e6, tol=1e-10, block=200)
    assert its == 1 and res <= 1e-12


def test_iteration_cap_raises():
    L, R, B = system(600, 2, 3)
    with pytest.raises(G.ConvergenceError):
        G.plain_solve(L, R, B, 2 * np.pi * 1e8, tol=1e-12, restart=5, maxiter=5, block=16)


def test_cr_project_equals_the_sparse_products():
    rng = np.random.default_rng(4)
    n, nl = 300, 120
    A = rng.standard_normal((n, n))
    L = A + A.T
    C = sparse.random(n, nl, density=0.05, random_state=5, format='csc')
    C.data = np.sign(C.data - 0.5)
    C.data[C.data == 0] = 1.0
    out = F.cr_project(C, L, workers=3)
    assert np.abs(out - M.qform(C, L, C)).max() <= 1e-12 * np.abs(out).max()
    assert np.abs(out - C.T.toarray() @ L @ C.toarray()).max() <= 1e-12 * np.abs(out).max()
    # a transposed basis or a dropped sign is caught
    C2 = C.copy()
    C2.data[0] *= -1
    assert np.abs(F.cr_project(C2, L, workers=3) - out).max() > 1e-6 * np.abs(out).max()
