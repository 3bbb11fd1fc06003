"""Oracle for the plain-preconditioned port GMRES (test infrastructure and the timed CPU baseline;
see oracle/__init__.py). Nothing in the product imports it.

Restates, in numpy:
  lockstep_gmres    src/pcbqs/mqs_iterative.py:109-170 (_lockstep_gmres: k right-hand sides at once,
                    right-preconditioned, modified Gram-Schmidt per column, complex Givens, columns
                    that converge drop out at the next restart)
  acceptance        src/pcbqs/peec_iterative.py:109-111 (true relative residual <= max(5 tol, 1e-13),
                    else ConvergenceError)
  block_jacobi      the device solve's plain preconditioner: exact inverses of the diagonal blocks of
                    256 consecutive free loops of Z_ff = Rl + j w Ll
with the loop operator applied as the reference applies a real inductance to complex currents: two
real products (src/pcbqs/mqs.py:669-678, _inductive_voltage).

Pinned by: tests/test_oracle_gmres.py (the solution of a seeded complex-symmetric system equals
numpy's dense solve; block-Jacobi with one block is the exact inverse, so one iteration suffices;
an iteration limit below convergence raises).
"""
import numpy as np


class ConvergenceError(RuntimeError):
    def __init__(self, residual, iterations):
        super().__init__(f'GMRES did not converge: relative residual {residual:.3g} after {iterations} iterations')
        self.relative_residual, self.iterations = residual, iterations


def loop_operator(L, Rl, omega):
    """X -> (Rl + j omega L) X for complex X (n, k): L real dense, Rl sparse or dense."""
    def op(X):
        re, im = np.ascontiguousarray(X.real), np.ascontiguousarray(X.imag)     # BLAS needs contiguous operands
        return Rl @ X + 1j * omega * (L @ re + 1j * (L @ im))
    return op


def block_jacobi(L, Rl, omega, block=256):
    """M: X -> blockdiag(Z_ff)^-1 X, blocks of `block` consecutive loops."""
    n = L.shape[0]
    Rd = Rl.toarray() if hasattr(Rl, 'toarray') else np.asarray(Rl)
    inv = []
    for b0 in range(0, n, block):
        b1 = min(n, b0 + block)
        inv.append((b0, b1, np.linalg.inv(Rd[b0:b1, b0:b1] + 1j * omega * L[b0:b1, b0:b1])))

    def M(X):
        Y = np.empty_like(X, dtype=complex)
        for b0, b1, A in inv:
            Y[b0:b1] = A @ X[b0:b1]
        return Y
    return M


def lockstep_gmres(op, M, B, X0, tol, maxiter, restart):
    """X (n, k), iterations per column, relative residuals; the algorithm of _lockstep_gmres."""
    n, k = B.shape
    X = np.array(X0, complex)
    bn = np.maximum(np.linalg.norm(B, axis=0), np.finfo(float).tiny)
    its = np.zeros(k, int)
    R = B - op(X)
    active = np.flatnonzero(np.linalg.norm(R, axis=0) / bn > tol)
    while len(active) and its[active].max() < maxiter:
        m = int(min(restart, maxiter - its[active].max()))
        r = R[:, active]
        beta = np.linalg.norm(r, axis=0)
        ka = len(active)
        V = np.zeros((m + 1, n, ka), complex)
        V[0] = r / beta
        H = np.zeros((m + 1, m, ka), complex)
        cs, sn = np.zeros((m, ka), complex), np.zeros((m, ka), complex)
        g = np.zeros((m + 1, ka), complex)
        g[0] = beta
        Z = np.zeros((m, n, ka), complex)
        done = np.zeros(ka, bool)
        last = np.full(ka, m - 1)
        for j in range(m):
            Z[j] = M(V[j])
            w = op(Z[j])
            for i in range(j + 1):
                h = np.einsum('nk,nk->k', V[i].conj(), w)
                H[i, j] = h
                w -= V[i] * h
            hn = np.linalg.norm(w, axis=0)
            H[j + 1, j] = hn
            V[j + 1] = w / np.where(hn > 0, hn, 1)
            for i in range(j):
                a, b = H[i, j].copy(), H[i + 1, j].copy()
                H[i, j] = cs[i].conj() * a + sn[i].conj() * b
                H[i + 1, j] = -sn[i] * a + cs[i] * b
            a, b = H[j, j].copy(), H[j + 1, j].copy()
            den = np.sqrt(abs(a) ** 2 + abs(b) ** 2)
            den = np.where(den > 0, den, 1)
            cs[j], sn[j] = a / den, b / den
            H[j, j] = cs[j].conj() * a + sn[j].conj() * b
            H[j + 1, j] = 0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j].conj() * g[j]
            newly = (~done) & (abs(g[j + 1]) / bn[active] <= tol)
            last[newly] = j
            done |= newly
            its[active[~done | newly]] += 1
            if done.all():
                break
        for c in range(ka):
            jj = last[c] + 1 if done[c] else j + 1
            y = np.linalg.solve(np.triu(H[:jj, :jj, c]), g[:jj, c]) if jj else np.zeros(0)
            X[:, active[c]] += np.einsum('j,jn->n', y, Z[:jj, :, c])
        R[:, active] = B[:, active] - op(X[:, active])
        res = np.linalg.norm(R[:, active], axis=0) / bn[active]
        active = active[res > tol]
    res = np.linalg.norm(R, axis=0) / bn
    return X, its, res


def plain_solve(L, Rl, B, omega, tol=1e-8, restart=100, maxiter=3000, block=256):
    """The free loops' solution X of (Rl + j w L) X = B (n, k) by block-Jacobi lockstep GMRES, with
    the acceptance rule; returns (X, iterations (the most of any column), worst relative residual)."""
    op = loop_operator(L, Rl, omega)
    M = block_jacobi(L, Rl, omega, block)
    X, its, res = lockstep_gmres(op, M, B, np.zeros_like(B, dtype=complex), tol, maxiter, restart)
    true = np.linalg.norm(B - op(X), axis=0) / np.maximum(np.linalg.norm(B, axis=0), np.finfo(float).tiny)
    worst = float(true.max()) if len(true) else 0.0
    if not np.isfinite(worst) or worst > max(5 * tol, 1e-13):
        raise ConvergenceError(worst, int(its.max()))
    return X, int(its.max()), worst
