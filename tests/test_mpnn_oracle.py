"""The message-passing network's CPU oracle: its BF16 rounding and its gradients.

The backward is pinned by central differences of the FP64 network (no rounding), the rounding by
known bit patterns."""
import numpy as np
from scipy import sparse

from oracle import mpnn as M


def random_graph(n, seed=0):
    rng = np.random.default_rng(seed)
    i = rng.integers(0, n, 4 * n)
    j = rng.integers(0, n, 4 * n)
    adj = sparse.csr_matrix((np.ones(len(i)), (i, j)), shape=(n, n))
    adj = ((adj + adj.T) > 0).astype(float)
    adj.setdiag(0)
    adj.eliminate_zeros()
    deg = np.asarray(adj.sum(1)).ravel()
    return (sparse.diags(1 / np.maximum(deg, 1)) @ adj).tocsr()


def test_bf16_rounding_known_values():
    one = 1.0
    assert M.bf16(np.array([one]))[0] == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert M.bf16(np.array([1 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even gives 1 + 2^-6
    assert M.bf16(np.array([1 + 3 * 2.0 ** -8]))[0] == 1 + 2.0 ** -6
    # just above the first tie rounds up
    assert M.bf16(np.array([1 + 2.0 ** -8 + 2.0 ** -20]))[0] == 1 + 2.0 ** -7
    x = np.random.default_rng(1).normal(size=1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-6))
    # idempotent, and the low 16 bits of the FP32 pattern are zero
    assert np.array_equal(M.bf16(r), r)
    assert np.all(r.astype(np.float32).view(np.uint32) & 0xFFFF == 0)


def test_gradients_match_central_differences():
    n, n_in, n_out = 40, 16, 16
    A = random_graph(n)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(n, n_in))
    p = M.init_params(n_in, n_out, seed=3)
    c = rng.normal(size=(n, n_out))
    y, cache = M.forward(p, X, A, bf16_on=False)
    g = M.backward(p, cache, A, c, bf16_on=False)

    def loss(q):
        return float(np.sum(c * M.forward(q, X, A, bf16_on=False)[0]))
    eps = 1e-6
    for name in ('W0', 'b0', 'W1', 'b2', 'W4', 'b4', 'Wo', 'bo'):
        flat = p[name].ravel()
        for idx in rng.choice(flat.size, 5, replace=False):
            q = {k: v.copy() for k, v in p.items()}
            q[name].ravel()[idx] += eps
            up = loss(q)
            q[name].ravel()[idx] -= 2 * eps
            dn = loss(q)
            fd = (up - dn) / (2 * eps)
            an = g[name].ravel()[idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, an)


def test_bf16_forward_close_to_fp64():
    n = 50
    A = random_graph(n, 4)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(n, 16))
    p = M.init_params(16, 16, seed=6)
    y64, _ = M.forward(p, X, A, bf16_on=False)
    y16, _ = M.forward(p, X, A, bf16_on=True)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < rel < 5e-2          # BF16 operands: visibly rounded, still the same network
