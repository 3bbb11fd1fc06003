"""Device kernels of the message-passing network against the CPU oracle (BF16 operands, FP32
accumulation on the matrix cores; the oracle rounds at the same points and sums in FP64)."""
import ctypes

import numpy as np
import pytest
from scipy import sparse

from oracle import mpnn as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32), device='cuda')


def graph(n, seed=0):
    rng = np.random.default_rng(seed)
    i, j = rng.integers(0, n, 5 * n), rng.integers(0, n, 5 * n)
    adj = sparse.csr_matrix((np.ones(len(i)), (i, j)), shape=(n, n))
    adj = ((adj + adj.T) > 0).astype(float)
    adj.setdiag(0)
    adj.eliminate_zeros()
    deg = np.asarray(adj.sum(1)).ravel()
    A = (sparse.diags(1 / np.maximum(deg, 1)) @ adj).tocsr()
    A.sort_indices()
    return A


def csr_dev(torch, A):
    return (torch.as_tensor(A.indptr.astype(np.int32), device='cuda'),
            torch.as_tensor(A.indices.astype(np.int32), device='cuda'),
            torch.as_tensor(A.data.astype(np.float32), device='cuda'))


@pytest.mark.parametrize('k1,k2,m', [(16, 0, 64), (64, 64, 64), (64, 0, 128), (64, 0, 16)])
@pytest.mark.parametrize('N', [1, 77, 3001])
def test_gemm_matches_bf16_oracle(torch, k1, k2, m, N):
    from pcbqs_amd.learned import _gemm
    rng = np.random.default_rng(N + k1 + m)
    A1, A2 = rng.normal(size=(N, k1)), rng.normal(size=(N, max(k2, 1)))
    W, b, res = rng.normal(size=(k1 + k2, m)), rng.normal(size=m), rng.normal(size=(N, m))
    A1f, A2f = A1.astype(np.float32).astype(np.float64), A2.astype(np.float32).astype(np.float64)
    Af = np.hstack([A1f, A2f[:, :k2]])
    z = M.bf16(Af) @ M.bf16(W.astype(np.float32)) + b.astype(np.float32)
    want = np.maximum(z, 0) + res.astype(np.float32)
    C, Z = torch.empty((N, m), device='cuda'), torch.empty((N, m), device='cuda')
    _gemm(torch, N, dev(torch, A1), k1, dev(torch, A2) if k2 else None, k2, dev(torch, W), m, dev(torch, b),
          dev(torch, res), 1, C, Z)
    torch.cuda.synchronize()
    assert rel(Z.cpu().numpy(), z) < 1e-6
    assert rel(C.cpu().numpy(), want) < 1e-6


@pytest.mark.parametrize('k1,k2,m', [(16, 0, 64), (64, 64, 64), (64, 0, 16)])
@pytest.mark.parametrize('N', [5, 1000, 70001])
def test_weight_gradient_matches_bf16_oracle(torch, k1, k2, m, N):
    from pcbqs_amd.learned import _gemm_tn
    rng = np.random.default_rng(N + m)
    A1, A2, G = rng.normal(size=(N, k1)), rng.normal(size=(N, max(k2, 1))), rng.normal(size=(N, m))
    Af = np.hstack([A1, A2[:, :k2]]).astype(np.float32).astype(np.float64)
    Gf = G.astype(np.float32).astype(np.float64)
    want_w, want_b = M.bf16(Af).T @ M.bf16(Gf), Gf.sum(0)
    dW, db = torch.empty((k1 + k2, m), device='cuda'), torch.empty(m, device='cuda')
    _gemm_tn(torch, N, dev(torch, A1), k1, dev(torch, A2) if k2 else None, k2, dev(torch, G), m, dW, db)
    torch.cuda.synchronize()
    assert rel(dW.cpu().numpy(), want_w) < 1e-5
    assert rel(db.cpu().numpy(), want_b) < 1e-5


def test_aggregate_matches_sparse_product(torch):
    from pcbqs_amd.learned import _aggregate
    A = graph(2500, 3)
    rng = np.random.default_rng(4)
    H, B1, B2 = rng.normal(size=(2500, 128)), rng.normal(size=(2500, 64)), rng.normal(size=(2500, 128))
    out = torch.empty((2500, 64), device='cuda')
    _aggregate(torch, 2500, csr_dev(torch, A), dev(torch, H), 128, out, base=dev(torch, B1), base2=dev(torch, B2),
               ldb2=128, in_offset=64)
    torch.cuda.synchronize()
    Hf = H.astype(np.float32).astype(np.float64)
    want = A.astype(np.float32).astype(np.float64) @ Hf[:, 64:] + B1.astype(np.float32) + B2.astype(np.float32)[:, :64]
    assert rel(out.cpu().numpy(), want) < 1e-6


def test_network_forward_and_gradients_match_oracle(torch):
    """The autograd wrapper over the device kernels against the oracle's forward and backward."""
    from pcbqs_amd.learned import MPNN, N_IN, N_OUT, param_names
    n = 3000
    A = graph(n, 7)
    rng = np.random.default_rng(8)
    X = rng.normal(size=(n, N_IN)).astype(np.float32).astype(np.float64)
    p = M.init_params(N_IN, N_OUT, seed=9)
    p = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
    c = rng.normal(size=(n, N_OUT)).astype(np.float32).astype(np.float64)
    Af = A.copy()
    Af.data = Af.data.astype(np.float32).astype(np.float64)
    y_want, cache = M.forward(p, X, Af, bf16_on=True)
    g_want = M.backward(p, cache, Af, c, bf16_on=True)
    net = MPNN(0, params=p)
    y = net(dev(torch, X), csr_dev(torch, A), csr_dev(torch, A.T.tocsr()))
    (y * dev(torch, c)).sum().backward()
    torch.cuda.synchronize()
    # FP32 accumulation order differs from the oracle's FP64 sums, which now and then moves a value
    # across a BF16 rounding boundary (2^-8 relative): the bounds allow that, not a wrong index
    assert rel(y.detach().cpu().numpy(), y_want) < 1e-3
    for name, t in zip(param_names(), net.params):
        assert rel(t.grad.cpu().numpy(), g_want[name]) < 5e-3, name
    # and the inference path gives the same output
    y2 = net.infer(dev(torch, X), csr_dev(torch, A))
    assert rel(y2.cpu().numpy(), y.detach().cpu().numpy()) < 1e-7
