"""The learned coarse-space correction of the loop solve.

A message-passing network over the branch graph emits, once per board, K = 4 channels per branch.
Each channel restricted to the Voronoi cell of every via and port (the cell of a branch is its
nearest via or port) is one enrichment pattern; on the free loops' cotree branches the patterns are
loop-space columns that join the stream-function hats of the two-level preconditioner. The network
is trained on solver-generated pairs: for converged port solutions x* (two-level GMRES to 1e-10 at
several frequencies) the loss is the fraction of x* - Pi_hats x* that the learned columns leave
uncaptured. Block-Jacobi GMRES from zero is the fallback when a two-level solve fails its acceptance
test (PortSolver.ports retries it and names the fallback in its info), and the true residual of
every solve judges the network. No counterpart in the reference (its iterative path uses a
graph-Laplacian preconditioner, src/pcbqs/mqs_precond.py:27-59).

Graph: one node per branch of the loop model (RWG edges, via segments, port sources). Two RWG edges
are neighbours when they share a triangle; a via or port branch neighbours the RWG edges at its two
end nodes. Aggregation is the neighbour mean.

Network (hidden 64, four layers): h = relu(X W0 + b0); four times a = mean_N(h),
h <- h + relu([h | a] W_l + b_l); y = h Wo + bo. The dense products run in hand-written BF16 MFMA
kernels with FP32 accumulation (forward and backward, libpcbqs_hip.so); torch only holds the
parameters and drives autograd, the projection loss and the optimiser.

Node features (16): position, layer, kind (via, port), edge length and normal direction, boundary
flag, distances to the nearest boundary vertex, via and port, tree membership, and the offset to
the branch's cell centre (two signed log distances) and the cell's kind.
"""
import ctypes
import os
import time

import numpy as np
from scipy import sparse
from scipy.spatial import cKDTree

from . import _lib
from ._lib import call

HIDDEN, LAYERS = 64, 4
N_IN, N_OUT, K_CHANNELS = 16, 16, 4
FREQS = (1e3, 1e5, 1e6, 1e7, 1e8)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'mpnn_c2.npz')


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('the learned correction runs on the GPU: no HIP device is visible')
    return torch


# ---- graph and features ---------------------------------------------------------------------

class BranchGraph:
    """Adjacency (mean weights) and static features of the branches of a LoopModel."""

    def __init__(self, model):
        b, r = model.board, model.rwg
        ne, nv, npo, nb = model.ne, model.nv, model.np, model.nb
        self.nb = nb
        rows = np.arange(ne)
        T = sparse.csr_matrix((np.ones(2 * ne), (np.r_[rows, rows], np.r_[r.tp, r.tm])), shape=(nb, b.nt))
        adj = (T @ T.T).tocsr()
        node_of = b.node_of
        rwg_nodes = sparse.csr_matrix((np.ones(2 * ne), (np.r_[rows, rows], np.r_[node_of[r.tp], node_of[r.tm]])),
                                      shape=(nb, b.n_nodes))
        special = np.arange(ne, nb)
        ends = model.ends[ne:]
        S = sparse.csr_matrix((np.ones(2 * len(special)), (np.r_[special, special], np.r_[ends[:, 0], ends[:, 1]])),
                              shape=(nb, b.n_nodes))
        link = (S @ rwg_nodes.T).tocsr()
        adj = adj + link + link.T
        adj.setdiag(0)
        adj.eliminate_zeros()
        adj.data[:] = 1.0
        deg = np.asarray(adj.sum(1)).ravel()
        A = sparse.diags(np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)) @ adj
        self.A = A.tocsr()
        self.A.sort_indices()
        self.AT = self.A.T.tocsr()
        self.AT.sort_indices()
        self.static, self.cell, self.n_cells = self._features(model, adj)
        self.cot = np.asarray(model.basis.cotree[:model.n_free])

    @staticmethod
    def _features(model, adj):
        b, r = model.board, model.rwg
        ne, nv, nb = model.ne, model.nv, model.nb
        cen = b.nodes[b.tris].mean(1)
        pos = np.zeros((nb, 2))
        pos[:ne] = (cen[r.tp] + cen[r.tm]) / 2
        for k, (x, y, _, _, _) in enumerate(b.vseg):
            pos[ne + k] = (x, y)
        z = np.zeros(nb)
        z[:ne] = (b.z_of[r.tp] + b.z_of[r.tm]) / 2
        for k, (_, _, _, z0, z1) in enumerate(b.vseg):
            z[ne + k] = (z0 + z1) / 2
        # a port branch sits at the mean position of its RWG neighbours
        deg = np.asarray(adj[ne + nv:, :ne].sum(1)).ravel()
        if nb > ne + nv:
            s = adj[ne + nv:, :ne] @ np.c_[pos[:ne], z[:ne]]
            s = s / np.maximum(deg, 1)[:, None]
            pos[ne + nv:], z[ne + nv:] = s[:, :2], s[:, 2]
        lo, hi = b.nodes.min(0), b.nodes.max(0)
        ext = np.maximum(hi - lo, 1e-12)
        h = float(r.l.mean()) if ne else 1.0
        F = np.zeros((nb, N_IN))
        F[:, 0:2] = (pos - lo) / ext * 2 - 1
        zl, zh = z.min(), z.max()
        F[:, 2] = (z - zl) / max(zh - zl, 1e-12) * 2 - 1
        F[ne:ne + nv, 3] = 1.0
        F[ne + nv:, 4] = 1.0
        F[:ne, 5] = np.log(r.l / h)
        d = cen[r.tm] - cen[r.tp]
        F[:ne, 6:8] = d / np.maximum(np.linalg.norm(d, axis=1), 1e-30)[:, None]
        from .coarse import boundary_vertices
        bnd = boundary_vertices(b.tris, len(b.nodes))
        F[:ne, 8] = (bnd[r.ends[:, 0]] | bnd[r.ends[:, 1]]).astype(float)

        def dist(points):
            if len(points) == 0:
                return np.full(nb, 10.0)
            dd, _ = cKDTree(points).query(pos)
            return np.log1p(dd / h)
        F[:, 9] = dist(b.nodes[bnd])
        F[:, 10] = dist(np.array([(x, y) for x, y, _, _, _ in b.vseg]).reshape(-1, 2))
        F[:, 11] = dist(pos[ne + nv:])
        tree = np.ones(nb)
        tree[model.basis.cotree] = 0.0
        F[:, 12] = tree
        # cells: the nearest via or port centre of every branch
        centres = pos[ne:]
        kinds = np.r_[np.zeros(nv), np.ones(nb - ne - nv)]
        if len(centres):
            _, cell = cKDTree(centres).query(pos)
        else:
            cell = np.zeros(nb, np.int64)
        off = pos - centres[cell] if len(centres) else np.zeros((nb, 2))
        F[:, 13:15] = np.sign(off) * np.log1p(np.abs(off) / h)
        F[:, 15] = kinds[cell] if len(centres) else 0.0
        return F.astype(np.float32), np.asarray(cell, np.int64), max(len(centres), 1)

    def batched(self, P, torch, device):
        """CSR of P disjoint copies (A and its transpose) as device tensors."""
        out = []
        for M in (self.A, self.AT):
            big = sparse.block_diag([M] * P, format='csr')
            big.sort_indices()
            out.append(tuple(torch.as_tensor(np.ascontiguousarray(a), device=device)
                             for a in (big.indptr.astype(np.int32), big.indices.astype(np.int32),
                                       big.data.astype(np.float32))))
        return out


# ---- the network on the device kernels -------------------------------------------------------

def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _gemm(torch, N, A1, k1, A2, k2, W, m, bias, res, relu, C, Zpre=None, lda1=None, lda2=None, ldc=None):
    call('pq_mpnn_gemm', N, _p(A1), lda1 or k1, k1, _p(A2), lda2 or max(k2, 1), k2, _p(W), m, _p(bias), _p(res),
         m, int(relu), _p(C), ldc or m, _p(Zpre), _stream(torch))


def _gemm_tn(torch, N, A1, k1, A2, k2, G, m, dW, db):
    need = ctypes.c_size_t()
    call('pq_mpnn_workspace', N, k1 + k2, m, ctypes.byref(need))
    work = torch.empty(need.value, dtype=torch.float32, device=G.device)
    call('pq_mpnn_gemm_tn', N, _p(A1), k1, k1, _p(A2), max(k2, 1), k2, _p(G), m, m, _p(work), need.value,
         _p(dW), _p(db), _stream(torch))


def _aggregate(torch, N, csr, inp, ldi, out, base=None, base2=None, ldb2=64, in_offset=0):
    ptr, col, w = csr
    pin = ctypes.c_void_p(inp.data_ptr() + 4 * in_offset)
    call('pq_mpnn_aggregate', N, _p(ptr), _p(col), _p(w), pin, ldi, _p(base), 64, _p(base2), ldb2, _p(out), 64,
         _stream(torch))


def _forward(torch, X, A, params, keep):
    """y (N, 16) and, with keep, the activations the backward needs."""
    N = X.shape[0]
    dev = X.device
    f = dict(dtype=torch.float32, device=dev)
    W0, b0 = params[0], params[1]
    h = torch.empty((N, HIDDEN), **f)
    z0 = torch.empty((N, HIDDEN), **f) if keep else None
    _gemm(torch, N, X, N_IN, None, 0, W0, HIDDEN, b0, None, 1, h, z0)
    hs, as_, zs = [], [], []
    for l in range(LAYERS):
        W, b = params[2 + 2 * l], params[3 + 2 * l]
        a = torch.empty((N, HIDDEN), **f)
        _aggregate(torch, N, A, h, HIDDEN, a)
        z = torch.empty((N, HIDDEN), **f) if keep else None
        hn = torch.empty((N, HIDDEN), **f)
        _gemm(torch, N, h, HIDDEN, a, HIDDEN, W, HIDDEN, b, h, 1, hn, z)
        if keep:
            hs.append(h); as_.append(a); zs.append(z)
        h = hn
    Wo, bo = params[2 + 2 * LAYERS], params[3 + 2 * LAYERS]
    y = torch.empty((N, N_OUT), **f)
    _gemm(torch, N, h, HIDDEN, None, 0, Wo, N_OUT, bo, None, 0, y)
    return y, (z0, hs, as_, zs, h)


def make_function(torch):
    class MPNNFunction(torch.autograd.Function):
        """y = MPNN(X) with forward and backward on the device kernels."""

        @staticmethod
        def forward(ctx, X, A, AT, *params):
            y, (z0, hs, as_, zs, hout) = _forward(torch, X, A, params, keep=True)
            ctx.AT = AT
            ctx.save_for_backward(X, z0, hout, *hs, *as_, *zs, *params)
            return y

        @staticmethod
        def backward(ctx, dy):
            saved = ctx.saved_tensors
            X, z0, hout = saved[0], saved[1], saved[2]
            hs, as_, zs = saved[3:3 + LAYERS], saved[3 + LAYERS:3 + 2 * LAYERS], saved[3 + 2 * LAYERS:3 + 3 * LAYERS]
            params = saved[3 + 3 * LAYERS:]
            N = X.shape[0]
            f = dict(dtype=torch.float32, device=X.device)
            dy = dy.contiguous()
            grads = [torch.empty_like(p) for p in params]
            Wo = params[2 + 2 * LAYERS]
            _gemm_tn(torch, N, hout, HIDDEN, None, 0, dy, N_OUT, grads[2 + 2 * LAYERS], grads[3 + 2 * LAYERS])
            dh = torch.empty((N, HIDDEN), **f)
            _gemm(torch, N, dy, N_OUT, None, 0, Wo.t().contiguous(), HIDDEN, None, None, 0, dh)
            for l in range(LAYERS - 1, -1, -1):
                W = params[2 + 2 * l]
                dz = torch.empty((N, HIDDEN), **f)
                call('pq_mpnn_relu_mask', dz.numel(), _p(zs[l]), _p(dh), _p(dz), _stream(torch))
                _gemm_tn(torch, N, hs[l], HIDDEN, as_[l], HIDDEN, dz, HIDDEN, grads[2 + 2 * l], grads[3 + 2 * l])
                D = torch.empty((N, 2 * HIDDEN), **f)
                _gemm(torch, N, dz, HIDDEN, None, 0, W.t().contiguous(), 2 * HIDDEN, None, None, 0, D)
                dhn = torch.empty((N, HIDDEN), **f)
                _aggregate(torch, N, ctx.AT, D, 2 * HIDDEN, dhn, base=dh, base2=D, ldb2=2 * HIDDEN, in_offset=HIDDEN)
                dh = dhn
            dz0 = torch.empty((N, HIDDEN), **f)
            call('pq_mpnn_relu_mask', dz0.numel(), _p(z0), _p(dh), _p(dz0), _stream(torch))
            _gemm_tn(torch, N, X, N_IN, None, 0, dz0, HIDDEN, grads[0], grads[1])
            return (None, None, None, *grads)
    return MPNNFunction


class MPNN:
    """Parameters (FP32 device tensors) and the device forward/backward of the network."""

    def __init__(self, device=0, seed=0, params=None):
        torch = _torch()
        self.torch = torch
        self.device = torch.device('cuda', device)
        if params is None:
            params = init_params(seed)
        names = param_names()
        self.params = [torch.tensor(np.asarray(params[k], np.float32), device=self.device).requires_grad_(True)
                       for k in names]
        self._fn = make_function(torch)

    def __call__(self, X, A, AT):
        return self._fn.apply(X, A, AT, *self.params)

    def infer(self, X, A):
        with self.torch.no_grad():
            y, _ = _forward(self.torch, X, A, [p.detach() for p in self.params], keep=False)
        return y

    def state(self):
        return {k: p.detach().cpu().numpy() for k, p in zip(param_names(), self.params)}

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, **self.state(), **{f'meta_{k}': np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS, device=0):
        with np.load(path, allow_pickle=False) as z:
            params = {k: z[k] for k in param_names()}
        return cls(device, params=params)


def param_names():
    return ['W0', 'b0'] + [f'{s}{l}' for l in range(1, LAYERS + 1) for s in ('W', 'b')] + ['Wo', 'bo']


def init_params(seed=0):
    """He-style random initialisation (a zero output layer would sit on a saddle of the projection
    loss: no columns, no gradient)."""
    rng = np.random.default_rng(seed)
    p = {'W0': rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, HIDDEN)), 'b0': np.zeros(HIDDEN)}
    for l in range(1, LAYERS + 1):
        p[f'W{l}'] = rng.normal(0, 0.5 * np.sqrt(2.0 / (2 * HIDDEN)), (2 * HIDDEN, HIDDEN))
        p[f'b{l}'] = np.zeros(HIDDEN)
    p['Wo'] = rng.normal(0, np.sqrt(1.0 / HIDDEN), (HIDDEN, N_OUT))
    p['bo'] = np.zeros(N_OUT)
    return p


# ---- learned columns, solver-generated pairs, training ----------------------------------------

def learned_columns(torch, graph, E, K=K_CHANNELS):
    """(n_free, n_cells * K) loop-space columns from the network output E (nb, >= K) (torch,
    differentiable): row l holds E[cot_l, :K] in the columns of its cell."""
    cot = torch.as_tensor(graph.cot, device=E.device)
    cell = torch.as_tensor(graph.cell[graph.cot], device=E.device)
    n = len(graph.cot)
    Phi = torch.zeros((n, graph.n_cells * K), dtype=E.dtype, device=E.device)
    rows = torch.arange(n, device=E.device)[:, None].expand(n, K)
    cols = cell[:, None] * K + torch.arange(K, device=E.device)[None, :]
    return Phi.index_put((rows, cols), E[cot, :K])


class BoardData:
    """One board's solver with the hat coarse space, its branch graph, and the converged port
    solutions' residuals after projection onto the hats (the training pairs)."""

    def __init__(self, board, grid=24, device=0, freqs=FREQS, tol=1e-10, keep_solver=True):
        from .solve import PortSolver
        torch = _torch()
        self.rwg = None
        self.solver = PortSolver(board, device=device)
        self.solver.set_stream_coarse(grid)
        phi = self.solver.coarse_phi
        self.phi_geo = phi.toarray() if sparse.issparse(phi) else phi
        self.graph = BranchGraph(self.solver.model)
        dev = torch.device('cuda', device)
        self.X = torch.as_tensor(self.graph.static, device=dev)
        self.A, self.AT = self.graph.batched(1, torch, dev)
        self.freqs = list(freqs)
        if freqs:
            _, self.info, X = self.solver.ports(self.freqs, precond='two-level', tol=tol, currents=True)
            sols = np.concatenate([X[i] for i in range(len(self.freqs))], axis=1)      # (n_free, 2 P F)
            Q, _ = torch.linalg.qr(torch.as_tensor(self.phi_geo, device=dev))
            R0 = torch.as_tensor(sols, device=dev)
            self.Q = Q
            self.R = R0 - Q @ (Q.T @ R0)                                                 # residual after hats
            self.R0norm = float((R0 ** 2).sum())
        self.rwg = int(self.solver.model.ne)
        if not keep_solver:                   # training needs only the graph and the pairs
            self.solver.close()
            self.solver = None

    def close(self):
        if self.solver is not None:
            self.solver.close()


def projection_loss(torch, bd, E):
    """Fraction of the hat residual R left uncaptured by the span of the learned columns."""
    Phi = learned_columns(torch, bd.graph, E).double()
    P = Phi - bd.Q @ (bd.Q.T @ Phi)
    G = P.T @ P
    eps = 1e-8 * torch.clamp(torch.diagonal(G).mean(), min=1e-30)
    G = G + eps * torch.eye(G.shape[0], dtype=G.dtype, device=G.device)
    C = P.T @ bd.R
    captured = (C * torch.linalg.solve(G, C)).sum()
    total = (bd.R ** 2).sum()
    return 1.0 - captured / total


def train(boards, steps=400, lr=2e-3, device=0, seed=0, log=print):
    """Train on solver pairs: boards is a list of BoardData. Returns (net, history)."""
    torch = _torch()
    net = MPNN(device, seed)
    opt = torch.optim.Adam(net.params, lr=lr)
    rng = np.random.default_rng(seed)
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        bd = boards[rng.integers(len(boards))]
        E = net(bd.X, bd.A, bd.AT)
        loss = projection_loss(torch, bd, E)
        opt.zero_grad()
        loss.backward()
        opt.step()
        history.append(float(loss.detach()))
        if log and (step % 50 == 0 or step == steps - 1):
            log(f'step {step} uncaptured {np.mean(history[-50:]):.4f} ({time.perf_counter() - t0:.1f} s)')
    return net, history


def enrichment(net, graph, dense=False):
    """The learned columns (n_free, n_cells * K) of a board, each scaled to unit norm (zero
    columns dropped): a scipy CSC matrix (row l holds the network's output at l's cotree branch in
    the K columns of its cell, the columns of learned_columns), or a dense array with dense=True."""
    torch = net.torch
    X = torch.as_tensor(graph.static, device=net.device)
    A, _ = graph.batched(1, torch, net.device)
    E = net.infer(X, A)
    K = K_CHANNELS
    vals = E[torch.as_tensor(graph.cot, device=E.device), :K].double().cpu().numpy()
    n = len(graph.cot)
    cols = (graph.cell[graph.cot][:, None] * K + np.arange(K)[None, :]).ravel()
    m = sparse.csc_matrix((vals.ravel(), (np.repeat(np.arange(n), K), cols)), shape=(n, graph.n_cells * K))
    m.eliminate_zeros()
    norm = np.sqrt(np.asarray(m.multiply(m).sum(0)).ravel())
    keep = np.flatnonzero(norm > 0)
    m = m[:, keep]
    m.data = m.data / np.repeat(norm[keep], np.diff(m.indptr))
    m.sort_indices()
    return np.ascontiguousarray(m.toarray()) if dense else m


def set_learned_coarse(solver, net, grid=36, graph=None):
    """The two-level coarse space of a solver: the grid x grid hats, the via loops, and the
    network's enrichment. Returns the seconds the network's part took. The default grid is finer
    than the 24 x 24 hats the shipped weights were trained against: at a 36 x 36 grid the enrichment
    still removes a quarter of the iterations, and weights retrained against 36 x 36 hats did no
    better at 1 MHz on the held-out board (86 iterations either way)."""
    t = time.perf_counter()
    graph = graph or BranchGraph(solver.model)
    extra = enrichment(net, graph)
    seconds = time.perf_counter() - t
    solver.set_stream_coarse(grid, extra_loops=extra)
    return seconds


def evaluate(net, bd):
    """Uncaptured fraction of the hat residual with the learned columns, and its size without them."""
    torch = net.torch
    with torch.no_grad():
        E = net.infer(bd.X, bd.A)
        loss = float(projection_loss(torch, bd, E))
    return dict(hat_residual=float((bd.R ** 2).sum()) / bd.R0norm, uncaptured=loss)


def compare(solver, net, freqs, grid=24, tol=1e-8):
    """Two-level GMRES (iterations, seconds) with the hats alone and with the learned enrichment."""
    rows = []
    solver.set_stream_coarse(grid)
    k0 = solver.coarse_k
    base = [solver.ports([f], precond='two-level', tol=tol)[1][0] for f in freqs]
    set_learned_coarse(solver, net, grid)
    k1 = solver.coarse_k
    learned = [solver.ports([f], precond='two-level', tol=tol)[1][0] for f in freqs]
    for f, a, b in zip(freqs, base, learned):
        rows.append(dict(frequency_Hz=float(f), hats=(a['iterations'], round(a['seconds'], 4), k0),
                         learned=(b['iterations'], round(b['seconds'], 4), k1)))
    return rows


def held_out(net, boards, grid, freq=1e6, tol=1e-8):
    """Two-level GMRES iterations at freq on held-out boards: (hats alone, hats + the network's
    enrichment) per board."""
    from .solve import PortSolver
    rows = []
    for b in boards:
        s = PortSolver(b)
        s.set_stream_coarse(grid)
        hats = s.ports([freq], precond='two-level', tol=tol)[1][0]['iterations']
        set_learned_coarse(s, net, grid)
        learned = s.ports([freq], precond='two-level', tol=tol)[1][0]['iterations']
        rows.append(dict(rwg=int(s.model.ne), hats=int(hats), learned=int(learned)))
        s.close()
    return rows


def main(argv=None):
    import argparse
    import json
    from . import board as boards
    ap = argparse.ArgumentParser(description='train the learned coarse-space correction on solver pairs')
    ap.add_argument('--boards', type=int, default=12, help='training boards: seeds 1..N of the plane-pair generator')
    ap.add_argument('--target', type=int, default=20000, help='triangles per board')
    ap.add_argument('--sizes', default=None, metavar='LO:HI',
                    help='triangles per board drawn uniformly from LO..HI (seeded per board) instead of --target')
    ap.add_argument('--steps', type=int, default=400)
    ap.add_argument('--grid', type=int, default=24)
    ap.add_argument('--curve', default=None, metavar='N1,N2,...',
                    help='also train on the first N1, N2, ... boards and report held-out iterations for each')
    ap.add_argument('--held-out', type=int, default=3, help='held-out boards (seeds 5000..) besides seed 0')
    ap.add_argument('--out', default=WEIGHTS)
    args = ap.parse_args(argv)
    _torch()
    lo_hi = tuple(int(v) for v in args.sizes.split(':')) if args.sizes else None

    def make(seed):
        if lo_hi is None:
            return boards.board_c2(seed, target=args.target)
        return boards.board_c2(seed, target=int(np.random.default_rng(seed).integers(lo_hi[0], lo_hi[1] + 1)))
    data = []
    t = time.perf_counter()
    for s in range(1, args.boards + 1):
        data.append(BoardData(make(s), grid=args.grid, keep_solver=False))
        print(f'board {s}: {data[-1].rwg} RWG, {data[-1].graph.nb} branches, hat residual '
              f'{float((data[-1].R ** 2).sum()) / data[-1].R0norm:.3f} ({time.perf_counter() - t:.1f} s)', flush=True)
    held = [boards.board_c2(0)] + [make(5000 + i) for i in range(args.held_out)]
    counts = sorted({int(c) for c in args.curve.split(',')} | {args.boards}) if args.curve else [args.boards]
    curve = []
    for count in counts:
        t = time.perf_counter()
        net, hist = train(data[:count], steps=args.steps)
        train_s = time.perf_counter() - t
        rows = held_out(net, held, args.grid)
        curve.append(dict(boards=count, steps=args.steps, train_s=train_s, final_uncaptured=float(np.mean(hist[-50:])),
                          held_out=rows))
        print('curve', json.dumps(curve[-1]), flush=True)
        if count == args.boards:
            net.save(args.out, boards=args.boards, target=args.target if lo_hi is None else -1, steps=args.steps,
                     grid=args.grid, final_uncaptured=float(np.mean(hist[-50:])))
    for bd in data:
        bd.close()
    print(json.dumps(dict(learning_curve=curve)), flush=True)


if __name__ == '__main__':
    main()
