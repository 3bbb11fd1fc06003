"""Device-resident port solve of one board: fill -> RWG combine -> loop projection -> GMRES.

The pipeline keeps everything in HBM: the moment planes, the branch inductance Lb (RWG edges and
via segments), the loop operator Ll = C^T Lb C and Rl = C^T R C, and the Krylov basis. precision
32 (the default) is the FP32-bulk pipeline: FP32 far pairs with FP64 analytic near pairs in local
frames, Lb stored in FP32. precision 64 is the plain FP64 method: FP64 planes in the reference's
layout and Lb in FP64. Either way Ll is formed and kept in FP64 (the system every residual, right-
hand side and port reduction uses) with a copy rounded to FP32 for the Krylov products, the
preconditioners and the FP32 factor. Port impedances follow the
reference's copper-port reduction (src/pcbqs/mqs_ports.py:35-86): Zp = Z_ss + Z_sf X with
X = -Z_ff^-1 Z_fs, here by lockstep GMRES (src/pcbqs/mqs_iterative.py:109-170, 298-367);
R = Re Zp and L = Im Zp / w (rl_equivalent, src/pcbqs/mqs_ports.py:72-86).
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr
from .backend import Mi355xBackend, padded
from .geometry import MU0, RWG, FillGeometry
from .loops import LoopModel

PRECONDITIONERS = {'none': 0, 'block-jacobi': 1, 'two-level': 2}
# The direct solve's refinement: relative residual and correction count of the reference's MixedLU
# (src/pcbqs/kernels.py:137: TOL, STEPS = 1e-12, 12).
DIRECT_TOL, DIRECT_STEPS = 1e-12, 12
# solver='auto' takes the direct solve up to this many free loops, and while a factorisation costs
# less than this many operator products per port and frequency (src/pcbqs/mqs_iterative.py:30-31).
DIRECT_LOOPS, PRODUCTS_PER_SOLVE = 8000, 60


def port_method(n_free, ports, n_freq):
    """'direct' or 'iterative' for solver='auto', the rule of the reference's port_method
    (src/pcbqs/mqs_iterative.py:260-272): direct while the free loops number at most
    min(8000, 60 x ports x frequencies)."""
    return 'direct' if n_free <= min(DIRECT_LOOPS, PRODUCTS_PER_SOLVE * ports * n_freq) else 'iterative'


def tree_arrays(basis):
    """The pq_loop_tree of a LoopBasis (and the arrays it points into, to keep alive with it)."""
    depth = np.asarray(basis.depth)
    kids = np.flatnonzero(np.asarray(basis.parent) >= 0)
    kids = kids[np.argsort(depth[kids], kind='stable')]
    n_levels = int(depth.max()) if len(depth) else 0
    level_ptr = np.searchsorted(depth[kids], np.arange(1, n_levels + 2)).astype(np.int32)
    arrays = [np.ascontiguousarray(a, t) for a, t in
              ((basis.parent, np.int32), (basis.upb, np.int32), (basis.ups, np.float32), (level_ptr, np.int32),
               (kids, np.int32), (basis.cotree, np.int32), (basis.cot_ends[:, 0], np.int32),
               (basis.cot_ends[:, 1], np.int32))]
    par, upb, ups, lp, ln, cot, ea, eb = arrays
    tree = _lib.LoopTree(len(par), ptr(par), ptr(upb), ptr(ups), n_levels, ptr(lp), ptr(ln), ptr(cot), ptr(ea),
                         ptr(eb))
    return tree, arrays


class PortSolver:
    """One board on one device. `ports(freqs)` returns Zp (F, P, P) complex and per-frequency info."""

    def __init__(self, board, device=0, near=3.0, outer_levels=1, precision=32, ideal_terminals=True, model=None,
                 projection=None):
        t0 = time.perf_counter()
        self.device = device
        self.model = model if model is not None else LoopModel(board, ideal_terminals=ideal_terminals)
        m = self.model
        t1 = time.perf_counter()
        g = FillGeometry.prepare(board.nodes, board.tris, board.z_of, outer_levels)
        # one non-blocking stream per solver: several boards can run at once from several threads
        stream = ctypes.c_void_p()
        call('pq_stream_create', device, ctypes.byref(stream))
        self._stream = stream
        self.backend = Mi355xBackend(g, 1, 16384, device)
        planes, self.analytic = self.backend.fill_device(near, outer_levels, False, precision, stream=stream)
        rwg = RWG(g.nodes, board.tris)                        # the fill's centred frame, same edges
        ne, nv = m.ne, m.nv
        nbr = ne + nv
        ldb = padded(nbr)
        lb_type = np.float64 if precision == 64 else np.float32
        Lb = DeviceArray((nbr, ldb), lb_type, device).zero()
        call('pq_stream_synchronize', None)                   # the zero fill ran on the null stream
        tri, alpha, vertex = rwg.combine_arrays()
        scale = 1.0 / rwg.l
        dev = [DeviceArray.from_host(a, device) for a in (tri, alpha, vertex, scale,
                                                          np.ascontiguousarray(g.cent, np.float64))]
        call('pq_combine_device', int(precision), planes.ptr, len(g.tris), planes.shape[-1], dev[4].ptr, dev[0].ptr,
             dev[1].ptr, dev[2].ptr, dev[3].ptr, ne, MU0 / (4 * np.pi), 64 if precision == 64 else 32, Lb.ptr, ldb,
             stream)
        call('pq_stream_synchronize', stream)
        for a in dev:
            a.close()
        planes.close()
        for k in range(nv):                                   # the via block
            row = np.ascontiguousarray(m.Lv[k], lb_type)
            call('pq_memcpy_h2d', Lb.offset((ne + k) * ldb + ne), ptr(row), row.nbytes)
        # the branches' self inductances (the graph preconditioner's D), from the device matrix's diagonal
        self.branch_ldiag = np.zeros(m.nb)
        diag = np.empty(nbr, lb_type)
        size = np.dtype(lb_type).itemsize
        call('pq_memcpy_d2h_strided', ptr(diag), Lb.ptr, size, (ldb + 1) * size, nbr)
        self.branch_ldiag[:nbr] = diag
        Rl = m.Rl.tocsr()
        Rl.sort_indices()
        rptr, rcol, rval = (np.ascontiguousarray(Rl.indptr, np.int32), np.ascontiguousarray(Rl.indices, np.int32),
                            np.ascontiguousarray(Rl.data, np.float64))
        handle = ctypes.c_void_p()
        self.projection = projection or os.environ.get('PCBQS_AMD_PROJECTION', 'tree')
        if self.projection == 'tree':
            tree, keep = tree_arrays(m.basis)
            call('pq_loops_create_tree', device, m.n_loops, m.n_free, ctypes.byref(tree), Lb.ptr, 8 * size, ldb, nbr,
                 ptr(rptr), ptr(rcol), ptr(rval), ctypes.byref(handle))
        elif self.projection == 'columns':
            C = m.basis.C.tocsc()
            colptr = np.ascontiguousarray(C.indptr, np.int32)
            rows = np.ascontiguousarray(C.indices, np.int32)
            sign = np.ascontiguousarray(C.data, np.float32)
            call('pq_loops_create', device, m.n_loops, m.n_free, ptr(colptr), ptr(rows), ptr(sign), len(rows), Lb.ptr,
                 8 * size, ldb, nbr, ptr(rptr), ptr(rcol), ptr(rval), ctypes.byref(handle))
        else:
            raise ValueError(f"projection must be 'tree' or 'columns', not {self.projection!r}")
        Lb.close()
        self._handle = handle
        self.port_cols = np.ascontiguousarray(m.port_cols, np.int32)
        self.precision = precision
        self.setup_seconds = dict(topology=t1 - t0, device=time.perf_counter() - t1)

    def set_coarse(self, phi, diagonalize=True):
        """Coarse space (n_free, k) of the two-level preconditioner (None or k = 0: none), a dense
        array or a scipy sparse matrix.

        With diagonalize, the span is re-expressed in the generalised eigenbasis of the Galerkin
        pair (Phi^T Rl Phi, Phi^T Ll Phi): Psi = Phi V with V^T Rc V = I and V^T Lc V = diag, so
        every frequency's coarse solve is a diagonal scaling. Directions with a Galerkin
        resistance below 1e-10 of the largest are dropped (they carry no loop current). The
        change of basis runs on the device (pq_loops_set_coarse_csc for sparse columns, which
        cross PCIe as their nonzeros only, else pq_loops_set_coarse_diagonal; coarse_path tells
        whether the Cholesky route or the eigenvector route was taken); diagonalize='host' does it
        with numpy instead (two eigh calls), False keeps Phi and inverts the k x k coarse matrix
        per frequency."""
        from scipy import sparse
        if phi is None:
            phi = np.zeros((self.model.n_free, 0))
        if phi.ndim != 2 or phi.shape[0] != self.model.n_free:
            raise ValueError('coarse space must have one row per free loop')
        k = phi.shape[1]
        self.coarse_path = None
        if sparse.issparse(phi) and k and diagonalize is True:
            m = sparse.csc_matrix(phi, dtype=np.float64)
            m.sort_indices()
            self.coarse_phi = m
            colptr = np.ascontiguousarray(m.indptr, np.int32)
            rows = np.ascontiguousarray(m.indices, np.int32)
            vals = np.ascontiguousarray(m.data, np.float64)
            kept, path = ctypes.c_int(), ctypes.c_int()
            call('pq_loops_set_coarse_csc', self._handle, k, ptr(colptr), ptr(rows), ptr(vals), 1e-10,
                 ctypes.byref(kept), ctypes.byref(path))
            self.coarse_k = kept.value
            self.coarse_path = {1: 'cholesky', 2: 'eigenvectors'}[path.value]
            return
        if sparse.issparse(phi):
            phi = phi.toarray()
        phi = np.ascontiguousarray(phi, np.float64)
        self.coarse_phi = phi
        if k and diagonalize is True:
            kept = ctypes.c_int()
            call('pq_loops_set_coarse_diagonal', self._handle, k, ptr(phi), 1e-10, ctypes.byref(kept))
            self.coarse_k = kept.value
            return
        call('pq_loops_set_coarse', self._handle, k, ptr(phi) if k else None, 0)
        if k and diagonalize:
            Rc, Lc = np.empty((k, k)), np.empty((k, k))
            call('pq_loops_coarse_matrices', self._handle, ptr(Rc), ptr(Lc))
            s, U = np.linalg.eigh(0.5 * (Rc + Rc.T))
            keep = s > 1e-10 * s.max()
            W = U[:, keep] / np.sqrt(s[keep])
            lam, Q = np.linalg.eigh(W.T @ (0.5 * (Lc + Lc.T)) @ W)
            phi = np.ascontiguousarray(phi @ (W @ Q))
            k = phi.shape[1]
            call('pq_loops_set_coarse', self._handle, k, ptr(phi), 1)
        self.coarse_k = k

    def set_stream_coarse(self, grid=8, vias=True, extra=None, extra_loops=None, diagonalize=True):
        """The geometric coarse space: stream-function hats of a grid x grid net per layer, the
        via-closed loops (vias=True), plus extra branch patterns (nb, k') and extra loop-space
        columns (n_free, k'', dense or sparse) when given. The columns stay sparse up to the device
        and repeats are removed (coarse.unique_columns)."""
        from scipy import sparse
        from .coarse import loop_restrict, stream_patterns, unique_columns, via_loop_columns
        patterns = stream_patterns(self.model, grid)
        if extra is not None and extra.shape[1]:
            patterns = sparse.hstack([patterns, sparse.csc_matrix(extra)]).tocsc()
        cols = [loop_restrict(self.model, patterns, dense=False)]
        if vias:
            cols.append(via_loop_columns(self.model, dense=False))
        if extra_loops is not None and extra_loops.shape[1]:
            cols.append(sparse.csc_matrix(extra_loops))
        self.set_coarse(unique_columns(sparse.hstack(cols, format='csc')), diagonalize=diagonalize)

    def prepare_block_jacobi(self):
        """The block-Jacobi preconditioner's per-board set-up (done by the first iterative solve
        otherwise): the 256-loop blocks' generalised eigenvectors, from which every frequency's block
        inverses are one batched GEMM. Returns 'modes', or 'gauss-jordan' where a block's resistance
        is not positive definite (each frequency then inverts its blocks)."""
        modal = ctypes.c_int()
        call('pq_loops_prepare_block_jacobi', self._handle, ctypes.byref(modal))
        return {1: 'modes', -1: 'gauss-jordan', 0: 'gauss-jordan'}[modal.value]

    @property
    def n_ports(self):
        return len(self.port_cols)

    @property
    def device_ports(self):
        """Columns the device solves in lockstep: the port count, or 16 for 9 to 15 ports (the
        extra columns repeat the last port and are dropped)."""
        P = self.n_ports
        return P if P <= 8 or P == 16 else 16

    def _split(self, x):
        """(n_free, P) complex -> (n_free, 2P) FP64 (real parts first), the library's layout."""
        return np.ascontiguousarray(np.hstack([x.real, x.imag]), np.float64)

    def _join(self, y):
        P = y.shape[1] // 2
        return y[:, :P] + 1j * y[:, P:]

    def rhs(self, frequency):
        """Free-loop right-hand sides -Z_fs (n_free, P) complex at one frequency."""
        P = self.n_ports
        B = np.zeros((self.model.n_free, 2 * P))
        call('pq_loops_rhs', self._handle, 2 * np.pi * float(frequency), P, ptr(self.port_cols), ptr(B))
        return self._join(B)

    def operator(self, X, frequency, precision=32):
        """Z_ff X = (Rl + j w Ll) X on the free loops, X (n_free, P) complex with P <= 8 or 16: with
        the FP32 copy of Ll (precision 32, the Krylov product) or the FP64 Ll (64, the system)."""
        X = np.asarray(X, complex).reshape(self.model.n_free, -1)
        Xs = self._split(X)
        W = np.empty_like(Xs)
        call('pq_loops_operator', self._handle, 2 * np.pi * float(frequency), X.shape[1], int(precision), ptr(Xs),
             ptr(W))
        return self._join(W)

    def coarse_solution(self, frequency):
        """The coarse-space solution Phi (Phi^T Z Phi)^-1 Phi^T b (n_free, P) complex of the port
        right-hand sides (needs a coarse space)."""
        P = self.n_ports
        X = np.zeros((self.model.n_free, 2 * P))
        call('pq_loops_coarse_solution', self._handle, 2 * np.pi * float(frequency), P, ptr(self.port_cols), ptr(X))
        return self._join(X)

    def ports(self, freqs, precond='block-jacobi', tol=1e-8, restart=100, maxiter=3000, currents=False, x0=None,
              solver='iterative', fallback_maxiter=3000):
        """Port impedance matrices (F, P, P) complex; info per frequency (iterations, residual, seconds).

        solver: 'iterative' (lockstep GMRES with `precond`, to `tol`), 'direct' (the free-loop
        system factored as L D L^T in FP32 on the device and refined in FP64 to a relative residual
        of 1e-12 in at most 12 corrections, as the reference's MixedLU; precond, tol, restart,
        maxiter and x0 do not apply, and `iterations` counts the corrections) or 'auto' (the
        reference's choice between them, `port_method`).
        x0: optional initial iterates, (F, n_free, P) complex or a callable frequency -> (n_free, P)
        (the learned correction's prediction); GMRES starts from zero otherwise.
        A two-level (learned) solve, or one from an initial iterate, that fails its acceptance test is
        solved again by block-Jacobi GMRES from zero (at most fallback_maxiter iterations): the plain
        method is the fallback and the judge; info then names it ('fallback_from')."""
        if solver not in ('iterative', 'direct', 'auto'):
            raise ValueError("solver must be 'iterative', 'direct' or 'auto'")
        if solver == 'auto':
            solver = port_method(self.model.n_free, self.n_ports, len(np.atleast_1d(freqs)))
        P = self.n_ports
        Pd = self.device_ports
        cols = self.port_cols if Pd == P else np.ascontiguousarray(
            np.r_[self.port_cols, np.full(Pd - P, self.port_cols[-1])], np.int32)
        f = np.atleast_1d(np.asarray(freqs, float))
        if f.ndim != 1 or not np.isfinite(f).all() or (f < 0).any():
            raise ValueError('port frequencies must be finite and nonnegative')
        Z = np.zeros((len(f), P, P), complex)
        X = np.zeros((len(f), self.model.n_free, 2 * P)) if currents else None
        info = []
        kind = PRECONDITIONERS[precond] if solver == 'iterative' else None
        batch = self._direct_batch(len(f)) if solver == 'direct' else 1
        if batch > 1 and not currents:
            # each chunk of frequencies factored and refined together (the same results as one by one)
            for i0 in range(0, len(f), batch):
                om = np.ascontiguousarray(2 * np.pi * f[i0:i0 + batch])
                nw = len(om)
                zp = np.zeros((nw, Pd, Pd, 2))
                st, rs = np.zeros(nw, np.int64), np.zeros(nw)
                t = time.perf_counter()
                call('pq_loops_solve_direct_batch', self._handle, nw, ptr(om), Pd, ptr(cols), DIRECT_TOL, DIRECT_STEPS,
                     ptr(zp), ptr(st), ptr(rs))
                seconds = (time.perf_counter() - t) / nw
                fms, nfb = ctypes.c_double(), ctypes.c_int64()
                call('pq_loops_factor_time', self._handle, ctypes.byref(fms))
                call('pq_loops_last_fallbacks', self._handle, ctypes.byref(nfb))
                for b in range(nw):
                    Z[i0 + b] = zp[b, :P, :P, 0] + 1j * zp[b, :P, :P, 1]
                    info.append(dict(frequency_Hz=float(f[i0 + b]), iterations=int(st[b]), matvecs=int(st[b]) + 1,
                                     relative_residual=float(rs[b]), seconds=seconds, preconditioner='direct',
                                     factor_ms=float(fms.value), chunk_fp64_fallbacks=int(nfb.value)))
            return Z, info
        for i, fi in enumerate(f):
            zp = np.zeros((Pd, Pd, 2))
            its, res = ctypes.c_int64(), ctypes.c_double()
            xh = np.zeros((self.model.n_free, 2 * Pd)) if currents else None
            if solver == 'direct' and batch > 1 and i % batch == 0:
                # factor the next frequencies together; the solves below refine on these factors
                om = np.ascontiguousarray(2 * np.pi * f[i:i + batch])
                call('pq_loops_factor_batch', self._handle, len(om), ptr(om), None)
            if solver == 'direct':
                call('pq_loops_solve_direct', self._handle, 2 * np.pi * fi, Pd, ptr(cols), DIRECT_TOL, DIRECT_STEPS,
                     ptr(zp), ptr(xh) if currents else None, ctypes.byref(its), ctypes.byref(res))
                Z[i] = zp[:P, :P, 0] + 1j * zp[:P, :P, 1]
                if currents:
                    X[i] = np.hstack([xh[:, :P], xh[:, Pd:Pd + P]])
                seconds, mv, fms, nfb = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
                call('pq_loops_last_solve', self._handle, None, ctypes.byref(mv), None, ctypes.byref(seconds))
                call('pq_loops_factor_time', self._handle, ctypes.byref(fms))
                call('pq_loops_last_fallbacks', self._handle, ctypes.byref(nfb))
                info.append(dict(frequency_Hz=float(fi), iterations=int(its.value), matvecs=int(mv.value),
                                 relative_residual=float(res.value), seconds=float(seconds.value),
                                 preconditioner='direct', factor_ms=float(fms.value), fp64_fallback=bool(nfb.value)))
                continue
            start = None
            if x0 is not None:
                guess = x0(fi) if callable(x0) else x0[i]
                if guess.shape != (self.model.n_free, P):
                    raise ValueError('initial iterate must be (n_free, P)')
                if Pd != P:
                    guess = np.hstack([guess, np.repeat(guess[:, -1:], Pd - P, axis=1)])
                start = self._split(guess)
            used = precond
            try:
                call('pq_loops_solve_ports', self._handle, 2 * np.pi * fi, Pd, ptr(cols), kind, float(tol),
                     int(restart), int(maxiter), ptr(zp), ptr(xh) if currents else None,
                     ptr(start) if start is not None else None, ctypes.byref(its), ctypes.byref(res))
            except _lib.ConvergenceError:
                if kind != PRECONDITIONERS['two-level'] and start is None:
                    raise
                used = 'block-jacobi'
                call('pq_loops_solve_ports', self._handle, 2 * np.pi * fi, Pd, ptr(cols),
                     PRECONDITIONERS['block-jacobi'], float(tol), int(restart), int(fallback_maxiter), ptr(zp),
                     ptr(xh) if currents else None, None, ctypes.byref(its), ctypes.byref(res))
            Z[i] = zp[:P, :P, 0] + 1j * zp[:P, :P, 1]
            if currents:
                X[i] = np.hstack([xh[:, :P], xh[:, Pd:Pd + P]])
            seconds, mv = ctypes.c_double(), ctypes.c_int64()
            call('pq_loops_last_solve', self._handle, None, ctypes.byref(mv), None, ctypes.byref(seconds))
            info.append(dict(frequency_Hz=float(fi), iterations=int(its.value), matvecs=int(mv.value),
                             relative_residual=float(res.value), seconds=float(seconds.value), preconditioner=used))
            if used != precond:
                info[-1]['fallback_from'] = precond if start is None else f'{precond} from an initial iterate'
        return (Z, info, X) if currents else (Z, info)

    def use_graphs(self, enable=True):
        """Replay each Arnoldi step as a captured graph (True), launch kernels one by one (False),
        or follow PCBQS_AMD_GRAPHS (None)."""
        call('pq_loops_use_graphs', self._handle, -1 if enable is None else int(bool(enable)))

    def tuning(self, columns=None):
        """(chunk, microseconds) of the autotuned operator product for 2P real columns (P ports by
        default): 0 for the symmetric upper-triangle kernel, else the whole-matrix kernel's
        column-chunk width; and its time per product."""
        nr = 2 * self.device_ports if columns is None else int(columns)
        chunk, us = ctypes.c_int(), ctypes.c_double()
        call('pq_loops_tuning', self._handle, nr, ctypes.byref(chunk), ctypes.byref(us))
        return chunk.value, us.value

    def _direct_batch(self, n_freq):
        """Frequencies whose direct factorisations run together: all of them up to 32, within about
        16 GB of factors (PCBQS_AMD_DIRECT_BATCH=0: one at a time)."""
        if n_freq < 2 or os.environ.get('PCBQS_AMD_DIRECT_BATCH', '1') == '0':
            return 1
        n = self.model.n_free
        nblk = (n + 63) // 64
        per = 4 * (2 * n * nblk * 64 + 2 * nblk * 64 * 64 + 8 * n * 64)
        return int(max(1, min(n_freq, 32, (16 << 30) // max(per, 1))))

    def graph_preconditioner(self, frequency):
        """The graph preconditioner (C_f^T D C_f)^-1 at one frequency, D the branches' self impedances
        (src/pcbqs/mqs_iterative.py:337-341): (LinearOperator on the free loops, factor bytes)."""
        from .precond import branch_diagonal, graph, graph_setup
        p = getattr(self, '_graph_setup', None) or graph_setup(self.model)
        self._graph_setup = p
        d = branch_diagonal(self.model.R.diagonal(), self.branch_ldiag, 2 * np.pi * frequency)
        return graph(p, d, device=self.device)

    def loop_matrix(self, rows=None, precision=32):
        """A host copy of the device's loop inductance Ll (its FP32 copy, or the FP64 matrix with
        precision 64; the first `rows` rows, all columns)."""
        n = self.model.n_loops
        rows = n if rows is None else int(rows)
        dev, ld = ctypes.c_void_p(), ctypes.c_int64()
        call('pq_loops_matrix', self._handle, int(precision), ctypes.byref(dev), ctypes.byref(ld))
        out = np.empty((rows, ld.value), np.float64 if precision == 64 else np.float32)
        call('pq_memcpy_d2h', ptr(out), dev, out.nbytes)
        return out[:, :n]

    def close(self):
        if getattr(self, '_handle', None) is not None and _lib._lib is not None:
            _lib._lib.pq_loops_destroy(self._handle)
        self._handle = None
        if getattr(self, '_stream', None) is not None and _lib._lib is not None:
            _lib._lib.pq_stream_destroy(self._stream)
        self._stream = None
        if getattr(self, 'backend', None) is not None:
            self.backend.close()

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass


def rl_equivalent(frequency, Z, tolerance=1e-7):
    """(R, L) of one port matrix, each made PSD (src/pcbqs/mqs_ports.py:72-86)."""
    if frequency <= 0:
        raise ValueError('RL extraction needs a positive reference frequency')
    mats = []
    for A in (Z.real, Z.imag / (2 * np.pi * frequency)):
        A = (A + A.T) / 2
        e, V = np.linalg.eigh(A)
        scale = max(abs(e).max(), 1e-30)
        if e.min() < -tolerance * scale:
            raise ValueError('non-passive copper reduction; refine/check topology')
        mats.append((V * np.maximum(e, scale * 1e-12)) @ V.T)
    return mats[0], mats[1]


def main(argv=None):
    """Port impedances of a seeded board: the set-up (fill, combine, projection), then the solve at
    the given frequencies; one JSON line with the set-up time and each solve's info."""
    import argparse
    import json
    from . import board as boards
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument('--board', choices=['c1', 'c2', 'c5'], default='c2')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--solver', choices=['iterative', 'direct', 'auto'], default='iterative')
    ap.add_argument('--precond', choices=['block-jacobi', 'two-level'], default='block-jacobi')
    ap.add_argument('--frequencies', type=float, nargs='+', default=[1e6])
    ap.add_argument('--tol', type=float, default=1e-8)
    ap.add_argument('--coarse', choices=['hats', 'learned'], default='learned',
                    help='two-level coarse space: 36 x 36 hats and via loops, or with the learned enrichment')
    ap.add_argument('--coarse-diagonalize', choices=['device', 'host'], default='device',
                    help="the coarse Galerkin pair's eigenbasis on the device (rocSOLVER) or with numpy on the host "
                         "(for runs under rocprofv3 --pmc, where rocSOLVER's dsyevd ended the process with a host "
                         "segmentation fault)")
    args = ap.parse_args(argv)
    b = {'c1': lambda: boards.board_c1(), 'c2': lambda: boards.board_c2(args.seed),
         'c5': lambda: boards.board_c5(args.seed)}[args.board]()
    s = PortSolver(b, device=0)
    print(f'set-up: {b.nt} triangles, {s.model.n_free} free loops, {s.setup_seconds}', flush=True)
    if args.precond == 'two-level' and args.solver != 'direct':
        how = True if args.coarse_diagonalize == 'device' else 'host'
        if args.coarse == 'learned':
            from .learned import MPNN, BranchGraph, enrichment
            s.set_stream_coarse(36, extra_loops=enrichment(MPNN.load(), BranchGraph(s.model)), diagonalize=how)
        else:
            s.set_stream_coarse(36, diagonalize=how)
        print(f'coarse space: {s.coarse_k} columns', flush=True)
    Z, info = s.ports(args.frequencies, solver=args.solver, precond=args.precond, tol=args.tol)
    print(json.dumps(dict(board=args.board, triangles=int(b.nt), free_loops=int(s.model.n_free),
                          setup=s.setup_seconds, solves=[{k: v for k, v in i.items() if np.isscalar(v)}
                                                         for i in info])), flush=True)
    s.close()


if __name__ == '__main__':
    main()
