"""The surface-charge (PEEC) solve of a board on the device: copper currents with the charges they
leave on the copper, the node potentials, and the sources' impedances, in a homogeneous dielectric.

Mirrors src/pcbqs/peec.py:164-292 (prepare, solve) for the boards of pcbqs_amd.board: copper
branches (RWG edges, via segments) and current sources (the ports) between terminal nodes; no
lumped passive parts, no frequency-dependent resistance. With branch incidence A (+1 at a branch's
first node) and exp(+j w t):
    Z i = A^T P q,    A i + j w q = 0,
P the node potential matrix (volts per coulomb, referred to infinity). The unknowns are the
passive branches' currents in the basis K = [loops | tree branches]: loops carry no charge, so P
enters only the tree's block (E = T^T P T, T = A K_tree); the tree's coordinates are scaled by j w,
and the electrostatic equilibrium of each copper component is split out first (both as the
reference does, to keep 1/w out of the conditioning and the small conduction charges out of
cancellation at low frequency).

Device work: the FP64 fill (its M0 plane gives P, all four planes the branch inductance), the
RWG combination, K^T Lb K and E on the device; per frequency the scaled dense system is formed in
HBM and solved by an FP64 LU with partial pivoting (pq_peec_solve). The sparse topology, the
tree's sparse LU (the equilibrium) and the small component system stay on the host, as in the
reference.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
from scipy import sparse
from scipy.sparse.csgraph import connected_components
from scipy.sparse.linalg import splu

from ._lib import DeviceArray, call, ptr
from .backend import Mi355xBackend, padded
from .geometry import EPS0, MU0, RWG, FillGeometry
from .loops import loop_basis, sheet_resistance, via_matrices


@dataclass
class Result:
    """The reference's Result fields for a charge solve (src/pcbqs/mqs.py Result): frequencies,
    branch currents (F, nb), source impedances {name: (F,)}, node charges and potentials (F, n_nodes)."""
    f: np.ndarray
    i: np.ndarray
    z: dict
    charge: np.ndarray
    potential: np.ndarray


class ChargeSolver:
    """One board's charge solve on one device (precision: FP64 throughout, as the reference)."""

    def __init__(self, board, device=0, er=1.0, near=3.0, outer_levels=1, ideal_terminals=True):
        if not (np.isfinite(er) and er > 0):
            raise ValueError('er must be a finite positive relative permittivity')
        self.board, self.device, self.er = board, device, float(er)
        r = RWG(board.nodes, board.tris)
        self.ne, self.nv, self.np_ = r.Ne, len(board.vseg), len(board.ports)
        node_of = board.node_of
        n_nodes = board.n_nodes
        self.n_nodes = n_nodes
        br = np.r_[np.c_[node_of[r.tp], node_of[r.tm]].reshape(-1, 2),
                   np.asarray(board.via_nodes, np.int64).reshape(-1, 2)]
        src = np.array([(board.terminals[a], board.terminals[b]) for _, a, b in board.ports], np.int64).reshape(-1, 2)
        ends = np.r_[br, src]
        nb, nc_br = len(ends), len(br)
        self.nb, self.names = nb, [name for name, _, _ in board.ports]
        is_src = np.r_[np.zeros(nc_br, bool), np.ones(self.np_, bool)]
        # the topology (peec.py:168-176): incidence, loops of the copper, the tree
        a, b = ends[:, 0], ends[:, 1]
        branch = np.arange(nb)
        self.A = sparse.coo_matrix((np.r_[np.ones(nb), -np.ones(nb)], (np.r_[a, b], np.r_[branch, branch])),
                                   shape=(n_nodes, nb)).tocsc()
        bpos = np.zeros((nb, 2))
        basis, _ = loop_basis(n_nodes, ends, is_src, nc_br, self.ne, bpos, ideal_terminals)
        passive, sources = np.flatnonzero(~is_src), np.flatnonzero(is_src)
        cycles = basis.C[:, :basis.n_free]
        tree = np.flatnonzero(basis.tree[:nc_br])
        K = sparse.hstack([cycles[passive], sparse.identity(nb, format='csc')[passive][:, tree]], format='csc')
        self.nl = basis.n_free
        self.K, self.passive, self.sources, self.tree = K, passive, sources, tree
        self.T = (self.A[:, passive] @ K[:, self.nl:]).tocsc()
        # the components of the copper graph and the equilibrium (peec.py:187-197)
        graph = sparse.coo_matrix((np.ones(len(passive)), (a[passive], b[passive])), shape=(n_nodes, n_nodes))
        ncomp, labels = connected_components(graph, directed=False)
        self.B = np.eye(ncomp)[labels]
        roots = np.unique(labels, return_index=True)[1]
        self.keep = np.ones(n_nodes, bool)
        self.keep[roots] = False
        self.tree_lu = splu(self.T[self.keep].astype(complex).tocsc()) if self.T.shape[1] else None
        self.source_components = self.A[:, sources].T @ self.B
        # the device: FP64 fill -> P (M0 plane) and Lb (combine) -> K^T Lb K, E
        g = FillGeometry.prepare(board.nodes, board.tris, board.z_of, outer_levels)
        backend = Mi355xBackend(g, 1, 16384, device)
        try:
            planes, _ = backend.fill_device(near, outer_levels, False, 64)
            nt = len(board.tris)
            order = np.argsort(node_of, kind='stable').astype(np.int32)
            nptr = np.r_[0, np.cumsum(np.bincount(node_of, minlength=n_nodes))].astype(np.int32)
            area = np.bincount(node_of, weights=g.area, minlength=n_nodes)
            d_ptr, d_tris, d_area = (DeviceArray.from_host(x, device) for x in (nptr, order, area))
            self.P = DeviceArray((n_nodes, n_nodes), np.float64, device)
            call('pq_potential_device', 64, planes.ptr, nt, planes.shape[-1], n_nodes, d_ptr.ptr, d_tris.ptr, d_area.ptr,
                 1.0 / (4 * np.pi * EPS0 * self.er), 64, self.P.ptr, n_nodes, None)
            call('pq_synchronize', device)
            for x in (d_ptr, d_tris, d_area):
                x.close()
            rg = RWG(g.nodes, board.tris)                       # the fill's centred frame, same edges
            tri, alpha, vertex = rg.combine_arrays()
            ldb = padded(nc_br)
            Lb = DeviceArray((nc_br, ldb), np.float64, device).zero()
            dev = [DeviceArray.from_host(x, device) for x in (tri, alpha, vertex, 1.0 / rg.l,
                                                              np.ascontiguousarray(g.cent, np.float64))]
            call('pq_combine_device', 64, planes.ptr, nt, planes.shape[-1], dev[4].ptr, dev[0].ptr, dev[1].ptr,
                 dev[2].ptr, dev[3].ptr, self.ne, MU0 / (4 * np.pi), 64, Lb.ptr, ldb, None)
            call('pq_synchronize', device)
            for x in dev:
                x.close()
            planes.close()
        finally:
            backend.close()
        Rv, Lv = via_matrices(board.vseg, board.sigma, board.plating)
        for k in range(self.nv):                              # the via block
            row = np.ascontiguousarray(Lv[k], np.float64)
            call('pq_memcpy_h2d', Lb.offset((self.ne + k) * ldb + self.ne), ptr(row), row.nbytes)
        R = sparse.block_diag((sheet_resistance(r, board.t_of, board.sigma), sparse.diags(Rv)), format='csr')
        RK = (K.T @ R @ K).tocsr()
        RK.sort_indices()
        colptr = np.ascontiguousarray(K.indptr, np.int32)
        rows = np.ascontiguousarray(K.indices, np.int32)
        sign = np.ascontiguousarray(K.data, np.float32)
        handle = ctypes.c_void_p()
        nK = K.shape[1]
        if nK:                          # (terminals covering whole conductors leave no unknowns)
            call('pq_loops_create', device, nK, nK, ptr(colptr), ptr(rows), ptr(sign), len(rows), Lb.ptr, 64, ldb,
                 nc_br, ptr(np.ascontiguousarray(RK.indptr, np.int32)), ptr(np.ascontiguousarray(RK.indices, np.int32)),
                 ptr(np.ascontiguousarray(RK.data, np.float64)), ctypes.byref(handle))
        else:
            handle = None
        Lb.close()
        self._handle = handle
        self.nK = nK
        Ph = self.P.host()
        self.Ph = Ph
        self.pscale = float(np.abs(Ph).max())
        ntr = nK - self.nl
        self.E = DeviceArray((max(ntr, 1), max(ntr, 1)), np.float64, device)
        if ntr:
            ta = np.ascontiguousarray(a[tree], np.int32)
            tb = np.ascontiguousarray(b[tree], np.int32)
            call('pq_peec_electro', device, self.P.ptr, n_nodes, n_nodes, ntr, ptr(ta), ptr(tb), self.pscale,
                 self.E.ptr, ntr)
        self._pinv_b = None

    def _equilibrium(self, excitation):
        """(qpart, vpart, source_voltage, y0): the copper components' electrostatic equilibrium for the
        sources' net currents, and the tree currents that carry the drive (peec.py:223-235)."""
        drive = np.asarray(self.A[:, self.sources] @ excitation).ravel()
        totals = self.source_components.T @ excitation
        qpart = np.zeros(self.n_nodes, complex)
        vpart = np.zeros(self.B.shape[1], complex)
        if np.any(totals):
            if self._pinv_b is None:
                self.capacitance_matrix()
            vpart = np.linalg.solve(self._component, -totals)
            qpart = self._pinv_b @ vpart
        y0 = (self.tree_lu.solve((-drive - qpart)[self.keep]) if self.tree_lu is not None
              else np.zeros(0, complex))
        return qpart, vpart, -self.source_components @ vpart, y0

    def _operator(self, x, omega):
        """(K^T R K + j w K^T Lb K) x for one complex vector (the FP64 product)."""
        xs = np.ascontiguousarray(np.c_[x.real, x.imag], np.float64)
        W = np.empty_like(xs)
        call('pq_loops_operator', self._handle, float(omega), 1, 64, ptr(xs), ptr(W))
        return W[:, 0] + 1j * W[:, 1]

    def solve(self, freqs, amps=None):
        """Result(f, i (F, nb), z {source: (F,)}, charge (F, n_nodes), potential (F, n_nodes))."""
        f = np.atleast_1d(np.asarray(freqs, float))
        if f.ndim != 1 or not np.isfinite(f).all() or (f <= 0).any():
            raise ValueError('frequencies must be finite and positive')
        excitation = np.array([(amps or {}).get(n, 1.0) for n in self.names], complex)
        qpart, vpart, source_voltage, y0 = self._equilibrium(excitation)
        nl, K, T = self.nl, self.K, self.T
        baseline = np.r_[np.zeros(nl, complex), y0]
        currents = np.zeros((len(f), self.nb), complex)
        charge = np.zeros((len(f), self.n_nodes), complex)
        potential = np.zeros_like(charge)
        impedance = {n: np.zeros(len(f), complex) for n in self.names}
        for fi, freq in enumerate(f):
            w = 2 * np.pi * freq
            jw = 1j * w
            if self.nK:
                rhs = -self._operator(baseline, w)
                B = np.ascontiguousarray(np.c_[rhs.real, rhs.imag])
                X = np.empty_like(B)
                zscale = ctypes.c_double()
                call('pq_peec_solve', self._handle, nl, self.E.ptr, max(self.nK - nl, 1), self.pscale, w, 1, ptr(B),
                     ptr(X), ctypes.byref(zscale))
                x = X[:, 0] + 1j * X[:, 1]
                alpha = zscale.value / self.pscale
                currents[fi, self.passive] = K[:, :nl] @ x[:nl] + K[:, nl:] @ (y0 + jw * alpha * x[nl:])
                tq = T @ x[nl:]
                ptq = self.Ph @ tq
                charge[fi] = qpart / jw - alpha * tq
                potential[fi] = self.B @ vpart / jw - alpha * ptq
                voltage = source_voltage / jw + alpha * (self.A[:, self.sources].T @ ptq)
            else:
                charge[fi] = qpart / jw
                potential[fi] = self.B @ vpart / jw
                voltage = source_voltage / jw
            currents[fi, self.sources] = excitation
            for j, name in enumerate(self.names):
                impedance[name][fi] = voltage[j] / excitation[j] if excitation[j] else np.nan
        return Result(f, currents, impedance, charge, potential)

    def capacitance_matrix(self):
        """(C, labels): the Maxwell capacitance matrix of the copper components (potentials referred
        to infinity) and each node's component, C = B^T P^-1 B with P solved on the device
        (src/pcbqs/peec.py:95-123)."""
        if self._pinv_b is None:
            n = self.n_nodes
            X = np.empty((n, self.B.shape[1]))
            steps, res, fms = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
            call('pq_spd_solve', self.device, n, self.P.ptr, n, self.B.shape[1], ptr(np.ascontiguousarray(self.B)),
                 1e-12, 12, ptr(X), ctypes.byref(steps), ctypes.byref(res), ctypes.byref(fms))
            self._pinv_b = X
            self._component = self.B.T @ X
        C = self.B.T @ self._pinv_b
        return (C + C.T) / 2, self.B.argmax(1)

    def potential_matrix(self):
        """The node potential matrix P (n_nodes square, FP64 host)."""
        return self.Ph

    def close(self):
        from . import _lib
        if getattr(self, '_handle', None) is not None and _lib._lib is not None:
            _lib._lib.pq_loops_destroy(self._handle)
        self._handle = None
        for x in ('P', 'E'):
            if getattr(self, x, None) is not None:
                getattr(self, x).close()
                setattr(self, x, None)

    def __del__(self):
        try:
            self.close()
        except Exception:                                     # noqa: BLE001 (interpreter shutdown)
            pass
