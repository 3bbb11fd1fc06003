"""The large single board end to end on one device (config 5): a seeded 4-layer 120 x 120 mm PDN
(about 67k triangles, 100k RWG, 200 vias, 8 ports) through the device fill (FP32 far / FP64 near
planes in HBM), the RWG combination (dense FP32 branch inductance in HBM), the loop projection,
plain block-Jacobi and learned two-level GMRES to 1e-8 at one frequency, and then H of the
learned solve's port currents at about a million points: a 1 mm grid over the board at heights
0.5-5 mm above its top layer (FP32 field kernel).

    python -m pcbqs_amd.scale [--seed 0] [--frequency 1e6] [--grid 24] [--heights 69] [--no-field] [--no-direct]

Unless --no-direct, the direct solve (FP32 L D L^T of the free-loop block on the device, FP64
refinement to 1e-12) runs at the same frequency after the two GMRES solves.

Prints one JSON line of timings, sizes and the agreement of the two solves.
"""
import argparse
import json
import time

import numpy as np


def field_sources(model, X, port_cols):
    """Sheets and via segments of the branch currents I_b = C l, l the loop currents of each port's
    excitation: the free loops' solution X (n_free, P) and a unit current in the port's own loop."""
    from .biot import rwg_currents
    b = model.board
    P = X.shape[1]
    lc = np.zeros((model.n_loops, P), complex)
    lc[:model.n_free] = X
    lc[port_cols, np.arange(P)] = 1.0
    Ib = model.basis.C @ lc
    a, bb = rwg_currents(model.rwg, Ib[:model.ne])
    sheets = []
    for lay in np.unique(b.layer_of):
        m = b.layer_of == lay
        sheets.append((b.nodes, b.tris[m], a[m], bb[m], float(b.z_of[m][0])))
    vs = np.asarray(b.vseg, float).reshape(-1, 5)
    p0 = np.c_[vs[:, 0], vs[:, 1], vs[:, 3]]
    p1 = np.c_[vs[:, 0], vs[:, 1], vs[:, 4]]
    return sheets, (p0, p1, Ib[model.ne:model.ne + model.nv])


def progress(stage, value):
    print(f'[scale] {stage}: {value}', flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description='Config-5 scale run on one device')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--frequency', type=float, default=1e6)
    ap.add_argument('--grid', type=int, default=24, help='hats per side of the two-level coarse space')
    ap.add_argument('--heights', type=int, default=69, help='field heights between 0.5 and 5 mm')
    ap.add_argument('--no-field', action='store_true')
    ap.add_argument('--no-direct', action='store_true')
    args = ap.parse_args(argv)
    from . import board as boards
    from .biot import h_field
    from .learned import MPNN, set_learned_coarse
    from ._lib import ConvergenceError
    from .solve import PortSolver
    t = time.perf_counter()
    b = boards.board_c5(args.seed)
    out = dict(workload=f'seeded 4-layer 120x120 mm PDN (seed {args.seed})', triangles=int(b.nt),
               vias=len({(round(x, 9), round(y, 9)) for x, y, _, _, _ in b.vseg}), via_segments=len(b.vseg),
               ports=len(b.ports), board_s=time.perf_counter() - t)
    t = time.perf_counter()
    solver = PortSolver(b, device=0)
    out['setup_s'] = time.perf_counter() - t
    out['setup_detail_s'] = solver.setup_seconds
    far_ms, near_ms, pairs = solver.backend.last_timing()
    nt = b.nt
    out.update(rwg=int(solver.model.ne), free_loops=int(solver.model.n_free),
               fill=dict(pairs=nt * (nt + 1) // 2, analytic=int(solver.analytic), far_ms=far_ms, near_ms=near_ms,
                         near_pairs=list(pairs), gpair_s=nt * (nt + 1) / 2 / ((far_ms + near_ms) * 1e-3) / 1e9,
                         planes_GB=4 * nt * (nt + 63) // 64 * 64 * 4 / 1e9),
               branch_inductance_GB=(solver.model.ne + solver.model.nv) ** 2 * 4 / 1e9)
    progress('set-up', out['setup_s'])
    f = args.frequency
    try:
        Zp, ip = solver.ports([f], precond='block-jacobi', tol=1e-8)
        out['plain'] = dict(seconds=ip[0]['seconds'], iterations=ip[0]['iterations'],
                            relative_residual=ip[0]['relative_residual'])
    except ConvergenceError as e:                # 3000 iterations without reaching 1e-8
        Zp, ip = None, None
        out['plain'] = dict(converged=False, message=str(e))
    progress('plain GMRES', out['plain'])
    t = time.perf_counter()
    net_s = set_learned_coarse(solver, MPNN.load(), args.grid)
    coarse_s = time.perf_counter() - t
    Zl, il, X = solver.ports([f], precond='two-level', tol=1e-8, currents=True)
    out['learned'] = dict(seconds=il[0]['seconds'], iterations=il[0]['iterations'],
                          relative_residual=il[0]['relative_residual'], coarse_columns=solver.coarse_k,
                          coarse_setup_s=coarse_s, network_s=net_s)
    progress('learned GMRES', out['learned'])
    if ip is not None:
        out['ratio'] = il[0]['seconds'] / ip[0]['seconds']
        out['port_z_rel_diff'] = float(np.abs(Zl - Zp).max() / np.abs(Zp).max())
    out['operator_product_us'] = solver.tuning()[1]
    if not args.no_direct:
        Zd, idd = solver.ports([f], solver='direct')
        nf = solver.model.n_free
        out['direct'] = dict(seconds=idd[0]['seconds'], factor_ms=idd[0]['factor_ms'],
                             factor_tflops=4 * nf ** 3 / 3 / (idd[0]['factor_ms'] * 1e-3) / 1e12,
                             corrections=idd[0]['iterations'], relative_residual=idd[0]['relative_residual'],
                             port_z_rel_diff_vs_learned=float(np.abs(Zd - Zl).max() / np.abs(Zl).max()))
        progress('direct solve', out['direct'])
    if not args.no_field:
        P = len(b.ports)
        sheets, segments = field_sources(solver.model, X[0][:, :P] + 1j * X[0][:, P:], solver.port_cols)
        lo, hi = b.nodes.min(0), b.nodes.max(0)
        xs = np.arange(lo[0], hi[0] + 1e-9, 1e-3)
        ys = np.arange(lo[1], hi[1] + 1e-9, 1e-3)
        top = float(b.z_of.max())
        zs = top + np.linspace(0.5e-3, 5e-3, args.heights)
        Xg, Yg, Zg = np.meshgrid(xs, ys, zs, indexing='ij')
        pts = np.c_[Xg.ravel(), Yg.ravel(), Zg.ravel()]
        timing = {}
        t = time.perf_counter()
        H = h_field(sheets, pts, segments, precision=32, timing=timing)
        total = time.perf_counter() - t
        K = H.shape[-1] if H.ndim == 3 else 1
        kk = 1 if K <= 1 else 2 if K <= 2 else 4 if K <= 4 else 8
        inter = timing['sources'] * timing['points']
        out['field'] = dict(points=timing['points'], sources=timing['sources'], segments=timing['segments'],
                            excitations=K, kernel_ms=timing['kernel_ms'], total_s=total,
                            interactions_per_s=inter / (timing['kernel_ms'] * 1e-3),
                            tflops=inter * (20 + 16 * kk) / (timing['kernel_ms'] * 1e-3) / 1e12,
                            max_abs_H=float(np.abs(H).max()))
        progress('field', out['field'])
    solver.close()
    print(json.dumps(out), flush=True)
    return out


if __name__ == '__main__':
    main()
