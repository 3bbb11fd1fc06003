"""Batched port sweeps: many uneven boards x ports x frequencies on one device.

Every board runs the device pipeline once (fill -> RWG combine -> loop projection -> coarse space)
and then solves its frequencies, all ports of a frequency as the columns of one lockstep GMRES
(src/pcbqs/mqs_iterative.py:109-170: every step one operator product for all columns). The host
half of the next board (RWG topology, spanning tree and loop basis, stream-function patterns) is
built on a worker thread while the device works on the current board. The result per board is the
reference's port response (src/pcbqs/mqs_ports.py:12-49): Z (F, P, P) and the solve statistics.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .loops import LoopModel


def _host(board, grid, precond, solver, model=None):
    model = LoopModel(board) if model is None else model
    phi = None
    if precond == 'two-level' and solver == 'iterative':
        from scipy import sparse
        from .coarse import loop_restrict, stream_patterns, unique_columns, via_loop_columns
        phi = unique_columns(sparse.hstack([loop_restrict(model, stream_patterns(model, grid), dense=False),
                                            via_loop_columns(model, dense=False)], format='csc'))
    return model, phi


def _board(i, board, freqs, precond, grid, tol, device, graphs, solver_kind):
    from .solve import PortSolver, port_method
    t = time.perf_counter()
    model = None
    if solver_kind == 'auto':
        model = LoopModel(board)
        solver_kind = port_method(model.n_free, len(board.ports), len(freqs))
    model, phi = _host(board, grid, precond, solver_kind, model)
    solver = PortSolver(board, device=device, model=model)
    solver.use_graphs(graphs)
    if phi is not None:
        solver.set_coarse(phi)
    setup = time.perf_counter() - t
    t = time.perf_counter()
    Z, info = solver.ports(freqs, precond=precond, tol=tol, solver=solver_kind)
    solve = time.perf_counter() - t
    solver.close()
    return dict(board=i, triangles=int(board.nt), layers=int(len(np.unique(board.layer_of))),
                ports=int(len(board.ports)), free_loops=int(model.n_free), solver=solver_kind, setup_s=setup,
                solve_s=solve,
                iterations=[r['iterations'] for r in info],
                worst_residual=max(r['relative_residual'] for r in info), Z=Z)


def _device_ports(solver):
    """(device column count, port loop columns) of a board: P itself for up to 8 ports or 16, else
    16 with the last port repeated (the extra columns are dropped)."""
    P = solver.n_ports
    Pd = solver.device_ports
    cols = solver.port_cols if Pd == P else np.r_[solver.port_cols, np.full(Pd - P, solver.port_cols[-1])]
    return Pd, np.ascontiguousarray(cols, np.int32)


def _sweep_batched(boards, freqs, precond, grid, tol, device, concurrency, log, solver_kind):
    """The sweep with every direct board's frequencies solved in captured multi-board launches
    (pq_loops_solve_direct_multi); boards the rule sends to GMRES keep the per-board route."""
    import ctypes
    from ._lib import call
    from .solve import DIRECT_STEPS, DIRECT_TOL, PortSolver, port_method
    t0 = time.perf_counter()

    def prepare(i, board):
        t = time.perf_counter()
        model = LoopModel(board)
        kind = solver_kind if solver_kind != 'auto' else port_method(model.n_free, len(board.ports), len(freqs))
        if kind != 'direct':
            return dict(i=i, kind=kind, model=model)
        s = PortSolver(board, device=device, model=model)
        return dict(i=i, kind=kind, model=model, solver=s, setup_s=time.perf_counter() - t)
    with ThreadPoolExecutor(max(1, int(concurrency))) as pool:
        prepared = list(pool.map(lambda a: prepare(*a), enumerate(boards)))
    setup_s = time.perf_counter() - t0
    results = [None] * len(boards)
    direct = [p for p in prepared if p['kind'] == 'direct']
    t1 = time.perf_counter()
    if direct:
        handles = (ctypes.c_void_p * len(direct))(*[p['solver']._handle.value for p in direct])
        dp = [_device_ports(p['solver']) for p in direct]
        Ps = np.array([d[0] for d in dp], np.int32)
        cols = (ctypes.c_void_p * len(direct))(*[d[1].ctypes.data for d in dp])
        om = np.ascontiguousarray(2 * np.pi * np.asarray(freqs, float))
        nw = len(om)
        zp = np.zeros(int(sum(nw * P * P * 2 for P in Ps)))
        st, rs = np.zeros(len(direct) * nw, np.int64), np.zeros(len(direct) * nw)
        stats = np.zeros(2, np.int64)
        call('pq_loops_solve_direct_multi', len(direct), handles, Ps.ctypes.data, cols, nw, om.ctypes.data,
             DIRECT_TOL, DIRECT_STEPS, zp.ctypes.data, st.ctypes.data, rs.ctypes.data, stats.ctypes.data)
        solve_s = time.perf_counter() - t1
        off = 0
        for k, p in enumerate(direct):
            P, Pb = int(Ps[k]), p['solver'].n_ports
            z = zp[off:off + nw * P * P * 2].reshape(nw, P, P, 2)
            off += nw * P * P * 2
            b = boards[p['i']]
            results[p['i']] = dict(board=p['i'], triangles=int(b.nt), layers=int(len(np.unique(b.layer_of))),
                                   ports=int(Pb), free_loops=int(p['model'].n_free), solver='direct',
                                   setup_s=p['setup_s'], solve_s=solve_s / len(direct),
                                   iterations=[int(v) for v in st[k * nw:(k + 1) * nw]],
                                   worst_residual=float(rs[k * nw:(k + 1) * nw].max()),
                                   Z=z[:, :Pb, :Pb, 0] + 1j * z[:, :Pb, :Pb, 1])
            p['solver'].close()
        launches, redone = int(stats[0]), int(stats[1])
    else:
        solve_s, launches, redone = 0.0, 0, 0
    for p in prepared:
        if p['kind'] != 'direct':
            results[p['i']] = _board(p['i'], boards[p['i']], freqs, precond, grid, tol, device, False, p['kind'])
    seconds = time.perf_counter() - t0
    if log:
        log(f'{len(direct)} direct boards: set-up {setup_s:.3f} s (host topology and device pipeline, '
            f'{concurrency} at once), {len(freqs)} frequencies solved in {launches} captured launches, '
            f'{solve_s:.3f} s ({redone} matrices solved again alone)')
    solves = sum(len(freqs) for _ in results)
    columns = sum(r['ports'] * len(freqs) for r in results)
    return results, dict(seconds=seconds, boards=len(results), frequencies=len(freqs), solves=solves,
                         port_columns=columns, port_columns_per_s=columns / seconds if seconds > 0 else 0.0,
                         precond=precond, grid=grid, concurrency=int(concurrency), graphs=False, solver=solver_kind,
                         boards_direct=len(direct), batched=True, graph_launches=launches, redone_alone=redone,
                         setup_s=setup_s, batched_solve_s=solve_s)


def sweep(boards, freqs, precond='two-level', grid=12, tol=1e-8, device=0, concurrency=4, graphs=False,
          log=None, solver='auto', batched=True):
    """(per-board results in board order, totals) of the port solves of every board at every
    frequency. solver='direct' factors each board's free-loop block per frequency on the device
    (FP32 L D L^T refined in FP64 to 1e-12, as the reference's dense path) instead of GMRES; precond,
    grid and tol then do not apply. solver='auto' (the default, as the reference's port_method)
    chooses per board: direct up to min(8000, 60 x ports x frequencies) free loops, else GMRES. `concurrency` boards run at once, each on its own stream from its own host thread
    (the library releases the interpreter during device calls). graphs=True (concurrency 1 only: a
    capture does not survive other threads' synchronising calls) replays each Arnoldi step as a
    captured graph; at these board sizes capturing a board's ~100 step graphs costs more than the
    launches it saves. batched=True (the default) solves every direct board's frequencies in captured
    multi-board launches (pq_loops_solve_direct_multi: boards of similar size together, one graph
    launch per group); batched=False keeps one board at a time."""
    if graphs and concurrency > 1:
        raise ValueError('captured steps need concurrency 1')
    freqs = [float(f) for f in np.atleast_1d(freqs)]
    if batched and solver in ('direct', 'auto') and not graphs:
        return _sweep_batched(boards, freqs, precond, grid, tol, device, concurrency, log, solver)
    t0 = time.perf_counter()
    results = [None] * len(boards)
    with ThreadPoolExecutor(max(1, int(concurrency))) as pool:
        futures = {pool.submit(_board, i, b, freqs, precond, grid, tol, device, graphs, solver): i
                   for i, b in enumerate(boards)}
        for fut in futures:
            r = fut.result()
            results[r['board']] = r
            if log:
                log(f"board {r['board']}: {r['triangles']} triangles, {r['ports']} ports, {r['free_loops']} loops, "
                    f"setup {r['setup_s']:.3f} s, {len(freqs)} frequencies in {r['solve_s']:.3f} s")
    seconds = time.perf_counter() - t0
    solves = sum(len(freqs) for _ in results)
    columns = sum(r['ports'] * len(freqs) for r in results)
    return results, dict(seconds=seconds, boards=len(results), frequencies=len(freqs), solves=solves,
                         port_columns=columns, port_columns_per_s=columns / seconds if seconds > 0 else 0.0,
                         precond=precond, grid=grid, concurrency=int(concurrency), graphs=graphs, solver=solver,
                         boards_direct=sum(r['solver'] == 'direct' for r in results))


def main(argv=None):
    import argparse
    import json
    from . import board as boards
    ap = argparse.ArgumentParser(description='batched port sweep over seeded uneven boards')
    ap.add_argument('--boards', type=int, default=64)
    ap.add_argument('--frequencies', type=int, default=20)
    ap.add_argument('--precond', default='two-level')
    ap.add_argument('--grid', type=int, default=12)
    ap.add_argument('--concurrency', type=int, default=4)
    ap.add_argument('--graphs', action='store_true')
    ap.add_argument('--solver', default='auto', choices=('auto', 'iterative', 'direct'))
    ap.add_argument('--per-board', action='store_true', help='one board at a time (no multi-board launches)')
    args = ap.parse_args(argv)
    freqs = np.logspace(3, 8, args.frequencies)
    bs = [boards.board_batch(s) for s in range(args.boards)]
    res, total = sweep(bs, freqs, precond=args.precond, grid=args.grid, concurrency=args.concurrency,
                       graphs=args.graphs, log=lambda m: print(m, flush=True), solver=args.solver,
                       batched=not args.per_board)
    print(json.dumps(total))


if __name__ == '__main__':
    main()
