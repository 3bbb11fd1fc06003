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
        from .coarse import loop_restrict, stream_patterns, via_loop_columns
        phi = np.hstack([loop_restrict(model, stream_patterns(model, grid)), via_loop_columns(model)])
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


def sweep(boards, freqs, precond='two-level', grid=12, tol=1e-8, device=0, concurrency=4, graphs=False,
          log=None, solver='auto'):
    """(per-board results in board order, totals) of the port solves of every board at every
    frequency. solver='direct' factors each board's free-loop block per frequency on the device
    (FP32 L D L^T refined in FP64 to 1e-12, as the reference's dense path) instead of GMRES; precond,
    grid and tol then do not apply. solver='auto' (the default, as the reference's port_method)
    chooses per board: direct up to min(8000, 60 x ports x frequencies) free loops, else GMRES. `concurrency` boards run at once, each on its own stream from its own host thread
    (the library releases the interpreter during device calls). graphs=True (concurrency 1 only: a
    capture does not survive other threads' synchronising calls) replays each Arnoldi step as a
    captured graph; at these board sizes capturing a board's ~100 step graphs costs more than the
    launches it saves."""
    if graphs and concurrency > 1:
        raise ValueError('captured steps need concurrency 1')
    freqs = [float(f) for f in np.atleast_1d(freqs)]
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
    args = ap.parse_args(argv)
    freqs = np.logspace(3, 8, args.frequencies)
    bs = [boards.board_batch(s) for s in range(args.boards)]
    res, total = sweep(bs, freqs, precond=args.precond, grid=args.grid, concurrency=args.concurrency,
                       graphs=args.graphs, log=lambda m: print(m, flush=True), solver=args.solver)
    print(json.dumps(total))


if __name__ == '__main__':
    main()
