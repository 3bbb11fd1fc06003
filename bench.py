"""Benchmark: fill Gpair/s and the port GMRES solve on the 30k-RWG seeded plane pair.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--seed S] [--no-cpu] [--no-gmres]

A step is one device-resident fill of the board's four moment planes (FP32 far pairs,
FP64 analytic near pairs) with the inputs already in HBM; value = n(n+1)/2 / seconds / 1e9
(the reference's pair count, src/pcbqs/static.py:39). The dominant kernel's duration comes
from HIP events recorded on its launch stream. The CPU baseline is the reference's own
compiled static_fill (built into oracle/_ref) on the same mesh, on this host's cores.

The GMRES part solves the board's 4 ports at 1 MHz to a relative residual of 1e-8 with plain
block-Jacobi GMRES and with the learned two-level preconditioner (seconds, iterations, ratio),
reports the direct solve at the same frequency beside them (FP32 L D L^T on the device refined to
1e-12: factorisation time and rate, solve time) and at eight frequencies in one call (factored
and refined together), and times the plain path's operator product on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X
FP32_VALU_PEAK = 157.3e12  # flop/s, MI355X (vendor)
F_FULL = 209               # flop per far pair, full moments
B_FULL_FP32 = 32           # bytes stored per unordered pair: 4 planes x 2 entries x 4 B


def measured_traffic(seed):
    """HBM bytes per launch of the far kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload (profiles/far32_tile_c2_30k_traffic.json, seed 0), else None."""
    path = os.path.join(ROOT, 'profiles', 'far32_tile_c2_30k_traffic.json')
    if seed != 0 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)['traffic_bytes']


def cpu_baseline(board):
    """Time the reference's compiled static_fill (oracle/_ref) on the whole board, or, where it
    is not built, the C restatement (oracle/libcpu_ref.so), on this host's cores."""
    from oracle import fill as F
    try:
        cores = len(os.sched_getaffinity(0))
    except OSError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    g = F.prepare(board.nodes, board.tris, board.z_of, 1)
    n = len(board.tris)
    mats = [np.zeros((n, n)) for _ in range(4)]              # touched before the clock starts
    core = F.ref_core()
    t = time.perf_counter()
    if core is not None:
        count, analytic = core.static_fill(F.ref_geometry_dict(g), g.cent, g.radius, g.z, 3.0, 1, False,
                                           cores, mats)
        kind = 'reference'
    else:
        _, count, analytic = F.cr_fill(g, 3.0, 1, False, cores)
        kind = 'port'
    seconds = time.perf_counter() - t
    del mats
    return dict(value=count / seconds / 1e9, unit='Gpair/s', cores=cores, kind=kind,
                sample=f'the whole fill of the same {n}-triangle board: static_fill, {count:,} unordered '
                       f'pairs ({analytic:,} analytic), FP64, near=3, outer_levels=1, one run',
                seconds=seconds)


def gmres_section(board, freq=1e6, grid=36, repeats=3, cpu=True):
    """Port solve at one frequency on the same board: plain block-Jacobi GMRES against two-level
    GMRES whose coarse space holds the grid x grid hats and the learned enrichment. Seconds are
    the solver's own wall clock per solve, the median of the repeats at that frequency; the first
    repeat also builds the frequency's block-Jacobi inverses and coarse matrices and is reported as
    first_solve_s. The per-board set-up of each method is reported beside it."""
    from pcbqs_amd.learned import MPNN, set_learned_coarse
    from pcbqs_amd.solve import PortSolver
    t = time.perf_counter()
    solver = PortSolver(board, device=0)
    board_setup = time.perf_counter() - t

    def run(precond):
        out = [solver.ports([freq], precond=precond, tol=1e-8) for _ in range(repeats)]
        secs = [o[1][0]['seconds'] for o in out]
        i = int(np.argsort(secs)[len(secs) // 2])
        return out[i][0][0], dict(out[i][1][0], first_solve_s=secs[0])
    Zp, plain = run('block-jacobi')
    # the direct solve at the same frequency: the first call forms and factors Z_ff (FP32
    # L D L^T, n^3 / 6 complex multiply-adds = 4 n^3 / 3 flop), the repeats reuse the factor
    dout = [solver.ports([freq], solver='direct') for _ in range(repeats)]
    dinfo = [o[1][0] for o in dout]
    nf = solver.model.n_free
    factor_ms = dinfo[0]['factor_ms']
    direct = dict(
        method='FP32 L D L^T of Rl + j omega Ll (64-wide panels, MFMA trailing updates, no pivoting), '
               'FP64 refinement to 1e-12',
        first_solve_s=dinfo[0]['seconds'], seconds=float(np.median([d['seconds'] for d in dinfo[1:]]))
        if repeats > 1 else dinfo[0]['seconds'], factor_ms=factor_ms, corrections=dinfo[0]['iterations'],
        relative_residual=dinfo[0]['relative_residual'],
        factor_tflops=4 * nf ** 3 / 3 / (factor_ms * 1e-3) / 1e12 if factor_ms > 0 else None,
        fp32_mfma_peak_tflops=157.3,
        port_z_rel_diff_vs_plain=float(np.abs(dout[0][0][0] - Zp).max() / np.abs(Zp).max()))
    # eight frequencies (10 kHz .. 100 MHz) in one call: factored and refined together
    sweep_f = np.logspace(4, 8, 8)
    t = time.perf_counter()
    _, binfo = solver.ports(sweep_f, solver='direct')
    bsec = time.perf_counter() - t
    bms = binfo[0]['factor_ms']
    direct['batched'] = dict(frequencies=len(sweep_f), seconds=bsec, seconds_per_frequency=bsec / len(sweep_f),
                             factor_ms_per_frequency=bms,
                             factor_tflops=4 * nf ** 3 / 3 / (bms * 1e-3) / 1e12 if bms > 0 else None,
                             worst_relative_residual=max(r['relative_residual'] for r in binfo))
    t = time.perf_counter()
    net = MPNN.load()
    net_s = set_learned_coarse(solver, net, grid)
    coarse_setup = time.perf_counter() - t
    Zl, learned = run('two-level')
    section = dict(
        frequency_Hz=freq, ports=solver.n_ports, free_loops=solver.model.n_free, tol=1e-8,
        plain=dict(preconditioner='block-jacobi (256-loop blocks)', seconds=plain['seconds'],
                   iterations=plain['iterations'], relative_residual=plain['relative_residual'],
                   first_solve_s=plain['first_solve_s']),
        learned=dict(preconditioner=f'two-level: {grid}x{grid} hats + via loops + learned enrichment',
                     seconds=learned['seconds'], iterations=learned['iterations'],
                     first_solve_s=learned['first_solve_s'],
                     relative_residual=learned['relative_residual'], coarse_columns=solver.coarse_k,
                     board_setup_s=coarse_setup, network_s=net_s),
        ratio=learned['seconds'] / plain['seconds'],
        port_z_rel_diff=float(np.abs(Zl - Zp).max() / np.abs(Zp).max()),
        board_setup_s=board_setup, direct=direct)
    chunk, us = solver.tuning()
    n = solver.model.n_free
    section['operator_product'] = dict(
        kernel='operator product (autotuned panel_kernel or sym_panel_kernel + operator_epilogue_kernel: (Rl + j omega Ll) X, Ll FP32, 8 real FP64 columns)', autotuned_chunk=chunk,
        microseconds=us, timed='the whole product: Ll X chunk partials, the resistive CSR part and the epilogue',
        algorithmic_bytes=4 * n * n, achieved_GBs=4 * n * n / (us * 1e-6) / 1e9,
        peak_GBs=HBM_PEAK / 1e9, frac=4 * n * n / (us * 1e-6) / HBM_PEAK if us > 0 else None)
    if cpu:
        L, R = host_loop_matrices(solver)
        section['cpu_operator'] = cpu_operator_sample(solver, plain['iterations'], L=L, R=R)
        section['direct']['cpu'] = cpu_direct_sample(L, R, 2 * np.pi * freq)
        del L, R
    solver.close()
    return section


def sweep_section(n_boards, n_freqs=20, concurrency=1):
    """The batched sweep: n_boards seeded uneven boards (1k-8k triangles, 2-4 layers, 4-16 ports) x
    n_freqs log-spaced frequencies 1 kHz - 100 MHz, two-level GMRES to 1e-8."""
    from pcbqs_amd import board as boards
    from pcbqs_amd.sweep import sweep
    bs = [boards.board_batch(s) for s in range(n_boards)]
    res, total = sweep(bs, np.logspace(3, 8, n_freqs), precond='two-level', grid=12, concurrency=concurrency)
    total.update(setup_s=sum(r['setup_s'] for r in res), solve_s=sum(r['solve_s'] for r in res),
                 triangles=sum(r['triangles'] for r in res), mean_iterations=float(np.mean(
                     [i for r in res for i in r['iterations']])))
    return total


def host_loop_matrices(solver):
    """The free-loop block of the device's FP64 Ll and Rl (CSR) on the host."""
    n = solver.model.n_free
    L = np.ascontiguousarray(solver.loop_matrix(n, precision=64)[:, :n])
    return L, solver.model.Rl.tocsr()[:n, :n]


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except OSError:
        return os.cpu_count() or 1


def cpu_direct_sample(L, R, omega):
    """The reference's dense path on the host: LAPACK's single-precision complex LU (scipy's
    lu_factor, as MixedLU, src/pcbqs/kernels.py:143-148) of the same Z_ff, timed once."""
    from scipy.linalg import lu_factor
    n = L.shape[0]
    Z = np.asarray(R.todense(), np.complex64)
    Z += np.complex64(1j * omega) * L.astype(np.float32)
    t = time.perf_counter()
    lu_factor(Z.T, overwrite_a=True, check_finite=False)
    seconds = time.perf_counter() - t
    return dict(factor_s=seconds, cores=host_cores(), kind='reference method (LAPACK cgetrf through scipy)',
                sample=f'one FP32 complex LU of the {n} x {n} free-loop block at {omega / 2 / np.pi:g} Hz '
                       f'(the refinement, a few solves and products, is not timed)')


def cpu_operator_sample(solver, iterations, products=5, L=None, R=None):
    """The plain path's operator product (Z_ff on 2P real columns, FP64 numpy on the host cores)
    timed over a few products and scaled to the plain solve's iteration count: a bounded sample of
    the CPU GMRES, whose time is ~98% operator products."""
    n, P = solver.model.n_free, solver.n_ports
    if L is None:
        L, R = host_loop_matrices(solver)
    X = np.random.default_rng(0).normal(size=(n, 2 * P))
    omega = 2 * np.pi * 1e6
    L @ X
    t = time.perf_counter()
    for _ in range(products):
        Y = R @ X
        Y[:, :P] -= omega * (L @ X[:, P:])
        Y[:, P:] += omega * (L @ X[:, :P])
    per = (time.perf_counter() - t) / products
    return dict(seconds_per_product=per, estimated_solve_s=per * iterations, cores=host_cores(), kind='port',
                sample=f'{products} FP64 operator products (n={n}, {2 * P} real columns) with numpy, '
                       f'scaled to the {iterations} iterations of the plain solve')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-gmres', action='store_true')
    ap.add_argument('--sweep', type=int, default=0, metavar='BOARDS',
                    help='also run the batched sweep over BOARDS uneven boards x 20 frequencies')
    ap.add_argument('--sweep-concurrency', type=int, default=4, help='boards at once in the sweep')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='BYTES',
                    help='first copy BYTES device-to-device (the scale of a counter run)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps(dict(error='this build runs on one GPU: every box here has one; '
                                    'the multi-GPU split is described, not run')))
        sys.exit(2)
    from pcbqs_amd import board as boards
    from pcbqs_amd.backend import Mi355xBackend
    from pcbqs_amd.geometry import FillGeometry
    if args.copy_probe:
        from pcbqs_amd._lib import DeviceArray, call
        src, dst = DeviceArray(args.copy_probe, np.uint8), DeviceArray(args.copy_probe, np.uint8)
        call('pq_memcpy_d2d', dst.ptr, src.ptr, args.copy_probe)
        src.close()
        dst.close()
    board = boards.board_c2(args.seed)
    n = board.nt
    pairs = n * (n + 1) // 2
    g = FillGeometry.prepare(board.nodes, board.tris, board.z_of, 1)
    backend = Mi355xBackend(g, 1, 16384, 0)
    planes, analytic = backend.fill_device(3.0, 1, False, 32)
    for _ in range(args.warmup):
        backend.fill_device(3.0, 1, False, 32, out=planes)
    far_ms, near_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        backend.fill_device(3.0, 1, False, 32, out=planes)
        f, nm, near_pairs = backend.last_timing()
        far_ms.append(f)
        near_ms.append(nm)
    seconds = (time.perf_counter() - t0) / args.steps
    far = float(np.mean(far_ms)) * 1e-3
    # the store ceiling of the same output pattern (the four planes' 64 x 64 tiles, no arithmetic)
    import ctypes
    from pcbqs_amd._lib import call
    probe_ms = ctypes.c_double()
    call('pq_store_probe', planes.ptr, n, planes.shape[-1], 4, ctypes.byref(probe_ms))
    store_probe = 16.0 * n * n / (probe_ms.value * 1e-3)
    # FP64 reference-layout fill (the parity mode), once, for the record
    planes.close()
    p64, _ = backend.fill_device(3.0, 1, False, 64)
    backend.fill_device(3.0, 1, False, 64, out=p64)
    f64, n64, _ = backend.last_timing()
    p64.close()
    achieved = pairs * B_FULL_FP32 / far
    line = dict(
        metric='fill Gpair/s; GMRES s to 1e-8 at 30k RWG, learned vs plain precond',
        value=pairs / seconds / 1e9, unit='Gpair/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=seconds * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None,
        dtype='fp32 far / fp64 near', data='synthetic',
        config=dict(workload=f'seeded 2-layer 60x60 mm plane pair (seed {args.seed}), {n} triangles, '
                             f'{board.info.get("cutouts")} cutouts, {len(board.vseg)} vias, 4 ports; '
                             f'near=3, outer_levels=1', triangles=n, pairs=pairs, analytic_pairs=analytic),
        roofline=dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                      frac=achieved / HBM_PEAK, traffic=measured_traffic(args.seed),
                      algorithmic_bytes=pairs * B_FULL_FP32,
                      kernel='far32_tile_kernel (FP32 far pairs, full moments, both reciprocal entries)',
                      valu_frac=pairs * F_FULL / far / FP32_VALU_PEAK,
                      store_probe_GBs=store_probe / 1e9, frac_of_store_probe=achieved / store_probe),
        kernels=dict(far_ms=float(np.mean(far_ms)), near_ms=float(np.mean(near_ms)),
                     near_pairs=list(near_pairs), fp64_far_ms=f64, fp64_near_ms=n64,
                     fp64_gpair_s=pairs / ((f64 + n64) * 1e-3) / 1e9),
    )
    if not args.no_gmres:
        line['gmres'] = gmres_section(board, cpu=not args.no_cpu)
    if args.sweep:
        line['sweep'] = sweep_section(args.sweep, concurrency=args.sweep_concurrency)
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(board)
    print(json.dumps(line))


if __name__ == '__main__':
    main()
