"""Benchmark: fill Gpair/s and the port GMRES solve on the 30k-RWG seeded plane pair.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--seed S] [--no-cpu] [--no-gmres] [--sweep BOARDS]

A step is one device-resident fill of the board's four moment planes (FP32 far pairs, FP64
analytic near pairs) with the inputs already in HBM; value = n(n+1)/2 / seconds / 1e9 (the
reference's pair count, src/pcbqs/static.py:39). The same loop over the plain FP64 fill (the
reference's precision and layout) is reported beside it (`fp64_fill`). The dominant kernel's
duration comes from HIP events recorded on its launch stream.

The GMRES part solves the board's 4 ports at 1 MHz to a relative residual of 1e-8 (the FP64
system's true residual) with plain block-Jacobi GMRES, with two-level GMRES whose coarse space is
the stream-function hats alone, and with the hats plus the learned enrichment; it reports each
method's set-up and an end-to-end figure (board set-up + the method's set-up + a 20-frequency
sweep), and the direct solve beside them.

The CPU leg (skipped with --no-cpu) times the oracle's FP64 fill (oracle/cpu_ref.c, bitwise the
reference core) on this host's cores, one warm-up and the median of three runs (cpu_baseline), and
the plain-preconditioned GMRES restated in numpy on the same FP64 system (cpu_gmres); it then uses
the oracle as the checker: every fill's per-plane error and every port solve's R and L error
against the oracle's FP64 chain (fill, combine, C^T L C, LU).

The compressed part (skipped with --no-compressed) builds the board's RWG inductance as an H-matrix
on the device (tol 1e-6, leaf 64) and, in the CPU leg, times the reference core's own native
construction of the same operator on every host core and compares the two operators' products.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X
FP32_VALU_PEAK = 157.3e12  # flop/s, MI355X (vendor)
FP64_VALU_PEAK = 78.6e12
F_FULL = 209               # flop per far pair, full moments
PLANES = ('M0', 'Mrx', 'Mry', 'Mrr')
FREQ = 1e6
SWEEP_FREQS = np.logspace(3, 8, 20)


def measured_traffic(seed):
    """HBM bytes per launch of the far kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload (profiles/far32_tile_c2_30k_traffic.json: seed 0 only), else None."""
    path = os.path.join(ROOT, 'profiles', 'far32_tile_c2_30k_traffic.json')
    if seed != 0 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)['traffic_bytes']


def sync():
    from pcbqs_amd._lib import call
    call('pq_synchronize', 0)


# ---- fills -------------------------------------------------------------------------------------

def timed_fills(backend, precision, steps, warmup):
    """Step-timed device fills of one precision: (planes, dict of rates and kernel times)."""
    n = backend.nt
    pairs = n * (n + 1) // 2
    planes, analytic = backend.fill_device(3.0, 1, False, precision)
    for _ in range(warmup):
        backend.fill_device(3.0, 1, False, precision, out=planes)
    sync()
    far_ms, near_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        backend.fill_device(3.0, 1, False, precision, out=planes)
        f, nm, near_pairs = backend.last_timing()
        far_ms.append(f)
        near_ms.append(nm)
    sync()
    seconds = (time.perf_counter() - t0) / steps
    far = float(np.mean(far_ms)) * 1e-3
    stored = pairs * 8 * (precision // 8)            # 4 planes x 2 reciprocal entries
    return planes, dict(gpair_s=pairs / seconds / 1e9, ms_per_step=seconds * 1e3, far_ms=far * 1e3,
                        near_ms=float(np.mean(near_ms)), near_pairs=list(near_pairs), analytic=analytic,
                        far_store_GBs=stored / far / 1e9, far_hbm_frac=stored / far / HBM_PEAK,
                        algorithmic_bytes=stored)


def store_probe(n, ld):
    """The store ceiling of the far kernel's output pattern (4 FP32 planes, 64 x 64 tiles, no
    arithmetic), on a scratch buffer of the planes' shape: bytes/s."""
    from pcbqs_amd._lib import DeviceArray, call
    buf = DeviceArray((4, n, ld), np.float32)
    ms = ctypes.c_double()
    call('pq_store_probe', buf.ptr, n, ld, 4, ctypes.byref(ms))
    buf.close()
    return 16.0 * n * n / (ms.value * 1e-3)


def drop_in_fill(backend):
    """The plugin's fill_symmetric (what `--fill-backend mi355x` callers get): four host FP64 (n, n)
    matrices, so its rate includes the device-to-host copy of 8 B x 4 x n^2."""
    n = backend.nt
    mats = [np.zeros((n, n)) for _ in range(4)]              # pages touched before the clock
    t = time.perf_counter()
    backend.fill_symmetric(mats, 3.0, 1, False)
    seconds = time.perf_counter() - t
    del mats
    pairs = n * (n + 1) // 2
    return dict(seconds=seconds, gpair_s_pcie_inclusive=pairs / seconds / 1e9, host_bytes=32 * n * n,
                note='one call, FP64 planes into caller-owned host memory (device-to-host copy included)')


# ---- port solves -------------------------------------------------------------------------------

def median_solve(solver, repeats, **kw):
    out = [solver.ports([FREQ], **kw) for _ in range(repeats)]
    secs = [o[1][0]['seconds'] for o in out]
    i = int(np.argsort(secs)[len(secs) // 2])
    return out[i][0][0], dict(out[i][1][0], first_solve_s=secs[0])


def sweep_seconds(solver, **kw):
    t = time.perf_counter()
    _, info = solver.ports(SWEEP_FREQS, **kw)
    return time.perf_counter() - t, int(sum(i['iterations'] for i in info))


def gmres_section(board, grid=36, repeats=3):
    """Port solves at FREQ on the board: plain block-Jacobi GMRES, two-level GMRES with the hats
    alone and with the hats plus the learned enrichment, and the direct solve. Seconds per solve are
    the solver's own clock, the median of the repeats (the first repeat also builds the frequency's
    block-Jacobi inverses and coarse matrices: first_solve_s). end_to_end: board set-up + the
    block-Jacobi set-up + the method's set-up + a 20-frequency sweep (1 kHz - 100 MHz). Returns
    (section, Z per method, solver)."""
    from pcbqs_amd import board as boards
    from pcbqs_amd.learned import MPNN, BranchGraph, enrichment
    from pcbqs_amd.solve import PortSolver
    # process warm-up, not timed: a process's first rocBLAS / rocSOLVER calls load their code objects
    # (~0.4 s); a small board's block-Jacobi and coarse set-ups take that before the timed board
    t = time.perf_counter()
    warm = PortSolver(boards.board_c2(1, target=3000), device=0)
    warm.prepare_block_jacobi()
    warm.set_stream_coarse(24)
    warm.close()
    warm_s = time.perf_counter() - t
    t = time.perf_counter()
    solver = PortSolver(board, device=0)
    board_setup = time.perf_counter() - t
    # block-Jacobi's per-board set-up (every iterative method below uses it): the 256-loop blocks'
    # generalised eigenvectors, after which each frequency's block inverses are one batched GEMM
    t = time.perf_counter()
    bj_route = solver.prepare_block_jacobi()
    bj_setup = time.perf_counter() - t
    Z = {}
    Z['plain'], plain = median_solve(solver, repeats, precond='block-jacobi', tol=1e-8)
    plain_sweep = sweep_seconds(solver, precond='block-jacobi', tol=1e-8)
    # hats alone (set-up timed as the median of the repeats; the first call, which also loads
    # rocBLAS/rocSOLVER, is reported beside it)
    hats_times = []
    for _ in range(repeats):
        t = time.perf_counter()
        solver.set_stream_coarse(grid)
        hats_times.append(time.perf_counter() - t)
    hats_setup = float(np.median(hats_times))
    hats_k, hats_path = solver.coarse_k, solver.coarse_path
    Z['hats'], hats = median_solve(solver, repeats, precond='two-level', tol=1e-8)
    hats_sweep = sweep_seconds(solver, precond='two-level', tol=1e-8)
    # hats + learned enrichment: the network's forward pass and features, then the coarse space
    t = time.perf_counter()
    net = MPNN.load()
    load_s = time.perf_counter() - t
    learned_times, net_times = [], []
    for _ in range(repeats):
        t = time.perf_counter()
        extra = enrichment(net, BranchGraph(solver.model))
        net_times.append(time.perf_counter() - t)
        solver.set_stream_coarse(grid, extra_loops=extra)
        learned_times.append(time.perf_counter() - t)
    mid = int(np.argsort(learned_times)[len(learned_times) // 2])
    learned_setup, net_s = learned_times[mid], net_times[mid]
    Z['learned'], learned = median_solve(solver, repeats, precond='two-level', tol=1e-8)
    learned_sweep = sweep_seconds(solver, precond='two-level', tol=1e-8)

    def method(name, info, setup, sw):
        return dict(preconditioner=name, seconds=info['seconds'], iterations=info['iterations'],
                    relative_residual=info['relative_residual'], first_solve_s=info['first_solve_s'],
                    setup_s=setup, end_to_end_s=board_setup + bj_setup + setup + sw[0], sweep_s=sw[0],
                    sweep_iterations=sw[1])
    section = dict(
        frequency_Hz=FREQ, ports=solver.n_ports, free_loops=solver.model.n_free, tol=1e-8,
        residual='true relative residual of the FP64 system (Ll kept in FP64; FP32 copy for the Krylov products)',
        board_setup_s=board_setup, sweep_frequencies=len(SWEEP_FREQS),
        block_jacobi=dict(setup_s=bj_setup, route=bj_route,
                          note='per board, shared by the three iterative methods and counted in each one\'s '
                               'end_to_end_s; a frequency\'s block inverses are then one batched GEMM'),
        process_warmup_s=warm_s,
        plain=method('block-jacobi (256-loop blocks)', plain, 0.0, plain_sweep),
        hats=method(f'two-level: {grid}x{grid} hats + via loops', hats, hats_setup, hats_sweep),
        learned=method(f'two-level: {grid}x{grid} hats + via loops + learned enrichment', learned, learned_setup,
                       learned_sweep),
        coarse_columns=dict(hats=hats_k, learned=solver.coarse_k))
    section['hats'].update(first_setup_s=hats_times[0], coarse_route=hats_path)
    section['learned'].update(network_s=net_s, first_setup_s=learned_times[0], weights_load_s=load_s,
                              coarse_route=solver.coarse_path,
                              setup_note='per-board set-up, median of the repeats: the network and its features '
                                         '(network_s), the hats, the sparse upload and the device diagonalisation')
    section['ratio'] = learned['seconds'] / plain['seconds']
    section['ratio_hats'] = hats['seconds'] / plain['seconds']
    section['ratio_end_to_end'] = section['learned']['end_to_end_s'] / section['plain']['end_to_end_s']
    # the direct solve at the same frequency: the first call forms and factors Z_ff (FP32
    # L D L^T, n^3 / 6 complex multiply-adds = 4 n^3 / 3 flop), the repeats reuse the factor
    dout = [solver.ports([FREQ], solver='direct') for _ in range(repeats)]
    dinfo = [o[1][0] for o in dout]
    Z['direct'] = dout[0][0][0]
    nf = solver.model.n_free
    factor_ms = dinfo[0]['factor_ms']
    direct = dict(
        method='FP32 L D L^T of Rl + j omega Ll (64-wide panels, MFMA trailing updates, no pivoting), '
               'refined against the FP64 system to 1e-12 (FP64 LU fallback where that stalls)',
        first_solve_s=dinfo[0]['seconds'], seconds=float(np.median([d['seconds'] for d in dinfo[1:]])),
        factor_ms=factor_ms, corrections=dinfo[0]['iterations'], relative_residual=dinfo[0]['relative_residual'],
        fp64_fallback=dinfo[0]['fp64_fallback'],
        factor_tflops=4 * nf ** 3 / 3 / (factor_ms * 1e-3) / 1e12 if factor_ms > 0 else None,
        fp32_mfma_peak_tflops=157.3)
    t = time.perf_counter()
    _, binfo = solver.ports(SWEEP_FREQS, solver='direct')
    bsec = time.perf_counter() - t
    bms = binfo[0]['factor_ms']
    direct['batched'] = dict(frequencies=len(SWEEP_FREQS), seconds=bsec, seconds_per_frequency=bsec / len(SWEEP_FREQS),
                             end_to_end_s=board_setup + bsec, factor_ms_per_frequency=bms,
                             factor_tflops=4 * nf ** 3 / 3 / (bms * 1e-3) / 1e12 if bms > 0 else None,
                             worst_relative_residual=max(r['relative_residual'] for r in binfo))
    section['direct'] = direct
    section['graph'] = graph_gmres(solver)
    chunk, us = solver.tuning()
    n = solver.model.n_free
    section['operator_product'] = dict(
        kernel='operator product (autotuned panel_kernel or sym_panel_kernel + operator_epilogue_kernel: '
               '(Rl + j omega Ll) X, Ll FP32, 8 real FP64 columns)', autotuned_chunk=chunk,
        microseconds=us, timed='the whole product: Ll X chunk partials, the resistive CSR part and the epilogue',
        algorithmic_bytes=4 * n * n, achieved_GBs=4 * n * n / (us * 1e-6) / 1e9,
        peak_GBs=HBM_PEAK / 1e9, frac=4 * n * n / (us * 1e-6) / HBM_PEAK if us > 0 else None)
    return section, Z, solver


def graph_gmres(solver):
    """The reference's own iterative preconditioner on the same system: host GMRES (scipy, restart
    100, rtol 1e-8) on the device FP64 loop operator with the graph preconditioner (C_f^T D C_f)^-1
    (src/pcbqs/mqs_iterative.py:337-341; its node Laplacian factored and solved on the device), for
    the first port's right-hand side at FREQ. Its iteration count sits beside block-Jacobi's and the
    two-level methods'; its seconds include every operator product's PCIe round trip."""
    from scipy.sparse.linalg import LinearOperator, gmres
    n = solver.model.n_free
    t = time.perf_counter()
    M, nbytes = solver.graph_preconditioner(FREQ)
    setup = time.perf_counter() - t
    b = solver.rhs(FREQ)[:, 0]
    A = LinearOperator((n, n), matvec=lambda x: solver.operator(x, FREQ, precision=64).ravel(), dtype=complex)
    its = []
    t = time.perf_counter()
    x, status = gmres(A, b, M=M, rtol=1e-8, restart=100, maxiter=5, callback=lambda r: its.append(r),
                      callback_type='pr_norm')
    seconds = time.perf_counter() - t
    res = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    M.laplacian.close()
    return dict(method='host GMRES (scipy, restart 100) with the graph preconditioner (node Laplacian factored '
                       'on the device), device FP64 operator products, one port', iterations=len(its),
                converged=status == 0, true_relative_residual=res, seconds=seconds, setup_s=setup,
                factor_bytes=int(nbytes))


def fp64_pipeline_ports(board):
    """Config 1's plain FP64 method on this board: FP64 planes, FP64 branch inductance and loop
    matrix, the direct solve at FREQ: (Z, seconds of set-up + solve, info)."""
    from pcbqs_amd.solve import PortSolver
    t = time.perf_counter()
    s = PortSolver(board, device=0, precision=64)
    Z, info = s.ports([FREQ], solver='direct')
    seconds = time.perf_counter() - t
    s.close()
    return Z[0], seconds, info[0]


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


# ---- the CPU leg: baselines on the host cores, then the oracle as the checker -------------------

def host_cpus():
    """(CPUs in the affinity mask, the cgroup CPU quota or None, the CPU model)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except OSError:
        affinity = os.cpu_count() or 1
    quota = None
    try:
        with open('/sys/fs/cgroup/cpu.max') as f:
            q, period = f.read().split()[:2]
        if q != 'max':
            quota = max(1, int(round(int(q) / int(period))))
    except (OSError, ValueError):
        pass
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return affinity, quota, model


def cpu_fill_baseline(board):
    """The oracle's FP64 fill of the whole board (oracle/cpu_ref.c, the reference core's pair code,
    64 x 64 dynamic OpenMP tiles) on every CPU of the affinity mask, and also at the cgroup quota's
    CPU count where that is smaller: one warm-up and the median of three runs each; the faster
    thread count is the baseline. Returns (baseline dict, the last run's planes)."""
    from oracle import fill as F
    affinity, quota, model = host_cpus()
    g = F.prepare(board.nodes, board.tris, board.z_of, 1)
    n = len(board.tris)
    counts = sorted({affinity} | ({quota} if quota and quota < affinity else set()))
    runs, planes = {}, None
    for threads in counts:
        F.cr_fill(g, 3.0, 1, False, threads)                     # warm-up
        secs = []
        for _ in range(3):
            t = time.perf_counter()
            planes, count, analytic = F.cr_fill(g, 3.0, 1, False, threads)
            secs.append(time.perf_counter() - t)
        runs[threads] = float(np.median(secs))
    best = min(runs, key=runs.get)
    out = dict(value=count / runs[best] / 1e9, unit='Gpair/s', cores=best, kind='port',
               sample=f'the whole fill of the same {n}-triangle board: {count:,} unordered pairs ({analytic:,} '
                      f'analytic), FP64, near=3, outer_levels=1; one warm-up, median of 3 runs',
               seconds=runs[best], cpu_model=model, affinity_cpus=affinity, cgroup_quota_cpus=quota,
               median_seconds_by_threads={str(k): v for k, v in runs.items()},
               binary='oracle/libcpu_ref.so, built from the tracked oracle/cpu_ref.c (bitwise equal to the '
                      'reference core: tests/test_oracle.py)')
    core = F.ref_core()
    if core is not None:                                          # calibration against the reference core
        t = time.perf_counter()
        F.ref_fill(core, g, 3.0, 1, False, best)
        out['reference_core_seconds'] = time.perf_counter() - t
    return out, planes


def plane_errors(dev_planes, ref_planes, local_cent=None, block=2048):
    """Per-plane relative Frobenius errors of device planes (4, n, ld) against the oracle's (n, n)
    FP64 planes; FP32 local-frame planes (local_cent given) are first taken to the reference layout."""
    n = ref_planes[0].shape[0]
    host = dev_planes.host()[:, :, :n]
    num, den = np.zeros(4), np.zeros(4)
    for r0 in range(0, n, block):
        r1 = min(n, r0 + block)
        b = [host[k][r0:r1].astype(np.float64) for k in range(4)]
        if local_cent is not None:
            cx, cy = local_cent[:, 0], local_cent[:, 1]
            ax, ay = cx[r0:r1, None], cy[r0:r1, None]
            mxt = host[1][:, r0:r1].T.astype(np.float64)
            myt = host[2][:, r0:r1].T.astype(np.float64)
            b = [b[0], b[1] + ax * b[0], b[2] + ay * b[0],
                 b[3] + ax * mxt + ay * myt + cx[None, :] * b[1] + cy[None, :] * b[2] +
                 (ax * cx[None, :] + ay * cy[None, :]) * b[0]]
        for k in range(4):
            ref = ref_planes[k][r0:r1]
            num[k] += np.sum((b[k] - ref) ** 2)
            den[k] += np.sum(ref ** 2)
    del host
    return dict(zip(PLANES, np.sqrt(num / den).tolist()))


def oracle_ports(board, ref_planes, freq):
    """The oracle's FP64 dense chain from its own fill planes: port Z (P, P) at freq."""
    from oracle import fill as F
    from oracle import mqs as M
    q = M.CopperModel(board.nodes, board.tris, board.z_of, board.t_of, board.node_of, board.terminals, board.vseg,
                      board.via_nodes, board.ports)
    g = F.prepare(board.nodes, board.tris, board.z_of, 1)
    q.assemble((g, *ref_planes))
    return q.port_response([freq])[0][0]


def port_errors(Z, Zo, freq):
    w = 2 * np.pi * freq
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    return dict(port_R=rel(Z.real, Zo.real), port_L=rel(Z.imag / w, Zo.imag / w))


def cpu_gmres(solver, iterations_gpu):
    """The plain-preconditioned solve restated in numpy (oracle/gmres.py: lockstep GMRES(100) of
    mqs_iterative._lockstep_gmres, block-Jacobi on 256-loop blocks, tol 1e-8, true-residual
    acceptance of peec_iterative._gmres) on the same FP64 system the device solves, all 4 ports,
    timed once on the host cores; and the reference's dense method's FP32 LU of the same block."""
    from oracle import gmres as G
    n = solver.model.n_free
    L = np.ascontiguousarray(solver.loop_matrix(n, precision=64)[:, :n])
    R = solver.model.Rl.tocsr()[:n, :n]
    B = solver.rhs(FREQ)
    omega = 2 * np.pi * FREQ
    t = time.perf_counter()
    X, its, res = G.plain_solve(L, R, B, omega, tol=1e-8, restart=100)
    seconds = time.perf_counter() - t
    affinity, quota, model = host_cpus()
    out = dict(seconds=seconds, iterations=its, relative_residual=res, gpu_iterations=iterations_gpu,
               cores=affinity, cpu_model=model, kind='port',
               sample=f'the whole plain solve of the {n}-loop system at {FREQ:g} Hz, {B.shape[1]} ports in lockstep, '
                      f'numpy FP64 (BLAS threads on the host cores)')
    from scipy.linalg import lu_factor
    Z = np.asarray(R.todense(), np.complex64)
    Z += np.complex64(1j * omega) * L.astype(np.float32)
    del L
    t = time.perf_counter()
    lu_factor(Z.T, overwrite_a=True, check_finite=False)
    out['reference_method_factor_s'] = time.perf_counter() - t
    out['reference_method'] = ("MixedLU's FP32 complex LU (LAPACK cgetrf through scipy) of the same block, "
                               'timed once (its refinement is not timed)')
    return out


def compressed_section(board, tol=1e-6, leaf=64, reps=20):
    """The compressed RWG inductance of the board built on the device (entry blocks, checked ACA,
    dense leaves; no dense matrix) and its product with 8 columns."""
    from pcbqs_amd.hmatrix import RwgEntries, build
    E = RwgEntries(board.nodes, board.tris, board.z_of)
    build(E, tol=tol, leaf=leaf).close()                          # warm-up (module load, first launches)
    E.close()
    t = time.perf_counter()
    E = RwgEntries(board.nodes, board.tris, board.z_of)
    provider_s = time.perf_counter() - t
    D = build(E, tol=tol, leaf=leaf)
    info = D.info
    x = np.random.default_rng(0).standard_normal((E.n, 8))
    y = D @ x
    out = dict(rwg=E.n, tol=tol, leaf=leaf, provider_s=provider_s, build_s=info['build_seconds'],
               aca_device_ms=info['aca_ms'], dense_device_ms=info['dense_ms'], blocks=info['blocks'],
               low_rank_blocks=info['low_rank_blocks'], bytes=info['bytes'], dense_bytes=info['dense_bytes'],
               evaluated_entries=info['evaluated_entries'], product_ms_8_columns=D.time_product(8, reps),
               note='device construction: the provider (geometry upload) and the build (tree, ACA, dense leaves, '
                    'assembly); the product is the H-matrix times 8 real columns, vectors in HBM')
    return out, D, E, x, y


def reference_core_build(board, tol=1e-6, leaf=64):
    """The reference core's own native construction of the same operator (RwgEntries, aca_blocks,
    dense_blocks; hmatrix_build.hpp), driven level by level as hmatrix.py:179-210 does, on every
    host core: (seconds, blocks, threads) or None where oracle/_ref is not built."""
    from oracle import fill as F
    from oracle import hmatrix as H
    core = F.ref_core()
    if core is None:
        return None
    g = F.prepare(board.nodes, board.tris, board.z_of, 1)
    sup, vx, al = H.rwg_arrays(g)
    lo, hi = H.edge_boxes(g, sup)
    affinity, quota, _ = host_cpus()
    threads = quota if quota and quota < affinity else affinity
    t = time.perf_counter()
    provider = core.RwgEntries(F.ref_geometry_dict(g), g.z, g.cent, g.radius, 3.0, 1, sup,
                               np.ascontiguousarray(vx), np.ascontiguousarray(al), H.MU0 / (4 * np.pi))
    blocks, pending = [], [(H.cluster_tree(lo, hi, leaf),) * 2]
    while pending:
        adm, dense, following = [], [], []
        for a, c in pending:
            if H.admissible(a, c, 1.0):
                adm.append((a, c))
            elif not a.children and not c.children:
                dense.append((a, c))
            else:
                following.extend(H.children(a, c))
        if adm:
            factors, _, _ = core.aca_blocks(provider, [a.ids.astype(np.int64) for a, _ in adm],
                                            [c.ids.astype(np.int64) for _, c in adm], tol, threads)
            for (a, c), f in zip(adm, factors):
                if f is not None:
                    blocks.append((a.ids, c.ids, f[0], f[1], False))
                elif not a.children and not c.children:
                    dense.append((a, c))
                else:
                    following.extend(H.children(a, c))
        if dense:
            for (a, c), blk in zip(dense, core.dense_blocks(provider, [a.ids.astype(np.int64) for a, _ in dense],
                                                            [c.ids.astype(np.int64) for _, c in dense], threads)):
                blocks.append((a.ids, c.ids, (blk + blk.T) / 2 if a is c else blk, None, a is c))
        pending = following
    return time.perf_counter() - t, blocks, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-gmres', action='store_true')
    ap.add_argument('--no-compressed', action='store_true', help='skip the compressed-operator build')
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
    planes32, f32 = timed_fills(backend, 32, args.steps, args.warmup)
    probe = store_probe(n, planes32.shape[-1])
    planes64, f64 = timed_fills(backend, 64, args.steps, args.warmup)
    far = f32['far_ms'] * 1e-3
    achieved = pairs * 32 / far
    line = dict(
        metric='fill Gpair/s; GMRES s to 1e-8 at 30k RWG, learned vs plain precond',
        value=f32['gpair_s'], unit='Gpair/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=f32['ms_per_step'], higher_is_better=True, scaling='weak', vs_baseline=None,
        dtype='fp32 far / fp64 near', data='synthetic',
        config=dict(workload=f'seeded 2-layer 60x60 mm plane pair (seed {args.seed}), {n} triangles, '
                             f'{board.info.get("cutouts")} cutouts, {len(board.vseg)} vias, 4 ports; '
                             f'near=3, outer_levels=1', triangles=n, pairs=pairs, analytic_pairs=f32['analytic']),
        roofline=dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                      frac=achieved / HBM_PEAK, traffic=measured_traffic(args.seed),
                      traffic_source='rocprofv3 FETCH_SIZE + WRITE_SIZE of the far kernel on this workload (seed 0), '
                                     'scaled by a copy kernel of known size: profiles/far32_tile_c2_30k_traffic.json',
                      algorithmic_bytes=pairs * 32,
                      kernel='far32_tile_kernel (FP32 far pairs, full moments, both reciprocal entries)',
                      valu_frac=pairs * F_FULL / far / FP32_VALU_PEAK,
                      store_probe_GBs=probe / 1e9, frac_of_store_probe=achieved / probe),
        kernels=dict(far_ms=f32['far_ms'], near_ms=f32['near_ms'], near_pairs=f32['near_pairs']),
        fp64_fill=dict(f64, dtype='fp64 (the reference precision and layout)',
                       kernel='far64 kernel + FP64 near kernel; far stores 64 B per unordered pair',
                       valu_frac_fp64=pairs * F_FULL / (f64['far_ms'] * 1e-3) / FP64_VALU_PEAK),
    )
    line['drop_in'] = drop_in_fill(backend)
    backend.close()
    Z, solver = None, None
    if not args.no_gmres:
        line['gmres'], Z, solver = gmres_section(board)
        Z['fp64_pipeline'], s64, i64 = fp64_pipeline_ports(board)
        line['gmres']['fp64_pipeline'] = dict(method='FP64 planes, FP64 Lb and Ll, direct solve', seconds=s64,
                                              relative_residual=i64['relative_residual'])
    if args.sweep:
        line['sweep'] = sweep_section(args.sweep, concurrency=args.sweep_concurrency)
    hm = None
    if not args.no_compressed:
        line['compressed'], *hm = compressed_section(board)
    if not args.no_cpu:
        line['cpu_baseline'], ref_planes = cpu_fill_baseline(board)
        line['fill_rel_err_vs_fp64'] = plane_errors(planes32, ref_planes, local_cent=g.cent)
        line['fp64_fill_rel_err_vs_oracle'] = plane_errors(planes64, ref_planes)
        if Z is not None:
            Zo = oracle_ports(board, ref_planes, FREQ)
            errs = {k: port_errors(v, Zo, FREQ) for k, v in Z.items()}
            line['port_L_rel_err_vs_oracle'] = {k: v['port_L'] for k, v in errs.items()}
            line['port_R_rel_err_vs_oracle'] = {k: v['port_R'] for k, v in errs.items()}
        del ref_planes
        if hm and reference_core_build(board) is not None:        # the first run is the warm-up
            from pcbqs_amd.hmatrix import DeviceHMatrix
            runs = [reference_core_build(board) for _ in range(3)]
            seconds, blocks, threads = sorted(runs, key=lambda r: r[0])[1]
            D, E, x, y = hm
            R = DeviceHMatrix(blocks, E.n)
            yr = R @ x
            R.close()
            line['compressed']['cpu_reference_build'] = dict(
                seconds=seconds, threads=threads, kind='reference', blocks=len(blocks),
                sample='the whole construction of the same operator by the reference core (oracle/_ref, built from '
                       'its sources): RwgEntries, aca_blocks and dense_blocks level by level, every host core; '
                       'one warm-up, the median of three runs')
            line['compressed']['cpu_over_gpu_build'] = seconds / line['compressed']['build_s']
            line['compressed']['product_rel_err_vs_reference_build'] = float(
                np.linalg.norm(y - yr) / np.linalg.norm(yr))
        if solver is not None:
            line['cpu_gmres'] = cpu_gmres(solver, line['gmres']['plain']['iterations'])
            line['cpu_gmres']['cpu_over_gpu_plain'] = line['cpu_gmres']['seconds'] / line['gmres']['plain']['seconds']
    planes32.close()
    planes64.close()
    if solver is not None:
        solver.close()
    if hm:
        hm[0].close()
        hm[1].close()
    print(json.dumps(line))


if __name__ == '__main__':
    main()
