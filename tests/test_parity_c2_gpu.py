"""Parity at config-2 size: the seeded 60 x 60 mm plane pair board_c2(0) (20,336 triangles,
30,023 RWG edges, 9,639 free loops, 4 ports), the board every bench figure is measured on.

- The plain FP64 device fill against the oracle's fill (oracle/cpu_ref.c, bitwise the reference
  core): per-plane relative Frobenius error <= 1e-10 and the same analytic pair count.
- The FP32-bulk fill (FP32 far pairs in local frames, FP64 near pairs) against the same FP64
  planes, per plane after conversion to the reference layout: <= 2e-6.
- Port R and L: the FP32-bulk pipeline (lockstep GMRES with block-Jacobi to 1e-8, and the direct
  solve refined to 1e-12) within 1e-6 of the oracle's FP64 dense chain (fill, combine, C^T L C,
  LU per frequency); the plain FP64 pipeline within 1e-9.
- The residuals the solves report are those of the FP64 system: recomputed on the host with the
  device's FP64 loop matrix, they meet the solves' tolerances.
The oracle chain runs once per module on the host cores (about 20 s on 8 cores)."""
import numpy as np
import pytest

from oracle import fill as F
from oracle import mqs as M

pytestmark = pytest.mark.gpu
FREQS = [1e3, 1e6]


@pytest.fixture(scope='module')
def c2():
    from pcbqs_amd import board as boards
    b = boards.board_c2(0)
    print(f'c2 oracle: {b.nt} triangles', flush=True)
    q = M.CopperModel(b.nodes, b.tris, b.z_of, b.t_of, b.node_of, b.terminals, b.vseg, b.via_nodes, b.ports)
    g = F.prepare(b.nodes, b.tris, b.z_of, 1)
    (m0, mx, my, rr), _, analytic = F.cr_fill(g, 3.0, 1, False)
    q.assemble((g, m0, mx, my, rr))
    Zo = q.port_response(FREQS)[0]
    print('c2 oracle done', flush=True)
    return dict(board=b, planes=(m0, mx, my, rr), analytic=analytic, Zo=Zo)


def plane_errors(host_rows, ref_planes, nt, convert=None, block=2048):
    """Per-plane relative Frobenius errors; host_rows(k, r0, r1) gives rows r0..r1 of plane k as
    float64 (nt columns); convert(blocks, r0, r1) maps the four planes' row blocks to the reference
    layout when given."""
    num, den = np.zeros(4), np.zeros(4)
    for r0 in range(0, nt, block):
        r1 = min(nt, r0 + block)
        blocks = [host_rows(k, r0, r1) for k in range(4)]
        if convert is not None:
            blocks = convert(blocks, r0, r1)
        for k in range(4):
            ref = ref_planes[k][r0:r1]
            num[k] += np.sum((blocks[k] - ref) ** 2)
            den[k] += np.sum(ref ** 2)
    return np.sqrt(num / den)


def test_fp64_and_fp32_bulk_fills_of_the_c2_board(c2):
    from pcbqs_amd.backend import Mi355xBackend
    from pcbqs_amd.geometry import FillGeometry
    b = c2['board']
    nt = b.nt
    g = FillGeometry.prepare(b.nodes, b.tris, b.z_of, 1)
    backend = Mi355xBackend(g, 1, 16384, 0)
    try:
        print('c2 fp64 fill', flush=True)
        planes, analytic = backend.fill_device(3.0, 1, False, 64)
        assert analytic == c2['analytic']
        ld = planes.shape[-1]

        def rows64(k, r0, r1):
            out = np.empty((r1 - r0, ld))
            from pcbqs_amd._lib import call, ptr
            call('pq_memcpy_d2h', ptr(out), planes.offset((k * nt + r0) * ld), out.nbytes)
            return out[:, :nt]
        err64 = plane_errors(rows64, c2['planes'], nt)
        print(f'c2 FP64 fill rel. Frobenius per plane: {err64}', flush=True)
        assert (err64 <= 1e-10).all(), err64
        planes.close()

        print('c2 fp32 fill', flush=True)
        planes, analytic = backend.fill_device(3.0, 1, False, 32)
        assert analytic == c2['analytic']
        host = planes.host()[:, :, :nt]                      # (4, nt, nt) FP32 local-frame planes
        planes.close()
        cx, cy = g.cent[:, 0], g.cent[:, 1]

        def convert(blocks, r0, r1):
            m0, mx, my, rr = blocks
            ax, ay = cx[r0:r1, None], cy[r0:r1, None]
            mxt = host[1][:, r0:r1].T.astype(np.float64)     # rows r0..r1 of mx^T
            myt = host[2][:, r0:r1].T.astype(np.float64)
            Rr = rr + ax * mxt + ay * myt + cx[None, :] * mx + cy[None, :] * my + \
                (ax * cx[None, :] + ay * cy[None, :]) * m0
            return [m0, mx + ax * m0, my + ay * m0, Rr]
        err32 = plane_errors(lambda k, r0, r1: host[k][r0:r1].astype(np.float64), c2['planes'], nt, convert)
        print(f'c2 FP32-bulk fill rel. Frobenius per plane: {err32}', flush=True)
        assert (err32 <= 2e-6).all(), err32
    finally:
        backend.close()


def rl(Z, freqs):
    w = 2 * np.pi * np.asarray(freqs)[:, None, None]
    return Z.real, Z.imag / w


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def host_residual(s, X, freq):
    """|B - (Rl + j w Ll64) X| / |B| per column on the host, with the device's FP64 loop matrix."""
    n = s.model.n_free
    L = s.loop_matrix(n, precision=64)[:, :n]
    Rl = s.model.Rl.tocsr()[:n, :n]
    B = s.rhs(freq)
    w = 2 * np.pi * freq
    r = B - (Rl @ X + 1j * w * (L @ X))
    return (np.linalg.norm(r, axis=0) / np.linalg.norm(B, axis=0)).max()


@pytest.mark.parametrize('solver', ['iterative', 'direct'])
def test_fp32_bulk_port_rl_at_c2_matches_the_fp64_oracle(c2, solver):
    from pcbqs_amd.solve import PortSolver
    print(f'c2 ports fp32-bulk {solver}', flush=True)
    s = PortSolver(c2['board'], device=0)
    try:
        Z, info, X = s.ports(FREQS, precond='block-jacobi', tol=1e-8, solver=solver, currents=True)
        R, L = rl(Z, FREQS)
        Ro, Lo = rl(c2['Zo'], FREQS)
        eR, eL = rel(R, Ro), rel(L, Lo)
        print(f'c2 {solver}: port R rel. err {eR:.2e}, port L rel. err {eL:.2e}', flush=True)
        assert eR < 1e-6 and eL < 1e-6
        tol = 1e-12 if solver == 'direct' else 5e-8
        for i, f in enumerate(FREQS):
            assert info[i]['relative_residual'] <= tol
            P = s.n_ports
            res = host_residual(s, X[i][:, :P] + 1j * X[i][:, P:], f)
            assert res <= 2 * tol, (f, res)
    finally:
        s.close()


def test_fp64_pipeline_port_rl_at_c2_matches_the_fp64_oracle(c2):
    """Config 1's method at config 2's size: FP64 planes, FP64 branch inductance, FP64 loop matrix;
    the direct solve (FP32 factor refined against the FP64 system to 1e-12)."""
    from pcbqs_amd.solve import PortSolver
    print('c2 ports fp64 pipeline', flush=True)
    s = PortSolver(c2['board'], device=0, precision=64)
    try:
        Z, info = s.ports(FREQS, solver='direct')
        R, L = rl(Z, FREQS)
        Ro, Lo = rl(c2['Zo'], FREQS)
        eR, eL = rel(R, Ro), rel(L, Lo)
        print(f'c2 FP64 pipeline: port R rel. err {eR:.2e}, port L rel. err {eL:.2e}', flush=True)
        assert eR < 1e-9 and eL < 1e-9
        assert all(i['relative_residual'] <= 1e-12 for i in info)
    finally:
        s.close()
