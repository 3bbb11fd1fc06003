"""Config 5 on the device: the seeded 4-layer 120 x 120 mm PDN board_c5(0) (66,934 triangles, 99,342
RWG edges, 32,211 free loops, 8 ports) through the whole pipeline: FP32-bulk fill and combine (the
planes and the branch inductance dense in HBM), the projection, and three solves at 1 MHz that share
nothing but the operator: plain block-Jacobi GMRES, learned two-level GMRES, and the direct solve.
No FP64 oracle runs at this size (a dense FP64 loop solve of 32k unknowns is far beyond the host's
budget in a test), so the checks are the ones the size allows:
- every solve's residual is the FP64 system's true residual (tol 1e-8 for GMRES, 1e-12 direct), and
  the three port impedance matrices agree to 1e-8;
- Biot-Savart of the solved currents in FP32 agrees with the FP64 kernel on 2,000 points of the
  config-5 field grid (0.5-5 mm above the top layer) to 2e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def c5():
    from pcbqs_amd import board as boards
    from pcbqs_amd.solve import PortSolver
    b = boards.board_c5(0)
    print(f'c5: {b.nt} triangles', flush=True)
    s = PortSolver(b, device=0)
    print(f'c5: {s.model.n_free} free loops, set-up {s.setup_seconds}', flush=True)
    yield b, s
    s.close()


def test_c5_plain_learned_and_direct_solves_agree(c5):
    from pcbqs_amd.learned import MPNN, set_learned_coarse
    b, s = c5
    f = 1e6
    Zp, ip = s.ports([f], precond='block-jacobi', tol=1e-8)
    print(f"c5 plain: {ip[0]['iterations']} iterations, {ip[0]['seconds']:.3f} s", flush=True)
    set_learned_coarse(s, MPNN.load(), 24)
    Zl, il = s.ports([f], precond='two-level', tol=1e-8)
    print(f"c5 learned: {il[0]['iterations']} iterations, {il[0]['seconds']:.3f} s, {s.coarse_k} coarse columns",
          flush=True)
    Zd, idd = s.ports([f], solver='direct')
    print(f"c5 direct: {idd[0]['seconds']:.3f} s, residual {idd[0]['relative_residual']:.2e}", flush=True)
    assert ip[0]['relative_residual'] <= 1e-8 and il[0]['relative_residual'] <= 1e-8
    assert idd[0]['relative_residual'] <= 1e-12
    rel = lambda a, c: np.abs(a - c).max() / np.abs(c).max()
    assert rel(Zl, Zd) <= 1e-8 and rel(Zp, Zd) <= 1e-8
    assert il[0]['iterations'] < ip[0]['iterations']


def test_c5_field_fp32_matches_fp64_on_a_point_subset(c5):
    from pcbqs_amd.biot import h_field
    from pcbqs_amd.scale import field_sources
    b, s = c5
    f = 1e6
    Z, info, X = s.ports([f], precond='block-jacobi', tol=1e-8, currents=True)
    P = s.n_ports
    sheets, segments = field_sources(s.model, X[0][:, :P] + 1j * X[0][:, P:], s.port_cols)
    rng = np.random.default_rng(5)
    lo, hi = b.nodes.min(0), b.nodes.max(0)
    top = float(b.z_of.max())
    pts = np.c_[rng.uniform(lo[0], hi[0], 2000), rng.uniform(lo[1], hi[1], 2000),
                top + rng.uniform(0.5e-3, 5e-3, 2000)]
    H32 = h_field(sheets, pts, segments, precision=32)
    H64 = h_field(sheets, pts, segments, precision=64)
    err = np.abs(H32 - H64).max() / np.abs(H64).max()
    print(f'c5 field: FP32 vs FP64 on 2,000 points: {err:.2e}', flush=True)
    assert np.isfinite(H64).all() and err <= 2e-5
