"""The batched sweep over uneven multi-layer boards: every board's port impedances against the FP64
oracle (1e-6), for 4 to 11 ports (9 to 15 are padded to 16 device columns), with two-level GMRES,
with the direct solve, and with the reference's automatic choice between them."""
import numpy as np
import pytest

from pcbqs_amd import board as boards
from tests.test_solve_gpu import oracle_ports, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('solver', ['iterative', 'direct', 'auto'])
def test_sweep_matches_the_oracle_on_uneven_boards(solver):
    from pcbqs_amd.sweep import sweep
    seeds = [3, 2, 0]                         # 4, 6 and 11 ports; 1.6k to 5.5k triangles; 4 layers
    bs = [boards.board_batch(s) for s in seeds]
    freqs = [1e3, 1e6, 1e8]
    res, total = sweep(bs, freqs, precond='two-level', grid=8, tol=1e-9, solver=solver)
    assert total['boards'] == 3 and total['solves'] == 9
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    for b, r in zip(bs, res):
        print(f"board {r['board']}: {b.nt} triangles, {r['ports']} ports, iterations {r['iterations']}", flush=True)
        Zo = oracle_ports(b, freqs)
        assert r['Z'].shape == Zo.shape
        assert rel(r['Z'].real, Zo.real) < 1e-6
        assert rel(r['Z'].imag / w, Zo.imag / w) < 1e-6
        assert r['worst_residual'] <= 1e-9
