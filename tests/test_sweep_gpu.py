"""The batched sweep over uneven multi-layer boards: every board's port impedances against the FP64
oracle (1e-6), for 4 to 11 ports (9 to 15 are padded to 16 device columns), with two-level GMRES,
with the direct solve, and with the reference's automatic choice between them."""
import numpy as np
import pytest

from pcbqs_amd import board as boards
from tests.test_solve_gpu import oracle_ports, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('solver,batched', [('iterative', True), ('direct', True), ('auto', True),
                                            ('direct', False)])
def test_sweep_matches_the_oracle_on_uneven_boards(solver, batched):
    """solver 'direct' and 'auto' with batched=True solve all boards x frequencies in captured
    multi-board launches (pq_loops_solve_direct_multi); batched=False one board at a time."""
    from pcbqs_amd.sweep import sweep
    seeds = [3, 2, 0]                         # 4, 6 and 11 ports; 1.6k to 5.5k triangles; 4 layers
    bs = [boards.board_batch(s) for s in seeds]
    freqs = [1e3, 1e6, 1e8]
    res, total = sweep(bs, freqs, precond='two-level', grid=8, tol=1e-9, solver=solver, batched=batched)
    assert total['boards'] == 3 and total['solves'] == 9
    if solver != 'iterative' and batched:
        assert total['graph_launches'] >= 1 and total['batched']
        print(f"graph launches {total['graph_launches']}, solved again alone {total['redone_alone']}", flush=True)
    w = 2 * np.pi * np.array(freqs)[:, None, None]
    for b, r in zip(bs, res):
        print(f"board {r['board']}: {b.nt} triangles, {r['ports']} ports, iterations {r['iterations']}", flush=True)
        Zo = oracle_ports(b, freqs)
        assert r['Z'].shape == Zo.shape
        assert rel(r['Z'].real, Zo.real) < 1e-6
        assert rel(r['Z'].imag / w, Zo.imag / w) < 1e-6
        assert r['worst_residual'] <= 1e-9


def test_multi_board_launch_equals_the_per_board_direct_solves():
    """The captured multi-board solve (identity-padded matrices, in-graph refinement) gives every
    board's port impedances of its own direct solve to 1e-10, and residuals at the reference's 1e-12."""
    from pcbqs_amd.sweep import sweep
    bs = [boards.board_batch(s) for s in (5, 1, 4, 3)]
    freqs = list(np.logspace(3, 8, 5))
    one, t1 = sweep(bs, freqs, solver='direct', batched=False)
    many, t2 = sweep(bs, freqs, solver='direct', batched=True)
    print(f"per-board {t1['seconds']:.3f} s, batched {t2['seconds']:.3f} s in {t2['graph_launches']} launches", flush=True)
    for a, b in zip(one, many):
        assert np.abs(a['Z'] - b['Z']).max() <= 1e-10 * np.abs(a['Z']).max()
        assert b['worst_residual'] <= 1e-12
