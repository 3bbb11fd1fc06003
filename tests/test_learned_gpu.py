"""The learned coarse-space correction on the device: training through the MFMA kernels lowers the
projection loss, and the committed network cuts two-level GMRES iterations on a held-out board while
the port impedances stay those of plain block-Jacobi GMRES (the judge)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def test_training_lowers_the_projection_loss(torch):
    from pcbqs_amd import board as boards
    from pcbqs_amd.learned import BoardData, train
    data = [BoardData(boards.board_c2(s, target=4000), grid=8, freqs=(1e3, 1e6)) for s in (101, 102)]
    net, hist = train(data, steps=80, log=None)
    for bd in data:
        bd.close()
    assert np.isfinite(hist).all()
    assert np.mean(hist[-10:]) < 0.8 * hist[0], (hist[0], hist[-10:])


def test_trained_network_cuts_iterations_on_a_held_out_board(torch):
    from pcbqs_amd import board as boards
    from pcbqs_amd.learned import MPNN, compare
    from pcbqs_amd.solve import PortSolver
    solver = PortSolver(boards.board_c2(0), device=0)          # seed 0: not among the training seeds 1..16
    net = MPNN.load()
    f = 1e6
    Zp, plain = solver.ports([f], precond='block-jacobi')
    rows = compare(solver, net, [f], grid=24)                  # leaves the learned coarse space set
    Zl, learned = solver.ports([f], precond='two-level')
    solver.close()
    hats_its, learned_its = rows[0]['hats'][0], rows[0]['learned'][0]
    print('plain', plain[0], 'hats/learned', rows[0])
    assert learned_its <= 0.9 * hats_its
    assert learned[0]['relative_residual'] <= 1e-8
    assert np.abs(Zl - Zp).max() / np.abs(Zp).max() < 1e-6
    assert learned[0]['seconds'] < 0.5 * plain[0]['seconds']


def test_training_on_boards_spanning_10k_to_40k_rwg(torch):
    """Config 4's board range: three seeded plane pairs of about 10k, 22k and 40k RWG edges feed one
    training run (the pairs generated by the solver itself); the projection loss falls."""
    from pcbqs_amd import board as boards
    from pcbqs_amd.learned import BoardData, train
    data = []
    for seed, target in ((201, 6800), (202, 15000), (203, 26500)):
        print(f'c4 board seed={seed} target={target}', flush=True)
        data.append(BoardData(boards.board_c2(seed, target=target), grid=8, freqs=(1e3, 1e6), keep_solver=False))
    rwg = sorted(bd.rwg for bd in data)
    print(f'c4 RWG per board: {rwg}', flush=True)
    assert 10_000 <= rwg[0] and rwg[-1] <= 40_000 and rwg[-1] >= 30_000
    net, hist = train(data, steps=120, log=None)
    assert np.isfinite(hist).all()
    assert np.mean(hist[-20:]) < 0.8 * np.mean(hist[:5]), (hist[:5], hist[-20:])
