"""The charge-solve oracle (oracle/peec.py) pinned on the reference's own physical checks
(tests/test_peec.py of the reference): plates with no copper between them pass displacement current
only, with the capacitance of the Maxwell matrix; KCL and potential = P q hold at every node; the
total charge is zero; the response is linear in the source amplitude; the impedance scales as 1 / er;
a closed copper path approaches the inductive (no-charge) solve."""
import numpy as np
import pytest

from oracle import mqs as M
from oracle import peec as Q
from pcbqs_amd import board as boards


@pytest.fixture(scope='module')
def plates():
    b = boards.board_plates()
    return b, Q.prepare(b)


def test_disconnected_plates_carry_displacement_current(plates):
    b, m = plates
    freqs = [1e2, 1e5, 1e7]
    i, q, v, z = Q.solve(m, freqs)
    C = Q.capacitance(m)
    labels = m['B'].argmax(1)
    drive = np.zeros(len(C))
    drive[labels[b.terminals['B']]] = 1
    drive[labels[b.terminals['F']]] = -1
    expected = 1 / (drive @ np.linalg.solve(C, drive))
    measured = -1 / (2 * np.pi * np.array(freqs[:2]) * z['I'][:2].imag)
    assert measured == pytest.approx(expected, rel=2e-5)
    assert np.all(z['I'].real >= -1e-6)
    parallel = Q.EPS0 * 0.01 ** 2 / 0.0005
    assert parallel < expected < 1.5 * parallel
    assert np.abs(q.sum(1)).max() < 1e-17
    ends = np.r_[np.asarray(M.CopperModel(b.nodes, b.tris, b.z_of, b.t_of, b.node_of, b.terminals, b.vseg,
                                          b.via_nodes, b.ports).br).reshape(-1, 2),
                 [(b.terminals['F'], b.terminals['B'])]]
    for fi, f in enumerate(freqs):
        balance = 2j * np.pi * f * q[fi]
        np.add.at(balance, ends[:, 0], i[fi])
        np.subtract.at(balance, ends[:, 1], i[fi])
        assert np.abs(balance).max() < 1e-9
        assert v[fi] == pytest.approx(m['P'] @ q[fi])


def test_linearity_and_dielectric_scaling(plates):
    b, m = plates
    i1, q1, _, z1 = Q.solve(m, [1e3, 1e6])
    i2, q2, _, z2 = Q.solve(m, [1e3, 1e6], amps={'I': 2j})
    assert i2 == pytest.approx(2j * i1)
    assert q2 == pytest.approx(2j * q1)
    assert z2['I'] == pytest.approx(z1['I'])
    m4 = Q.prepare(boards.board_plates(whole=True), er=4.0)
    ma = Q.prepare(boards.board_plates(whole=True))
    assert Q.solve(m4, [1e3])[3]['I'][0] == pytest.approx(Q.solve(ma, [1e3])[3]['I'][0] / 4, rel=1e-10)


def test_closed_copper_path_approaches_the_inductive_solve():
    b = boards.board_strip()
    freqs = [1e-3, 1e2, 1e5]
    m = Q.prepare(b, ideal_terminals=False)
    z_peec = Q.solve(m, freqs)[3]['I']
    q = Q.copper_model(b, ideal_terminals=False)
    q.assemble()
    z_mqs = q.port_response(freqs)[0][:, 0, 0]
    assert z_peec == pytest.approx(z_mqs, rel=1e-5)
