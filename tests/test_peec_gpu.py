"""The charge (PEEC) solve on the device (pcbqs_amd.peec.ChargeSolver: FP64 fill, P and K^T Lb K in
HBM, the scaled dense system factored by an FP64 LU per frequency) against the oracle's dense
restatement (oracle/peec.py) on the same boards: currents, charges, potentials and the source
impedance; and the reference's physical checks on the device's own results."""
import numpy as np
import pytest

from oracle import peec as Q
from pcbqs_amd import board as boards

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('name', ['plates', 'plates-whole', 'strip', 'c1'])
def test_charge_solve_matches_the_oracle(name):
    from pcbqs_amd.peec import ChargeSolver
    b, ideal = {'plates': (boards.board_plates(), True), 'plates-whole': (boards.board_plates(whole=True), True),
                'strip': (boards.board_strip(), False), 'c1': (boards.board_c1(target=800), True)}[name]
    freqs = [1e2, 1e5, 1e7]
    print(f'peec {name}: {b.nt} triangles', flush=True)
    s = ChargeSolver(b, device=0, ideal_terminals=ideal)
    r = s.solve(freqs)
    i, q, v, z = Q.solve(Q.prepare(b, ideal_terminals=ideal), freqs)
    for fi in range(len(freqs)):
        assert rel(r.i[fi], i[fi]) < 1e-8, (name, fi)
        assert rel(r.charge[fi], q[fi]) < 1e-8, (name, fi)
        assert rel(r.potential[fi], v[fi]) < 1e-8, (name, fi)
    for name_ in z:
        assert rel(r.z[name_], z[name_]) < 1e-9, name_
    s.close()


def test_device_plates_physics():
    """The reference's plates checks on the device's results: displacement-current capacitance equal
    to 1 / (d^T C^-1 d) of the device's Maxwell matrix, inside the parallel-plate bracket, zero total
    charge, KCL at every node, potential = P q, linearity in the amplitude."""
    from pcbqs_amd.geometry import EPS0
    from pcbqs_amd.peec import ChargeSolver
    b = boards.board_plates()
    s = ChargeSolver(b, device=0)
    freqs = [1e2, 1e5, 1e7]
    r = s.solve(freqs)
    C, labels = s.capacitance_matrix()
    drive = np.zeros(len(C))
    drive[labels[b.terminals['B']]] = 1
    drive[labels[b.terminals['F']]] = -1
    expected = 1 / (drive @ np.linalg.solve(C, drive))
    measured = -1 / (2 * np.pi * np.array(freqs[:2]) * r.z['I'][:2].imag)
    assert measured == pytest.approx(expected, rel=2e-5)
    assert EPS0 * 1e-4 / 5e-4 < expected < 1.5 * EPS0 * 1e-4 / 5e-4
    assert np.abs(r.charge.sum(1)).max() < 1e-17
    A = s.A
    P = s.potential_matrix()
    for fi, f in enumerate(freqs):
        assert np.abs(2j * np.pi * f * r.charge[fi] + A @ r.i[fi]).max() < 1e-9
        assert r.potential[fi] == pytest.approx(P @ r.charge[fi])
    r2 = s.solve(freqs, amps={'I': 2j})
    assert r2.i == pytest.approx(2j * r.i)
    assert r2.z['I'] == pytest.approx(r.z['I'])
    s.close()
