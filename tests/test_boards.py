"""The seeded board generators (CPU): reproducible from the seed, the sizes the configurations
name, and connected the way the loop solve needs (every via and port on meshed copper)."""
import numpy as np

from pcbqs_amd import board as boards
from pcbqs_amd.geometry import RWG


def test_config5_board_is_seeded_and_sized():
    a, b = boards.board_c5(0), boards.board_c5(0)
    assert np.array_equal(a.tris, b.tris) and np.array_equal(a.nodes, b.nodes)
    assert 60000 <= a.nt <= 74000                              # about 67k triangles
    assert sorted(np.unique(a.layer_of)) == [0, 1, 2, 3]
    assert len(a.ports) == 8
    xy = {(round(x, 9), round(y, 9)) for x, y, _, _, _ in a.vseg}
    assert len(xy) == 200                                      # 200 vias, each joining >= 2 layers
    assert np.ptp(a.nodes[:, 0]) <= 0.120 + 1e-12 and np.ptp(a.nodes[:, 1]) <= 0.120 + 1e-12
    assert 90000 <= RWG(a.nodes, a.tris).Ne <= 110000          # about 100k RWG
    c = boards.board_c5(1)
    assert c.nt != a.nt or not np.array_equal(c.tris, a.tris)


def test_config5_via_segments_join_layers_with_copper_at_the_via():
    b = boards.board_c5(0)
    cen = b.nodes[b.tris].mean(1)
    for x, y, _, za, zb in b.vseg:
        assert za < zb
        for z in (za, zb):
            on = b.z_of == z
            assert np.hypot(cen[on, 0] - x, cen[on, 1] - y).min() < 2e-3
