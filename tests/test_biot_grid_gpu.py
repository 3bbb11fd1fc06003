"""GPU parity of the FFT near-field grid (pq_biot_grid: deposit on the host, linear FFT convolution
on the device) against the oracle's restatement (scipy.fft) to 1e-9, and against direct
Biot-Savart on the device to 1 % on the copper fixture's solution (tests/test_mqs.py:122-133),
closer with a refined grid."""
import numpy as np
import pytest

from oracle import biot as B
from tests.test_biot_grid_oracle import copper_solution, random_sheets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('refine', [1, 2])
def test_device_grid_matches_the_oracle(refine):
    from pcbqs_amd import biot as D
    sheets = random_sheets(2, k=3)
    x, y, heights = np.linspace(0.003, 0.015, 13), np.linspace(0.001, 0.011, 9), [0.002, -0.0015, 0.0004]
    timing = {}
    got = D.h_grid(sheets, x, y, heights, refine=refine, timing=timing)
    want = B.h_grid(sheets, x, y, heights, refine=refine)
    print(f'grid refine {refine}: {timing}', flush=True)
    assert got.shape == want.shape == (13, 9, 3, 3, 3)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize('refine', [1, 2])
def test_device_grid_matches_biot_savart_on_the_copper_fixture(refine):
    from pcbqs_amd import biot as D
    sheets, segs = copper_solution()
    x, y, heights = np.linspace(.012, .018, 25), np.linspace(.007, .013, 25), [.0025, -.002]
    H = D.h_grid(sheets, x, y, heights, segs, refine=refine)
    X, Y, Z = np.meshgrid(x[::6], y[::6], heights, indexing='ij')
    pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    ref = D.h_field(sheets, pts, segs).reshape(5, 5, 2, 3, 2)
    err = np.abs(H[::6, ::6] - ref).max() / np.abs(ref).max()
    print(f'copper grid refine {refine}: {err:.2e}', flush=True)
    assert err < 0.01
    np.testing.assert_allclose(H, B.h_grid(sheets, x, y, heights, segs, refine=refine), rtol=1e-9,
                               atol=1e-9 * np.abs(ref).max())


def test_single_excitation_and_errors():
    from pcbqs_amd import biot as D
    sheets = [(n, t, a[:, 0], b[:, :, 0], z) for n, t, a, b, z in random_sheets(3, k=1)]
    x, y = np.linspace(0.004, 0.012, 5), np.linspace(0.002, 0.008, 4)
    H = D.h_grid(sheets, x, y, [0.001])
    assert H.shape == (5, 4, 1, 3)
    np.testing.assert_allclose(H, B.h_grid(sheets, x, y, [0.001]), rtol=1e-9, atol=1e-9 * np.abs(H).max())
    with pytest.raises(ValueError, match='copper'):
        D.h_grid(sheets, x, y, [sheets[0][4]])
    with pytest.raises(ValueError, match='uniform'):
        D.h_grid(sheets, np.array([0.0, 0.001, 0.003]), y, [0.001])
