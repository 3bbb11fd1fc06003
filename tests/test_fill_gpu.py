"""GPU parity of the fill: the HIP path against the oracle (oracle/cpu_ref.c, which equals the
reference core bitwise) on the same seeded meshes. FP64: <= 1e-10 relative Frobenius per plane,
identical pair and analytic counts, bitwise reciprocity and repeatability. FP32 bulk (far FP32,
near FP64, local frame): relative Frobenius <= 2e-6 per plane after conversion to the reference
layout, and L within 1e-5."""
import numpy as np
import pytest

from oracle import fill as F
from oracle import mqs as M
from tests.meshes import two_layer

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def to_reference_layout(planes, cent):
    """FP32 local-frame planes -> the reference's centred-frame planes (float64)."""
    m0, mx, my, rr = (np.asarray(p, np.float64) for p in planes)
    cx, cy = cent[:, 0], cent[:, 1]
    Mx = mx + cx[:, None] * m0
    My = my + cy[:, None] * m0
    Rr = rr + cx[:, None] * mx.T + cy[:, None] * my.T + cx[None, :] * mx + cy[None, :] * my + \
        (cx[:, None] * cx[None, :] + cy[:, None] * cy[None, :]) * m0
    return m0, Mx, My, Rr


@pytest.fixture(scope='module')
def meshes():
    return {'islands': None, 'two_layer': two_layer(20, 12, seed=0), 'strips': two_layer(48, 30, seed=1)}


@pytest.mark.parametrize('level', [0, 1, 2])
@pytest.mark.parametrize('scalar_only', [False, True])
@pytest.mark.parametrize('near', [0.0, 0.5, 3.0, np.inf])
def test_fill_symmetric_matches_oracle_on_the_island_mesh(mesh, level, scalar_only, near):
    from pcbqs_amd.static import static_moments
    print(f'fill_symmetric islands level={level} scalar={scalar_only} near={near}', flush=True)
    (local, *gpu), info = static_moments(*mesh, near=near, outer_levels=level, scalar_only=scalar_only,
                                         device=0, return_info=True)
    g = F.prepare(*mesh, level)
    ref, pairs, analytic = F.cr_fill(g, near, level, scalar_only)
    assert (info['pairs'], info['analytic_pairs']) == (pairs, analytic)
    for a, b in zip(gpu, ref):
        assert a == pytest.approx(b, rel=2e-10, abs=1e-26)
    assert np.array_equal(gpu[0], gpu[0].T)
    if not scalar_only:
        assert np.array_equal(gpu[3], gpu[3].T)


@pytest.mark.parametrize('name', ['two_layer', 'strips'])
def test_fill_symmetric_matches_oracle_on_two_layer_meshes(meshes, name):
    from pcbqs_amd.static import static_moments
    nodes, tris, z = meshes[name]
    print(f'fill_symmetric {name} nt={len(tris)}', flush=True)
    (local, *gpu), info = static_moments(nodes, tris, z, device=0, return_info=True)
    g = F.prepare(nodes, tris, z, 1)
    ref, pairs, analytic = F.cr_fill(g, 3.0, 1, False)
    assert (info['pairs'], info['analytic_pairs']) == (pairs, analytic)
    for a, b in zip(gpu, ref):
        assert rel(a, b) < 1e-10
    assert np.array_equal(gpu[0], gpu[0].T) and np.array_equal(gpu[3], gpu[3].T)
    again = static_moments(nodes, tris, z, device=0)
    assert all(np.array_equal(a, b) for a, b in zip(gpu, again[1:]))


@pytest.mark.parametrize('precision', [64, 32])
def test_device_fill_matches_oracle(meshes, precision):
    from pcbqs_amd.backend import Mi355xBackend
    nodes, tris, z = meshes['strips']
    print(f'fill_device precision={precision} nt={len(tris)}', flush=True)
    g = F.prepare(nodes, tris, z, 1)
    backend = Mi355xBackend(g, 1, 16384, 0)
    planes, analytic = backend.fill_device(3.0, 1, False, precision)
    host = planes.host()[:, :, :len(tris)]
    ref, pairs, ranalytic = F.cr_fill(g, 3.0, 1, False)
    assert analytic == ranalytic
    if precision == 64:
        for a, b in zip(host, ref):
            assert rel(a, b) < 1e-10
        assert np.array_equal(host[0], host[0].T)
    else:
        conv = to_reference_layout(host, g.cent)
        for k, (a, b) in enumerate(zip(conv, ref)):
            assert rel(a, b) < 2e-6, k
        assert np.array_equal(host[0], host[0].T) and np.array_equal(host[3], host[3].T)
    scalar, analytic_s = backend.fill_device(3.0, 1, True, precision)
    assert analytic_s == ranalytic
    assert rel(scalar.host()[0, :, :len(tris)].astype(np.float64), ref[0]) < (1e-10 if precision == 64 else 2e-6)
    far_ms, near_ms, near_pairs = backend.last_timing()
    assert far_ms > 0 and sum(near_pairs) == ranalytic
    backend.close()


@pytest.mark.parametrize('precision', [64, 32])
def test_device_combine_matches_oracle(meshes, precision):
    from pcbqs_amd.backend import Mi355xBackend, combine_device
    from pcbqs_amd.geometry import RWG
    nodes, tris, z = meshes['strips']
    print(f'combine precision={precision}', flush=True)
    g = F.prepare(nodes, tris, z, 1)
    backend = Mi355xBackend(g, 1, 16384, 0)
    planes, _ = backend.fill_device(3.0, 1, False, precision)
    r = RWG(g.nodes, tris)
    scale = 1 / r.l
    L = combine_device(planes, precision, r, g.cent, 1e-7, scale, 64).host()[:, :r.Ne]
    ref_planes, _, _ = F.cr_fill(g, 3.0, 1, False)
    m0, mx, my, rr = ref_planes
    Lref = F.cr_combine(M.RWG(g.nodes, tris), rr, mx, my, m0, 1e-7, scale)
    assert np.array_equal(L, L.T)
    assert rel(L, Lref) < (1e-10 if precision == 64 else 1e-5)
    backend.close()


def test_edge_inductance_from_host_moments(mesh):
    from pcbqs_amd.static import edge_inductance, static_moments
    from pcbqs_amd.geometry import RWG
    nodes, m0, mx, my, rr = static_moments(*mesh, device=0)
    r = RWG(nodes, mesh[1])
    L, info = edge_inductance(r, rr, mx, my, m0, 1e-7, 1 / r.l)
    ref = F.cr_combine(M.RWG(nodes, mesh[1]), rr, mx, my, m0, 1e-7, 1 / r.l)
    assert L == pytest.approx(ref, rel=1e-12, abs=1e-12 * np.abs(ref).max())
    assert np.linalg.eigvalsh(L).min() > 0


def test_store_probe_writes_its_pattern():
    """The store-ceiling probe (bench.py's reference beside the fill's roofline) writes r - c into
    every (r, c < n) entry of each plane and reports a positive time."""
    import ctypes
    from pcbqs_amd._lib import DeviceArray, call
    n, ld, planes = 200, 256, 3
    out = DeviceArray((planes, n, ld), np.float32).zero()
    ms = ctypes.c_double()
    call('pq_store_probe', out.ptr, n, ld, planes, ctypes.byref(ms))
    host = out.host()
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    for p in range(planes):
        np.testing.assert_array_equal(host[p, :, :n], (r - c).astype(np.float32))
        assert not host[p, :, n:].any()
    assert ms.value > 0
    out.close()
