"""The C ABI library: built, loadable, exporting every entry point include/pcbqs_amd.h declares;
and no silent CPU path when no device is visible."""
import ctypes
import os
import re

import numpy as np
import pytest

from pcbqs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, 'include', 'pcbqs_amd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(pq_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIBRARY)
    declared = declared_entry_points()
    assert declared, 'no declarations found'
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.exported_symbols())


def test_error_reporting_without_a_device():
    if _lib.device_count() > 0:
        pytest.skip('a device is visible')
    from pcbqs_amd.backend import Mi355xBackend
    from pcbqs_amd.geometry import FillGeometry
    g = FillGeometry.prepare(np.array([[0., 0.], [1., 0.], [0., 1.]]), np.array([[0, 1, 2]]), np.zeros(1), 1)
    with pytest.raises(RuntimeError, match='MI355X fill unavailable'):
        Mi355xBackend(g, 1, 16384, None)
    with pytest.raises((RuntimeError, MemoryError)):
        _lib.DeviceArray((4,), np.float64)


def test_rwg_topology_matches_the_oracle():
    from oracle import mqs as M
    from pcbqs_amd.geometry import RWG
    from tests.meshes import two_layer
    nodes, tris, _ = two_layer(9, 7, seed=4)
    a, b = RWG(nodes, tris), M.RWG(nodes, tris)
    assert a.Ne == b.Ne
    for name in ('tp', 'tm', 'vp', 'vm', 'l'):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


def test_geometry_prepare_matches_the_oracle_bitwise():
    from oracle import fill as F
    from pcbqs_amd.geometry import FillGeometry
    from tests.meshes import two_layer
    nodes, tris, z = two_layer(9, 7, seed=5)
    for level in (0, 1, 2):
        a, b = FillGeometry.prepare(nodes, tris, z, level), F.prepare(nodes, tris, z, level)
        for name in ('nodes', 'cent', 'radius', 'points', 'z'):
            assert np.array_equal(getattr(a, name), getattr(b, name)), name
        for k in a.levels:
            assert all(np.array_equal(x, y) for x, y in zip(a.levels[k], b.levels[k]))
    with pytest.raises(ValueError, match='counter-clockwise'):
        FillGeometry.prepare(nodes, tris[:, ::-1], z, 1)
