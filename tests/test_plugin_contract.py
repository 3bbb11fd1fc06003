"""The drop-in boundary: `mi355x` found through the pcbqs.static_backends entry-point group and
used through the plugin protocol only (factory, parallel_tiles, min_batch_pairs, evaluate,
fill_symmetric, the error types). The GPU cases compare with the oracle (oracle/cpu_ref.c)."""
import inspect
import sys

import numpy as np
import pytest

from tests import plugin_contract as pc


@pytest.fixture
def resolver(tmp_path):
    path = pc.entry_point_path(tmp_path)
    yield pc.resolve
    sys.path.remove(path)


def test_mi355x_resolves_to_the_package_factory(resolver):
    from pcbqs_amd.backend import make_backend
    factory = resolver('mi355x')
    assert factory is make_backend
    assert list(inspect.signature(factory).parameters) == ['geometry', 'workers', 'chunk', 'device']


def test_unknown_backend_names_do_not_resolve(resolver):
    with pytest.raises(ValueError, match='0 providers'):
        resolver('no-such-backend')


def test_protocol_attributes_are_class_level():
    """parallel_tiles and min_batch_pairs are read before any call, so they must not need a device."""
    from pcbqs_amd.backend import Mi355xBackend
    assert Mi355xBackend.parallel_tiles is False
    assert isinstance(Mi355xBackend.min_batch_pairs, int) and Mi355xBackend.min_batch_pairs >= 1


@pytest.mark.gpu
def test_resolved_backend_fills_and_evaluates_like_the_oracle(resolver):
    from oracle import fill as F
    from pcbqs_amd.geometry import FillGeometry
    from tests.meshes import two_layer
    nodes, tris, z = two_layer(20, 12, seed=0)
    n = len(tris)
    for level in (0, 1, 2):
        print(f'plugin contract level={level} nt={n}', flush=True)
        g = FillGeometry.prepare(nodes, tris, z, level)
        backend = resolver('mi355x')(g, 1, 16384, 0)
        ref_g = F.prepare(nodes, tris, z, level)
        rng = np.random.default_rng(level)
        p, q = rng.integers(0, n, 3000).astype(np.int32), rng.integers(0, n, 3000).astype(np.int32)
        p[:50] = q[:50]
        d = ref_g.z[p] - ref_g.z[q]
        kind = F.kinds(ref_g, p, q, d, 3.0, level)
        for scalar in (False, True):
            out = backend.evaluate(p, q, d, kind, scalar)
            assert isinstance(out, np.ndarray) and out.dtype == np.float64
            assert out.shape == (len(p), 1 if scalar else 6) and np.isfinite(out).all()
            ref = F.cr_pairs(ref_g, p, q, d, kind, scalar)
            assert out == pytest.approx(ref, rel=1e-11, abs=1e-26)
        # a kind naming a level the geometry did not prepare is the caller's error
        bad = np.full(len(p), {0: 1, 1: 2, 2: 1}[level], np.int32)
        with pytest.raises(ValueError, match='prepared analytic level'):
            backend.evaluate(p, q, d, bad, False)
        if level == 1:
            planes = [np.empty((n, n)) for _ in range(4)]
            analytic = backend.fill_symmetric(planes, 3.0, 1, False)
            ref, pairs, ref_analytic = F.cr_fill(ref_g, 3.0, 1, False)
            assert isinstance(analytic, int) and analytic == ref_analytic
            for a, b in zip(planes, ref):
                assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-10
            with pytest.raises(ValueError, match='writable C-contiguous'):
                backend.fill_symmetric([np.empty((n, n), np.float32)] * 4, 3.0, 1, False)
        backend.close()


@pytest.mark.gpu
def test_out_of_range_pairs_raise_value_error(resolver):
    from pcbqs_amd.geometry import FillGeometry
    from tests.meshes import two_layer
    nodes, tris, z = two_layer(6, 4, seed=0)
    backend = resolver('mi355x')(FillGeometry.prepare(nodes, tris, z, 1), 1, 16384, 0)
    p = np.array([0, len(tris)], np.int32)
    with pytest.raises(ValueError, match='invalid triangle index'):
        backend.evaluate(p, p, np.zeros(2), np.full(2, -1, np.int32), False)
    backend.close()
