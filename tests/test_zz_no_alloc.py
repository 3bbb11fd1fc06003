"""pq_pairs_noalloc, the explicit pair entry that allocates no device memory (its staging buffers
live in a fixed device array of the library), against pq_pairs on the same inputs: bitwise equal
moments on two calls in a row, the second with fewer pairs (the array keeps the first call's
contents, which must not leak into the second); full and scalar moments; a call beyond the stated
maximum is refused on the host before any device work."""
import ctypes

import numpy as np
import pytest

from pcbqs_amd._lib import call, ptr

pytestmark = pytest.mark.gpu


def pairs(handle, name, p, q, d, kind, scalar):
    out = np.full((len(p), 1 if scalar else 6), np.nan)
    call(name, handle, ptr(p), ptr(q), ptr(d), ptr(kind), len(p), int(scalar), ptr(out))
    return out


def test_noalloc_pairs_equal_the_allocating_entry_point():
    from oracle import fill as F
    from pcbqs_amd.backend import Mi355xBackend
    from pcbqs_amd.geometry import FillGeometry
    from tests.meshes import two_layer
    nodes, tris, z = two_layer(48, 30, seed=2)
    n = len(tris)
    level = 1
    g = FillGeometry.prepare(nodes, tris, z, level)
    ref_g = F.prepare(nodes, tris, z, level)
    backend = Mi355xBackend(g, 1, 16384, 0)
    try:
        rng = np.random.default_rng(7)
        for count in (300_000, 7_000):
            print(f'noalloc pairs count={count} nt={n}', flush=True)
            p = rng.integers(0, n, count).astype(np.int32)
            q = rng.integers(0, n, count).astype(np.int32)
            q[: count // 20] = np.clip(p[: count // 20] + rng.integers(-3, 4, count // 20), 0, n - 1)
            d = np.ascontiguousarray(g.z[p] - g.z[q])
            kind = np.ascontiguousarray(F.kinds(ref_g, p, q, d, 3.0, level), np.int32)
            assert (kind >= 0).any() and (kind < 0).any()
            for scalar in (False, True):
                a = pairs(backend._handle, 'pq_pairs', p, q, d, kind, scalar)
                b = pairs(backend._handle, 'pq_pairs_noalloc', p, q, d, kind, scalar)
                assert np.isfinite(b).all()
                assert np.array_equal(a, b), (count, scalar)
        big = np.zeros((1 << 20) + 1, np.int32)
        out = np.empty((len(big), 6))
        with pytest.raises(ValueError, match='at most 1048576 pairs'):
            call('pq_pairs_noalloc', backend._handle, ptr(big), ptr(big), ptr(np.zeros(len(big))), ptr(big),
                 len(big), 0, ptr(out))
    finally:
        backend.close()
