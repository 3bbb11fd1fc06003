"""Host-side geometry: the prepared fill geometry and the RWG basis (topology only).

FillGeometry mirrors src/pcbqs/static_backends.py:94-134 (same fields, same validation
and error messages) so that a geometry prepared here and one prepared by the reference's
FillEngine are interchangeable for the `mi355x` backend. The arrays are computed with the
same elementary numpy operations as the reference (centring, centroid means, bounding
radii, rule points): the near/far rule compares sums of these values against a threshold,
so they must agree to the last bit for the two sides to choose the same quadrature.

RWG mirrors src/pcbqs/_numerics/base.py:10-41: interior edges numbered in order of first
appearance over the triangles' edges (t-major, edges (a,b), (b,c), (c,a)); T+ is the
triangle where the edge appears first; the free vertices, lengths and areas.
"""
from dataclasses import dataclass

import numpy as np

MU0 = 4e-7 * np.pi
C0 = 299792458.0
EPS0 = 1.0 / (MU0 * C0 ** 2)

_a1, _a2 = 0.445948490915965, 0.091576213509771
BAR6 = np.array([[1 - 2 * _a1, _a1, _a1], [_a1, 1 - 2 * _a1, _a1], [_a1, _a1, 1 - 2 * _a1],
                 [1 - 2 * _a2, _a2, _a2], [_a2, 1 - 2 * _a2, _a2], [_a2, _a2, 1 - 2 * _a2]])
W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
BAR3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
W3 = np.ones(3) / 3


def _split4(V):
    m01, m12, m20 = (V[:, 0] + V[:, 1]) / 2, (V[:, 1] + V[:, 2]) / 2, (V[:, 2] + V[:, 0]) / 2
    return np.concatenate([np.stack(t, 1) for t in ((V[:, 0], m01, m20), (m01, V[:, 1], m12),
                                                      (m20, m12, V[:, 2]), (m01, m12, m20))])


def _level(nodes, tris, level):
    """(V, L, tau, un, area, P, W) of one observation level (src/pcbqs/_numerics/sing.py:56-69)."""
    V = nodes[tris]
    e = np.roll(V, -1, axis=1) - V
    L = np.linalg.norm(e, axis=2)
    tau = e / L[..., None]
    un = np.stack([tau[..., 1], -tau[..., 0]], -1)
    area = 0.5 * np.abs(e[:, 0, 0] * (-e[:, 2, 1]) - e[:, 0, 1] * (-e[:, 2, 0]))
    bar, weights = BAR6, W6
    if level:
        ref = np.eye(3)[None]
        for _ in range(level):
            ref = _split4(ref)
        bar = np.einsum('qk,tkc->tqc', BAR6, ref).reshape(-1, 3)
        weights = np.tile(W6, len(ref)) / len(ref)
    P = np.einsum('qk,tkc->tqc', bar, V)
    return V, L, tau, un, area, P, np.array(weights, copy=True)


@dataclass(frozen=True)
class FillGeometry:
    nodes: np.ndarray
    tris: np.ndarray
    z: np.ndarray
    cent: np.ndarray
    radius: np.ndarray
    points: np.ndarray
    levels: dict

    @classmethod
    def prepare(cls, nodes, tris, z, outer_levels):
        nodes, tris, z = np.asarray(nodes, float), np.asarray(tris), np.asarray(z, float)
        if (nodes.ndim != 2 or nodes.shape[1] != 2 or not len(nodes) or
                tris.ndim != 2 or tris.shape[1] != 3 or not len(tris) or tris.dtype.kind not in 'iu' or
                z.shape != (len(tris),) or not np.isfinite(nodes).all() or not np.isfinite(z).all() or
                tris.min() < 0 or tris.max() >= len(nodes)):
            raise ValueError('fill requires finite planar nodes, integer triangles and one z per triangle')
        nodes = np.ascontiguousarray(nodes - nodes.mean(0))
        tris = np.array(tris, dtype=np.int32, order='C', copy=True)
        z = np.array(z, order='C', copy=True)
        V = nodes[tris]
        edge = V[:, 1:] - V[:, :1]
        if (edge[:, 0, 0] * edge[:, 1, 1] - edge[:, 0, 1] * edge[:, 1, 0] <= 0).any():
            raise ValueError('fill triangles must be nondegenerate and counter-clockwise')
        cent = V.mean(1)
        radius = np.linalg.norm(V - cent[:, None], axis=2).max(1)
        points = np.einsum('qk,tkc->tqc', BAR3, V)
        levels = {level: _level(nodes, tris, level) for level in {0, outer_levels}}
        for array in (nodes, tris, z, cent, radius, points, *(a for g in levels.values() for a in g)):
            array.setflags(write=False)
        return cls(nodes, tris, z, cent, radius, points, levels)

    @property
    def area(self):
        return self.levels[0][4]


class RWG:
    """RWG basis on the interior edges of a planar mesh (topology and the combine's arrays)."""

    def __init__(self, nodes, tris):
        nodes, tris = np.asarray(nodes, float), np.asarray(tris)
        self.nodes, self.tris = nodes, tris
        V = nodes[tris]
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        self.area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
        if not (self.area > 0).all():
            raise ValueError('triangles must be counter-clockwise')
        self.cent = V.mean(1)
        nt = len(tris)
        a, b = tris, np.roll(tris, -1, axis=1)                     # half-edge i of t: (tris[t,i], tris[t,i+1])
        lo, hi = np.minimum(a, b).ravel().astype(np.int64), np.maximum(a, b).ravel().astype(np.int64)
        key = lo * (int(nodes.shape[0]) + 1) + hi
        order = np.argsort(key, kind='stable')                     # equal keys keep appearance order
        sk = order[np.r_[True, key[order][1:] != key[order][:-1]]] # first appearance of each key
        starts = np.flatnonzero(np.r_[True, key[order][1:] != key[order][:-1]])
        counts = np.diff(np.r_[starts, len(order)])
        interior = counts == 2
        first = sk[interior]
        second = order[starts[interior] + 1]
        rank = np.argsort(first, kind='stable')                    # edges in first-appearance order
        first, second = first[rank], second[rank]
        t1, i1 = np.divmod(first, 3)
        t2, i2 = np.divmod(second, 3)
        self.Ne = len(first)
        self.tp, self.tm = t1.astype(np.int64), t2.astype(np.int64)
        self.vp, self.vm = nodes[tris[t1, (i1 + 2) % 3]], nodes[tris[t2, (i2 + 2) % 3]]
        u, v = lo[first], hi[first]
        self.ends = np.stack([u, v], 1)
        pu, pv = nodes[u], nodes[v]
        self.l = np.linalg.norm(pv - pu, axis=1)
        self.nt = nt

    def combine_arrays(self):
        """(tri (2,Ne) int32, alpha (2,Ne), vertex (2,Ne,2)): static.py:116-118."""
        tri = np.ascontiguousarray(np.stack([self.tp, self.tm]), np.int32)
        alpha = np.ascontiguousarray(np.stack([self.l / (2 * self.area[self.tp]), -self.l / (2 * self.area[self.tm])]))
        vertex = np.ascontiguousarray(np.stack([self.vp, self.vm]), float)
        return tri, alpha, vertex
