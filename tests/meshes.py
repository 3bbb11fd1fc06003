"""Seeded test meshes: structured two-layer grids with jittered interior nodes."""
import numpy as np


def grid_layer(nx, ny, w, h, x0=0.0, y0=0.0, jitter=0.0, rng=None):
    xs, ys = np.linspace(x0, x0 + w, nx + 1), np.linspace(y0, y0 + h, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    nodes = np.c_[X.ravel(), Y.ravel()]
    if jitter and rng is not None:
        interior = ((X > x0) & (X < x0 + w) & (Y > y0) & (Y < y0 + h)).ravel()
        nodes[interior] += rng.uniform(-jitter, jitter, (interior.sum(), 2)) * np.array([w / nx, h / ny])
    idx = lambda i, j: i * (ny + 1) + j
    tris = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris += [[a, b, c], [a, c, d]] if (i + j) % 2 == 0 else [[a, b, d], [b, c, d]]
    return nodes, np.array(tris)


def two_layer(nx=20, ny=12, w=0.02, h=0.012, gap=2e-4, seed=0, jitter=0.25):
    rng = np.random.default_rng(seed)
    n1, t1 = grid_layer(nx, ny, w, h, jitter=jitter, rng=rng)
    n2, t2 = grid_layer(nx // 2, ny // 2, w * 0.6, h * 0.5, x0=0.002, y0=0.003, jitter=jitter, rng=rng)
    nodes = np.r_[n1, n2]
    tris = np.r_[t1, t2 + len(n1)]
    z = np.r_[np.zeros(len(t1)), np.full(len(t2), gap)]
    return nodes, tris, z
