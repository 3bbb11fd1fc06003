"""Seeded boards and their meshes: the workloads of the fill and loop/port solve.

The reference meshes copper with shapely + a constrained Delaunay triangulation
(src/pcbqs/pdn/mesh.py:175, src/pcbqs/_numerics/geo.py:249); neither is available here, so
this module builds structured meshes instead: each copper layer is a grid of square cells
split into two counter-clockwise triangles (diagonals alternating like a chequerboard), and
a cutout removes the triangles whose centroid lies inside it. What the solve needs from a
mesh is kept as the reference builds it (src/pcbqs/mqs.py:347-425):
  - one sheet per layer at its z, thickness t; triangles stacked layer by layer, each layer on
    its own nodes;
  - a terminal is the triangles whose centroid lies in its pad box, merged into one node;
  - a via joins the node nearest its centre on each layer it spans; the triangles touching
    that node (its fan) are merged into one node;
  - branches: the RWG edges (node of T+ -> node of T-), then the via segments, then the ports
    (current sources between two terminals).

  board_c1(...)   a 20 x 5 mm strip on layer 1 over a 20 x 20 mm plane on layer 2, 0.2 mm apart,
                  a via at the strip's far end, one port between the strip's start and the plane
  board_c2(seed)  a 60 x 60 mm plane pair with 8-32 circular cutouts (r 0.5-3 mm) and 16-64
                  vias (0.3 mm drill) at uniform positions, ~20k triangles (~30k RWG), 4 ports
  board_batch     uneven multi-layer boards for batched sweeps
  board_plates    two square plates over each other with a current source between pads on them (the
                  reference's charge-solve case, tests/test_peec.py:9-18)
  board_strip     one rectangular sheet with a current source between pads at its two ends (a closed
                  copper path, tests/test_peec.py:86-101)
"""
from dataclasses import dataclass, field

import numpy as np

SIGMA_CU = 5.8e7
T_CU = 35e-6


@dataclass
class Board:
    nodes: np.ndarray               # (N, 2) m
    tris: np.ndarray                # (nt, 3) counter-clockwise
    z_of: np.ndarray                # (nt,) m
    t_of: np.ndarray                # (nt,) m
    layer_of: np.ndarray            # (nt,) layer index
    node_of: np.ndarray             # (nt,) graph node of each triangle
    terminals: dict                 # name -> graph node
    vseg: list                      # [(x, y, r, z0, z1)] m
    via_nodes: list                 # [(node a, node b)]
    ports: list                     # [(name, terminal a, terminal b)]
    sigma: float = SIGMA_CU
    plating: float = 25e-6
    info: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return int(self.node_of.max()) + 1

    @property
    def nt(self):
        return len(self.tris)


def grid_sheet(x0, y0, w, h, nx, ny):
    """(nodes (N,2), tris (2 nx ny, 3)) of a w x h rectangle at (x0, y0), counter-clockwise."""
    xs, ys = np.linspace(x0, x0 + w, nx + 1), np.linspace(y0, y0 + h, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    nodes = np.c_[X.ravel(), Y.ravel()]
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing='ij')
    i, j = i.ravel(), j.ravel()
    a, b = i * (ny + 1) + j, (i + 1) * (ny + 1) + j
    c, d = b + 1, a + 1
    even = (i + j) % 2 == 0
    t1 = np.where(even[:, None], np.c_[a, b, c], np.c_[a, b, d])
    t2 = np.where(even[:, None], np.c_[a, c, d], np.c_[b, c, d])
    tris = np.stack([t1, t2], 1).reshape(-1, 3)
    return nodes, tris


def _compact(nodes, tris):
    used = np.unique(tris)
    remap = np.full(len(nodes), -1)
    remap[used] = np.arange(len(used))
    return nodes[used], remap[tris]


def _cut(nodes, tris, circles):
    """Drop triangles whose centroid is inside any circle (x, y, r)."""
    if not len(circles):
        return nodes, tris
    cen = nodes[tris].mean(1)
    keep = np.ones(len(tris), bool)
    for x, y, r in circles:
        keep &= np.hypot(cen[:, 0] - x, cen[:, 1] - y) > r
    return _compact(nodes, tris[keep])


def _keep_rects(nodes, tris, rects):
    """Keep the triangles whose centroid lies in any rectangle (x0, y0, x1, y1)."""
    cen = nodes[tris].mean(1)
    keep = np.zeros(len(tris), bool)
    for x0, y0, x1, y1 in rects:
        keep |= (cen[:, 0] >= x0) & (cen[:, 0] <= x1) & (cen[:, 1] >= y0) & (cen[:, 1] <= y1)
    return _compact(nodes, tris[keep])


class _Builder:
    """Stacks layers and resolves terminals, vias and ports into graph nodes."""

    def __init__(self):
        self.nodes, self.tris, self.z, self.t, self.layer = [], [], [], [], []
        self.n_nodes = 0
        self.n_tris = 0
        self.layer_tris = {}
        self.layer_z = {}

    def add_layer(self, key, nodes, tris, z, t=T_CU):
        self.nodes.append(nodes)
        self.tris.append(tris + self.n_nodes)
        self.layer_tris[key] = np.arange(self.n_tris, self.n_tris + len(tris))
        self.layer_z[key] = z
        self.z.append(np.full(len(tris), z))
        self.t.append(np.full(len(tris), t))
        self.layer.append(np.full(len(tris), key))
        self.n_nodes += len(nodes)
        self.n_tris += len(tris)

    def build(self, terminals, vias, ports, **info):
        nodes, tris = np.vstack(self.nodes), np.vstack(self.tris)
        z_of, t_of, layer_of = np.concatenate(self.z), np.concatenate(self.t), np.concatenate(self.layer)
        cen = nodes[tris].mean(1)
        term_tris = {}
        for name, (lay, (x0, y0, x1, y1)) in terminals.items():
            idx = self.layer_tris[lay]
            c = cen[idx]
            inside = idx[(c[:, 0] >= x0) & (c[:, 0] <= x1) & (c[:, 1] >= y0) & (c[:, 1] <= y1)]
            if not len(inside):
                raise ValueError(f'terminal {name} on none of the meshed copper')
            term_tris[name] = inside
        fans = []                                  # per via: [(layer, fan triangles)]
        for x, y, r, layers in vias:
            on = []
            for lay in layers:
                idx = self.layer_tris[lay]
                lnodes = np.unique(tris[idx])
                c = lnodes[np.argmin(np.hypot(nodes[lnodes, 0] - x, nodes[lnodes, 1] - y))]
                fan = idx[(tris[idx] == c).any(1)]
                if len(fan):
                    on.append((lay, fan))
            fans.append(on)
        parent = np.arange(len(tris))

        def find(a):
            root = a
            while parent[root] != root:
                root = parent[root]
            while parent[a] != root:
                parent[a], a = root, parent[a]
            return root

        def merge(ts):
            r0 = find(ts[0])
            for t in ts[1:]:
                parent[find(t)] = r0
        for ts in term_tris.values():
            merge(ts)
        for on in fans:
            for _, fan in on:
                merge(fan)
        roots = np.array([find(t) for t in range(len(tris))])
        _, node_of = np.unique(roots, return_inverse=True)
        vseg, via_nodes = [], []
        for (x, y, r, layers), on in zip(vias, fans):
            on = sorted(on, key=lambda lf: self.layer_z[lf[0]])
            for (la, fa), (lb, fb) in zip(on, on[1:]):
                vseg.append((x, y, r, self.layer_z[la], self.layer_z[lb]))
                via_nodes.append((int(node_of[fa[0]]), int(node_of[fb[0]])))
        term_node = {n: int(node_of[ts[0]]) for n, ts in term_tris.items()}
        return Board(nodes, tris, z_of, t_of, layer_of, node_of, term_node, vseg, via_nodes, list(ports),
                     info=dict(info, triangles=len(tris), terminal_triangles={n: len(v) for n, v in term_tris.items()}))


def board_c1(cells=None, strip=(20e-3, 5e-3), plane=(20e-3, 20e-3), gap=0.2e-3, target=2000):
    """The CPU-runnable plain case: a 20 x 5 mm strip (layer 1, z = gap) over a 20 x 20 mm plane
    (layer 0, z = 0), a via at the strip's far end, a port from the strip's start to the plane."""
    sw, sh = strip
    pw, ph = plane
    # cells of equal size on both: n_plane = (pw/h)(ph/h) 2, n_strip = (sw/h)(sh/h) 2
    h = cells or np.sqrt(2 * (sw * sh + pw * ph) / target)
    px, py = max(2, round(pw / h)), max(2, round(ph / h))
    sx, sy = max(2, round(sw / h)), max(2, round(sh / h))
    b = _Builder()
    b.add_layer(0, *grid_sheet(0, 0, pw, ph, px, py), z=0.0)
    y0 = (ph - sh) / 2
    b.add_layer(1, *grid_sheet(0, y0, sw, sh, sx, sy), z=gap)
    hx = pw / px
    ym = ph / 2
    term = {'s': (1, (-1e-9, y0 - 1e-9, 1.01 * hx, y0 + sh + 1e-9)),
            'g': (0, (-1e-9, y0 - 1e-9, 1.01 * hx, y0 + sh + 1e-9))}
    vias = [(sw - 0.5 * hx, ym, 0.15e-3, (0, 1))]
    return b.build(term, vias, [('p', 's', 'g')], config='c1', gap=gap)


def board_c2(seed=0, size=60e-3, gap=0.2e-3, target=20000, ports=4):
    """Seeded 2-layer plane pair (config 2): cutouts, vias, ports, ~target triangles."""
    rng = np.random.default_rng(seed)
    n_cut, n_via = int(rng.integers(8, 33)), int(rng.integers(16, 65))
    cuts = [(int(rng.integers(0, 2)), rng.uniform(0, size), rng.uniform(0, size), rng.uniform(0.5e-3, 3e-3))
            for _ in range(n_cut)]
    removed = sum(np.pi * r * r for _, _, _, r in cuts) / size ** 2 / 2
    n = int(round(np.sqrt(target / (4 * (1 - removed)))))
    b = _Builder()
    for lay in (0, 1):
        nodes, tris = grid_sheet(0, 0, size, size, n, n)
        nodes, tris = _cut(nodes, tris, [(x, y, r) for l, x, y, r in cuts if l == lay])
        b.add_layer(lay, nodes, tris, z=lay * gap)
    h = size / n

    def clear(x, y, margin):
        return all(np.hypot(x - cx, y - cy) > r + margin for _, cx, cy, r in cuts)
    vias = []
    while len(vias) < n_via:
        x, y = rng.uniform(h, size - h, 2)
        if clear(x, y, 2 * h):
            vias.append((x, y, 0.15e-3, (0, 1)))
    terms, port_list = {}, []
    while len(port_list) < ports:
        x, y = rng.uniform(3 * h, size - 3 * h, 2)
        if not clear(x, y, 3 * h):
            continue
        k = len(port_list)
        box = (x - 0.75 * h, y - 0.75 * h, x + 0.75 * h, y + 0.75 * h)
        terms[f'p{k}+'], terms[f'p{k}-'] = (1, box), (0, box)
        port_list.append((f'p{k}', f'p{k}+', f'p{k}-'))
    return b.build(terms, vias, port_list, config='c2', seed=seed, cutouts=n_cut, n_vias=n_via, cell=h, gap=gap)


def board_batch(seed, triangles=(1000, 8000), layers=(2, 4), ports=(4, 16), gap=0.2e-3):
    """An uneven board for batched sweeps: 2-4 layers of rectangular pours 5-40 mm joined by vias."""
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(layers[0], layers[1] + 1))
    target = int(rng.integers(triangles[0], triangles[1] + 1))
    dims = [(rng.uniform(5e-3, 40e-3), rng.uniform(5e-3, 40e-3)) for _ in range(nl)]
    W = min(d[0] for d in dims)
    H = min(d[1] for d in dims)
    area = sum(w * h for w, h in dims)
    h = np.sqrt(2 * area / target)
    b = _Builder()
    for lay, (w, hh) in enumerate(dims):
        b.add_layer(lay, *grid_sheet(0, 0, w, hh, max(2, round(w / h)), max(2, round(hh / h))), z=lay * gap)
    npt = int(rng.integers(ports[0], ports[1] + 1))
    vias = []
    for _ in range(max(4, npt)):
        x, y = rng.uniform(0.1 * W, 0.9 * W), rng.uniform(0.1 * H, 0.9 * H)
        vias.append((x, y, 0.15e-3, tuple(range(nl))))
    terms, port_list = {}, []
    for k in range(npt):
        x, y = rng.uniform(0.1 * W, 0.9 * W), rng.uniform(0.1 * H, 0.9 * H)
        la, lb = sorted(rng.choice(nl, 2, replace=False))
        box = (x - 0.75 * h, y - 0.75 * h, x + 0.75 * h, y + 0.75 * h)
        terms[f'p{k}+'], terms[f'p{k}-'] = (int(lb), box), (int(la), box)
        port_list.append((f'p{k}', f'p{k}+', f'p{k}-'))
    return b.build(terms, vias, port_list, config='batch', seed=seed, layers=nl)


def board_c5(seed=0, size=120e-3, n_vias=200, ports=8, gap=0.2e-3, target=67000):
    """Config 5: a seeded 4-layer PDN, size x size. The outer layers are planes with 8-32 circular
    cutouts (r 0.5-3 mm); each inner layer is the union of 4-8 rectangular pours (20-60 mm). One grid
    spacing serves every layer (about `target` triangles in all); n_vias through vias join every
    layer with copper at their position; `ports` ports between two such layers."""
    rng = np.random.default_rng(seed)
    cuts = {lay: [(rng.uniform(0, size), rng.uniform(0, size), rng.uniform(0.5e-3, 3e-3))
                  for _ in range(int(rng.integers(8, 33)))] for lay in (0, 3)}
    pours = {}
    for lay in (1, 2):
        rects = []
        for _ in range(int(rng.integers(4, 9))):
            w, h = rng.uniform(20e-3, 60e-3, 2)
            x, y = rng.uniform(0, size - w), rng.uniform(0, size - h)
            rects.append((x, y, x + w, y + h))
        pours[lay] = rects

    def copper(lay, x, y, margin=0.0):
        if lay in cuts:
            return (margin <= x <= size - margin and margin <= y <= size - margin and
                    all(np.hypot(x - cx, y - cy) > r + margin for cx, cy, r in cuts[lay]))
        return any(x0 + margin <= x <= x1 - margin and y0 + margin <= y <= y1 - margin
                   for x0, y0, x1, y1 in pours[lay])
    probe = rng.uniform(0, size, (4000, 2))
    cover = sum(np.mean([copper(lay, x, y) for x, y in probe]) for lay in range(4))
    n = int(round(np.sqrt(target / (2 * cover))))
    h = size / n
    b = _Builder()
    for lay in range(4):
        nodes, tris = grid_sheet(0, 0, size, size, n, n)
        if lay in cuts:
            nodes, tris = _cut(nodes, tris, cuts[lay])
        else:
            nodes, tris = _keep_rects(nodes, tris, pours[lay])
        b.add_layer(lay, nodes, tris, z=lay * gap)
    vias = []
    while len(vias) < n_vias:
        x, y = rng.uniform(2 * h, size - 2 * h, 2)
        on = tuple(lay for lay in range(4) if copper(lay, x, y, 2 * h))
        if len(on) >= 2:
            vias.append((x, y, 0.15e-3, on))
    terms, port_list = {}, []
    while len(port_list) < ports:
        x, y = rng.uniform(3 * h, size - 3 * h, 2)
        on = [lay for lay in range(4) if copper(lay, x, y, 3 * h)]
        if len(on) < 2:
            continue
        la, lb = sorted(rng.choice(on, 2, replace=False))
        k = len(port_list)
        box = (x - 0.75 * h, y - 0.75 * h, x + 0.75 * h, y + 0.75 * h)
        terms[f'p{k}+'], terms[f'p{k}-'] = (int(lb), box), (int(la), box)
        port_list.append((f'p{k}', f'p{k}+', f'p{k}-'))
    return b.build(terms, vias, port_list, config='c5', seed=seed, cell=h, gap=gap, layers=4,
                   pours={k: len(v) for k, v in pours.items()})


def board_plates(h=0.5e-3, size=10e-3, cells=10, whole=False, pad=1e-3):
    """Two size x size plates at z = 0 (layer B) and z = h (layer F), not joined by copper, with a
    current source 'I' from a pad on F to one on B (the whole plates when `whole`)."""
    b = _Builder()
    for lay, z in ((0, 0.0), (1, h)):
        b.add_layer(lay, *grid_sheet(0, 0, size, size, cells, cells), z=z)
    box = (0, 0, size, size) if whole else (0, 0, pad, pad)
    return b.build({'F': (1, box), 'B': (0, box)}, [], [('I', 'F', 'B')], config='plates', h=h)


def board_strip(length=5e-3, width=2e-3, cells=(8, 3), pad=1e-3):
    """One length x width sheet with a current source 'I' from a pad at x = 0 to one at x = length."""
    b = _Builder()
    b.add_layer(0, *grid_sheet(0, 0, length, width, *cells), z=0.0)
    return b.build({'a': (0, (0, 0, pad, width)), 'b': (0, (length - pad, 0, length, width))}, [], [('I', 'a', 'b')],
                   config='strip')
