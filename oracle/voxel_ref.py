"""CPU oracle of the voxel passes around the SIMP loop (test infrastructure, never product code).

Restates, in numpy/scipy, what the reference does before and after the
optimisation:

* signed-distance regions of ``anneal/kernel.py`` (Frame, Box, Sphere,
  Cylinder, Union, make_region; kernel.py:19-151);
* ``Problem._connected`` and ``Problem.boundary`` (problem.py:276-326) with
  the anchor node selection of ``Anchor.nodes`` (problem.py:103-113);
* ``export.output_field``, ``choose_level`` and ``keep_load_paths``
  (export.py:8-118).

Only ``tests/`` and ``bench.py``'s CPU-baseline leg use it.  The objects a
test builds here (``Anchor``, ``Prob``) carry exactly the attributes the
reference's Problem and Anchor expose to those functions.

Pinning (``tests/test_voxel_oracle.py``):
  * SDFs: closed forms (distance to a box corner / face, sphere and cylinder
    radii, pipe wall), sign inside/outside, Union = minimum.
  * ``label``: ``scipy.ndimage.label`` against a brute-force BFS partition.
  * ``dijkstra``: scipy's distances against a brute-force Bellman-Ford on
    small voxel graphs; the repaired field reconnects every load anchor.
  * ``choose_level``: the bracket invariant count(f > lo) > target >= count(f > hi)
    and agreement with an exhaustive search over the sorted field.
  * ``boundary``: loads of each anchor sum to its force per case, fixed dofs =
    selected nodes x dofs, SpecError on an anchor without material.
"""
import numpy as np
from scipy import ndimage


# ---- kernel.py ----------------------------------------------------------------
def _unit(v):
    v = np.asarray(v, dtype=float)
    n = np.linalg.norm(v)
    if n < 1e-12:
        raise ValueError("zero-length direction")
    return v / n


class Frame:
    """kernel.py:19-38: origin, unit z (normal), x made orthogonal, y = z x x."""

    def __init__(self, origin, z=(0, 0, 1), x=None):
        self.o = np.asarray(origin, dtype=float)
        self.z = _unit(z)
        if x is None:
            x = (1, 0, 0) if abs(self.z[0]) < 0.9 else (0, 1, 0)
        x = np.asarray(x, dtype=float)
        x = x - self.z * (x @ self.z)
        self.x = _unit(x)
        self.y = np.cross(self.z, self.x)

    def to_local(self, p):
        d = np.asarray(p, dtype=float) - self.o
        return np.stack([d @ self.x, d @ self.y, d @ self.z], axis=-1)


def _box_sdf(q, half):
    """kernel.py:41-45."""
    d = np.abs(q) - half
    outside = np.linalg.norm(np.maximum(d, 0), axis=-1)
    inside = np.minimum(np.max(d, axis=-1), 0)
    return outside + inside


class Box:
    """kernel.py:56-82: axis box (min, max) or framed box (frame, size, z0)."""

    def __init__(self, min=None, max=None, frame=None, size=None, z0=0.0):
        if frame is None:
            self.lo, self.hi = np.asarray(min, float), np.asarray(max, float)
            self.frame = None
        else:
            self.frame, self.size, self.z0 = frame, np.asarray(size, float), float(z0)

    def sdf(self, p):
        if self.frame is None:
            c, half = (self.lo + self.hi) / 2, (self.hi - self.lo) / 2
            return _box_sdf(np.asarray(p) - c, half)
        q = self.frame.to_local(p)
        q[..., 2] -= self.z0 + self.size[2] / 2
        return _box_sdf(q, self.size / 2)

    def bounds(self):
        if self.frame is None:
            return self.lo, self.hi
        s = self.size / 2
        corners = np.array([[sx, sy, self.z0 + sz] for sx in (-s[0], s[0]) for sy in (-s[1], s[1])
                            for sz in (0, self.size[2])])
        w = self.frame.o + corners[:, :1] * self.frame.x + corners[:, 1:2] * self.frame.y + corners[:, 2:3] * self.frame.z
        return w.min(0), w.max(0)


class Sphere:
    """kernel.py:84-92."""

    def __init__(self, center, r):
        self.c, self.r = np.asarray(center, float), float(r)

    def sdf(self, p):
        return np.linalg.norm(np.asarray(p) - self.c, axis=-1) - self.r

    def bounds(self):
        return self.c - self.r, self.c + self.r


class Cylinder:
    """kernel.py:95-116: capped cylinder a -> b, radius r; r_in > 0 makes a pipe."""

    def __init__(self, a, b, r, r_in=0.0):
        self.a, self.b = np.asarray(a, float), np.asarray(b, float)
        self.r, self.r_in = float(r), float(r_in)
        self.L = np.linalg.norm(self.b - self.a)
        self.ax = (self.b - self.a) / self.L

    def sdf(self, p):
        d = np.asarray(p) - self.a
        t = d @ self.ax
        rad = np.linalg.norm(d - t[..., None] * self.ax, axis=-1)
        dr = rad - self.r
        if self.r_in > 0:
            dr = np.maximum(dr, self.r_in - rad)
        dz = np.maximum(-t, t - self.L)
        return np.linalg.norm(np.maximum(np.stack([dr, dz], -1), 0), axis=-1) + np.minimum(np.maximum(dr, dz), 0)

    def bounds(self):
        return np.minimum(self.a, self.b) - self.r, np.maximum(self.a, self.b) + self.r


class Union:
    """kernel.py:118-125."""

    def __init__(self, shapes):
        self.shapes = list(shapes)

    def sdf(self, p):
        return np.min([s.sdf(p) for s in self.shapes], axis=0)

    def bounds(self):
        b = [s.bounds() for s in self.shapes]
        return np.min([x[0] for x in b], 0), np.max([x[1] for x in b], 0)


# ---- problem.py ---------------------------------------------------------------
class SpecError(ValueError):
    pass


class Anchor:
    """The attributes of problem.py:28-57 that boundary / _connected / keep_load_paths read."""

    def __init__(self, name, role, region, dofs="xyz", loads=None, bores=None):
        self.name, self.role, self.region = name, role, region
        self.dofs = ["xyz".index(c) for c in dofs]
        self.loads = {k: np.asarray(v, float) for k, v in (loads or {}).items()}
        if bores:
            self._bores = [("wall", np.asarray(a, float), np.asarray(b, float), float(r)) for a, b, r in bores]

    def nodes(self, grid, cand_nodes, cand_xyz):
        """problem.py:103-113."""
        h = grid.h
        if getattr(self, "_bores", None):
            sel = np.zeros(cand_nodes.size, bool)
            for _, a, b, r in self._bores:
                wall = Cylinder(a, b, r + 0.9 * h, max(r - 0.9 * h, 1e-3))
                sel |= wall.sdf(cand_xyz) <= 0
            return cand_nodes[sel]
        return cand_nodes[self.region.sdf(cand_xyz) <= 0.5 * h]


class Prob:
    """Holder of the Problem attributes the restated functions read."""

    def __init__(self, grid, anchors, fill, active=None, frozen=None, solid_fill=None, void_fill=None,
                 cases=None, log=None):
        n = grid.nelem
        self.grid, self.anchors = grid, list(anchors)
        self.fill = np.asarray(fill, float)
        self.active = (self.fill > 0) if active is None else np.asarray(active, bool)
        self.frozen = np.zeros(n, bool) if frozen is None else np.asarray(frozen, bool)
        self.solid_fill = np.zeros(n) if solid_fill is None else np.asarray(solid_fill, float)
        self.void_fill = np.zeros(n) if void_fill is None else np.asarray(void_fill, float)
        if cases is None:
            cases = []
            for a in self.anchors:
                for c in a.loads:
                    if c not in cases:
                        cases.append(c)
        self.cases = list(cases)
        self.log = log or (lambda *a: None)


def connected(prob, mask):
    """Problem._connected (problem.py:276-300)."""
    g = prob.grid
    lab, n = ndimage.label(g.to3d(mask))
    lab = g.to1d(lab)
    if n <= 1:
        return mask
    keep = np.zeros(n + 1, bool)
    cand = np.flatnonzero(mask)
    en = g.element_nodes(cand)
    for a in prob.anchors:
        if a.role != "fixed":
            continue
        nodes = set(a.nodes(g, np.unique(en), g.node_coords(np.unique(en))).tolist())
        if not nodes:
            continue
        hit = np.isin(en, list(nodes)).any(axis=1)
        keep[np.unique(lab[cand[hit]])] = True
    keep[0] = False
    dropped = n - int(keep[1:].sum())
    if dropped:
        prob.log(f"  WARNING: {dropped} piece(s) of the design space touch no fixed anchor and were removed.")
    return keep[lab] & mask


def boundary(prob, active):
    """Problem.boundary (problem.py:302-326)."""
    g = prob.grid
    act = np.flatnonzero(active)
    nodes = np.unique(g.element_nodes(act))
    xyz = g.node_coords(nodes)
    fixed, F = [], {c: np.zeros(3 * g.nnode) for c in prob.cases}
    info = {}
    for a in prob.anchors:
        sel = a.nodes(g, nodes, xyz)
        if sel.size == 0:
            raise SpecError(f"Anchor '{a.name}' touches no material.")
        info[a.name] = int(sel.size)
        if a.role == "fixed":
            fixed += [3 * sel + d for d in a.dofs]
        else:
            for c, v in a.loads.items():
                for d in range(3):
                    np.add.at(F[c], 3 * sel + d, v[d] / sel.size)
    fixed = np.unique(np.concatenate(fixed)) if fixed else np.zeros(0, int)
    return fixed, F, info


# ---- export.py ----------------------------------------------------------------
def output_field(prob, x):
    """export.py:8-14."""
    f = x * prob.fill
    f = np.maximum(f, np.where(prob.active, prob.solid_fill, 0.0))
    f = np.minimum(f, 1.0 - prob.void_fill)
    f[prob.frozen] = np.maximum(f[prob.frozen], prob.fill[prob.frozen])
    return f


def choose_level(prob, f, x, target=None, raw=False):
    """export.py:17-28 (``target`` overrides sum(x fill); ``raw`` returns the midpoint before the clamp)."""
    if target is None:
        target = float((x * prob.fill).sum())
    lo, hi = 0.05, 0.95
    for _ in range(40):
        t = 0.5 * (lo + hi)
        v = float((f > t).sum())
        if v > target:
            lo = t
        else:
            hi = t
    if raw:
        return 0.5 * (lo + hi)
    return float(np.clip(0.5 * (lo + hi), 0.3, 0.7))


def anchor_groups(prob):
    """export.py:40-45: elements within 2.5 h of each anchor region (active only)."""
    g = prob.grid
    centres = g.elem_centers()
    groups = []
    for a in prob.anchors:
        sel = (a.region.sdf(centres) <= 2.5 * g.h) & prob.active
        if sel.any():
            groups.append((a.name, a.role, np.flatnonzero(sel)))
    return groups


def broken(prob, groups, f, need):
    """export.py:48-52."""
    g = prob.grid
    lab, _ = ndimage.label(g.to3d(f > need))
    lab = g.to1d(lab)
    fixed_labels = {int(l) for n, r, e in groups if r == "fixed" for l in np.unique(lab[e]) if l > 0}
    return [(n, e) for n, r, e in groups if r == "load" and not (set(np.unique(lab[e]).tolist()) & fixed_labels)]


def path_graph(prob, f, need):
    """export.py:70-83: CSR graph over the active voxels and the fixed sources."""
    from scipy.sparse import coo_matrix
    g = prob.grid
    act = np.flatnonzero(prob.active)
    pos = np.full(g.nelem, -1)
    pos[act] = np.arange(act.size)
    i, j, k = g.elem_ijk(act)
    rows, cols = [], []
    for ok, step in ((i < g.nx - 1, 1), (j < g.ny - 1, g.nx), (k < g.nz - 1, g.nx * g.ny)):
        a_ = act[ok]
        b_ = a_ + step
        m = pos[b_] >= 0
        rows.append(pos[a_[m]])
        cols.append(pos[b_[m]])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    cost = np.maximum(0.0, need - f[act]) + 1e-3
    W = coo_matrix((np.concatenate([cost[c], cost[r]]), (np.concatenate([r, c]), np.concatenate([c, r]))),
                   shape=(act.size, act.size)).tocsr()
    return act, pos, W


def shortest_paths(prob, groups, f, need, pred_rule="heap"):
    """export.py:81-83: scipy dijkstra, min_only from the fixed groups' voxels.
    Returns (act, pos, dist, pred) with dist/pred over the active voxels.

    pred_rule "heap": scipy's predecessors (the reference's).  Where several
    neighbours realise a voxel's distance, scipy keeps the one its Fibonacci
    heap finalised first, which depends on the order the sources are listed in
    (tests/test_voxel_oracle.py shows the same graph and source set giving a
    different tree when the sources are reversed), not on the geometry.
    pred_rule "smallest": the defined rule the GPU path uses, same distances,
    predecessor = the smallest voxel index among the active face neighbours v
    with dist[v] + cost[u] == dist[u] (FP64, left to right), -1 for sources.
    Both rules give the same predecessor wherever it is unique."""
    from scipy.sparse.csgraph import dijkstra
    act, pos, W = path_graph(prob, f, need)
    fixed_e = [pos[e] for n, ro, e in groups if ro == "fixed"]
    sources = np.unique(np.concatenate(fixed_e)) if fixed_e else np.zeros(0, int)
    sources = sources[sources >= 0]
    dist, pred, _ = dijkstra(W, indices=sources, min_only=True, return_predecessors=True)
    if pred_rule == "smallest":
        pred = smallest_index_predecessors(prob.grid, act, pos, dist, f, need, sources)
    elif pred_rule != "heap":
        raise ValueError(pred_rule)
    return act, pos, dist, pred


def smallest_index_predecessors(g, act, pos, dist, f, need, sources):
    """Predecessor (position in act, -1 if none) = the smallest-voxel-index active face neighbour realising the
    distance; sources and unreachable voxels get -1."""
    cost = np.maximum(0.0, need - f[act]) + 1e-3
    pred = np.full(act.size, -1, dtype=np.int64)
    i, j, k = g.elem_ijk(act)
    best = np.full(act.size, np.iinfo(np.int64).max)
    for ok, step in ((i > 0, -1), (i < g.nx - 1, 1), (j > 0, -g.nx), (j < g.ny - 1, g.nx),
                     (k > 0, -g.nx * g.ny), (k < g.nz - 1, g.nx * g.ny)):
        e = act + step
        q = np.where(ok, pos[np.clip(e, 0, g.nelem - 1)], -1)
        hit = (q >= 0) & np.isfinite(dist)
        hit[hit] = dist[q[hit]] + cost[hit] == dist[hit]
        take = hit & (e < best)
        best[take] = e[take]
        pred[take] = q[take]
    pred[sources] = -1
    return pred


def keep_load_paths(prob, f, level, log=print, margin=0.1, max_drop=0.15, repair=True, pred_rule="heap"):
    """export.py:31-118 (pred_rule: see shortest_paths)."""
    g = prob.grid
    groups = anchor_groups(prob)
    f = f.copy()
    state = {"need": min(level + margin, 0.95)}
    level0 = level
    cut = broken(prob, groups, f, state["need"])
    while cut and level > level0 - max_drop + 1e-9:
        level = max(0.1, level - 0.02)
        state["need"] = min(level + margin, 0.95)
        cut = broken(prob, groups, f, state["need"])
        if level <= 0.1:
            break
    if level < level0:
        log(f"  iso level lowered {level0:.2f} -> {level:.2f} to keep load paths connected")
    if not cut:
        return f, level
    if not repair:
        log(f"  WARNING: anchor(s) {[n for n, _ in cut]} are not connected to a support and path repair is off. "
            f"Turn repair on, or raise volumeFraction or iterations.")
        return f, level
    need = state["need"]
    act, pos, dist, pred = shortest_paths(prob, groups, f, need, pred_rule)
    added = 0
    nb = [1, -1, g.nx, -g.nx, g.nx * g.ny, -g.nx * g.ny]
    for name, e in cut:
        tgt = pos[e]
        tgt = tgt[tgt >= 0]
        start = tgt[np.argmin(dist[tgt])]
        if not np.isfinite(dist[start]):
            log(f"  WARNING: anchor '{name}' cannot be reached from any support inside the design space.")
            continue
        path = []
        v = start
        while v >= 0:
            path.append(act[v])
            v = pred[v]
        path = np.array(path)
        ring = np.unique(np.concatenate([path + d for d in nb]))
        ring = ring[(ring >= 0) & (ring < g.nelem)]
        ring = ring[prob.active[ring]]
        before = f.copy()
        f[ring] = np.maximum(f[ring], np.minimum(need + 0.05, prob.fill[ring]))
        f[path] = np.maximum(f[path], np.minimum(1.0, prob.fill[path]))
        added += int((f > level).sum() - (before > level).sum())
        log(f"  reconnected anchor '{name}' to the supports along a {path.size}-voxel path "
            f"({added} voxels of material added)")
    still = [n for n, _ in broken(prob, groups, f, need)]
    if still:
        log(f"  WARNING: anchor(s) {still} are still thinly connected. Raise volumeFraction or iterations.")
    return f, level


# ---- brute-force checkers used to pin the above --------------------------------
def bfs_components(grid, mask):
    """Face-connected partition by breadth-first search: label = smallest voxel index."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    lab = np.full(grid.nelem, -1, np.int64)
    m = np.asarray(mask, bool)
    for s in range(grid.nelem):
        if not m[s] or lab[s] >= 0:
            continue
        lab[s] = s
        stack = [s]
        while stack:
            v = stack.pop()
            i, j, k = v % nx, (v // nx) % ny, v // (nx * ny)
            for ok, u in ((i > 0, v - 1), (i < nx - 1, v + 1), (j > 0, v - nx), (j < ny - 1, v + nx),
                          (k > 0, v - nx * ny), (k < nz - 1, v + nx * ny)):
                if ok and m[u] and lab[u] < 0:
                    lab[u] = s
                    stack.append(u)
    return lab


def bellman_ford(W, sources):
    """Left-to-right FP sums relaxed to the fixed point (the floating-point shortest distances)."""
    W = W.tocsr()
    n = W.shape[0]
    dist = np.full(n, np.inf)
    dist[np.asarray(sources, int)] = 0.0
    changed = True
    while changed:
        changed = False
        for v in range(n):
            if not np.isfinite(dist[v]):
                continue
            for q in range(W.indptr[v], W.indptr[v + 1]):
                u = W.indices[q]
                nd = dist[v] + W.data[q]
                if nd < dist[u]:
                    dist[u] = nd
                    changed = True
    return dist
