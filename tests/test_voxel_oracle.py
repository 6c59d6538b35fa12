"""Pins the voxel-pass oracle (oracle/voxel_ref.py) against closed forms and brute force (CPU only)."""
import numpy as np
import pytest
from scipy import ndimage

from oracle.fem_ref import Grid
from oracle import voxel_ref as V


def _partition(lab):
    """Canonical form of a labelling: each voxel -> smallest voxel index of its class (-1 outside)."""
    lab = np.asarray(lab)
    out = np.full(lab.size, -1, np.int64)
    inside = np.flatnonzero(lab > 0) if lab.min() >= 0 else np.flatnonzero(lab >= 0)
    keys = lab[inside]
    first = {}
    for v, k in zip(inside, keys):
        first.setdefault(int(k), int(v))
    out[inside] = [first[int(k)] for k in keys]
    return out


# ---- kernel.py signed distances ----------------------------------------------
def test_box_sdf_closed_forms():
    b = V.Box([0, 0, 0], [2, 4, 6])
    p = np.array([[5.0, 2, 3],      # 3 beyond the x face
                  [1, 2, 3],        # centre: -1 (nearest face x)
                  [3, 5, 7],        # beyond a corner by (1, 1, 1)
                  [1, 3.5, 3],      # inside, 0.5 from the y face
                  [2, 2, 3]])       # on the x face
    np.testing.assert_allclose(b.sdf(p), [3.0, -1.0, np.sqrt(3.0), -0.5, 0.0], atol=1e-15)


def test_framed_box_matches_axis_box_when_rotated():
    # frame with normal z and x axis along world y: local x = world y, local y = -world x
    fr = V.Frame([1.0, 2.0, 3.0], z=(0, 0, 1), x=(0, 1, 0))
    ob = V.Box(frame=fr, size=[4.0, 2.0, 6.0], z0=-3.0)
    ab = V.Box([0.0, 0.0, 0.0], [2.0, 4.0, 6.0])
    rng = np.random.default_rng(0)
    p = rng.uniform(-3, 9, size=(500, 3))
    np.testing.assert_allclose(ob.sdf(p), ab.sdf(p), atol=1e-12)


def test_sphere_cylinder_pipe_union():
    s = V.Sphere([1, 1, 1], 2.0)
    assert s.sdf(np.array([[4.0, 1, 1]]))[0] == 1.0
    c = V.Cylinder([0, 0, 0], [0, 0, 10], 2.0)
    p = np.array([[5.0, 0, 5], [0, 0, 12], [0, 0, 5], [3, 4, 13], [1, 0, -1]])
    # radial 3; beyond the cap by 2; inside -2 (radial wins over the caps at 5); corner (3, 3); cap 1
    np.testing.assert_allclose(c.sdf(p), [3.0, 2.0, -2.0, np.sqrt(9 + 9), 1.0], atol=1e-12)
    pipe = V.Cylinder([0, 0, 0], [0, 0, 10], 2.0, 1.0)
    np.testing.assert_allclose(pipe.sdf(np.array([[1.5, 0, 5], [0.25, 0, 5]])), [-0.5, 0.75], atol=1e-12)
    u = V.Union([s, c])
    q = np.random.default_rng(1).uniform(-4, 12, size=(300, 3))
    np.testing.assert_array_equal(u.sdf(q), np.minimum(s.sdf(q), c.sdf(q)))


# ---- connected components -----------------------------------------------------
@pytest.mark.parametrize("p,seed", [(0.2, 0), (0.31, 1), (0.5, 2), (0.8, 3)])
def test_ndimage_label_is_face_connected_partition(p, seed):
    g = Grid((9, 7, 5), 1.0, (0, 0, 0))
    m = np.random.default_rng(seed).random(g.nelem) < p
    lab, _ = ndimage.label(g.to3d(m))
    np.testing.assert_array_equal(_partition(g.to1d(lab)), V.bfs_components(g, m))


# ---- shortest paths ------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dijkstra_distances_equal_bellman_ford(seed):
    rng = np.random.default_rng(seed)
    g = Grid((7, 5, 4), 1.0, (0, 0, 0))
    fill = np.where(rng.random(g.nelem) < 0.85, 1.0, 0.0)
    fill[0] = 1.0
    prob = V.Prob(g, [], fill)
    f = rng.random(g.nelem)
    need = 0.6
    act, pos, W = V.path_graph(prob, f, need)
    from scipy.sparse.csgraph import dijkstra
    src = np.unique(rng.choice(act.size, size=3, replace=False))
    d_ref = dijkstra(W, indices=src, min_only=True)
    np.testing.assert_array_equal(d_ref, V.bellman_ford(W, src))


def _beam(nx=24, ny=6, nz=6, h=1.0):
    g = Grid((nx, ny, nz), h, (0, 0, 0))
    wall = V.Anchor("wall", "fixed", V.Box([-1, -1, -1], [0.5 * h, ny * h + 1, nz * h + 1]))
    tip = V.Anchor("tip", "load", V.Box([nx * h - 0.5 * h, -1, -1], [nx * h + 1, ny * h + 1, nz * h + 1]),
                   loads={"down": [0, 0, -100.0], "side": [0, 40.0, 0]})
    return g, [wall, tip]


def test_keep_load_paths_repair_reconnects():
    g, anchors = _beam()
    rng = np.random.default_rng(3)
    prob = V.Prob(g, anchors, np.ones(g.nelem))
    f = 0.3 * rng.random(g.nelem)          # nothing above the need: every load anchor is cut
    logs = []
    f2, level = V.keep_load_paths(prob, f, 0.5, log=logs.append)
    assert level == pytest.approx(0.34)       # 8 steps of 0.02: 0.5 - 8 * 0.02 in floating point
    assert any("reconnected anchor 'tip'" in s for s in logs)
    groups = V.anchor_groups(prob)
    assert V.broken(prob, groups, f2, min(level + 0.1, 0.95)) == []


def _tie_problem(shape=(14, 8, 6), seed=0):
    """Uniform cost in voids and small corner anchors: most voxels have several equal-cost predecessors."""
    nx, ny, nz = shape
    g = Grid(shape, 1.0, (0, 0, 0))
    anchors = [V.Anchor("corner", "fixed", V.Box([-1, -1, -1], [0.5, 2.0, 2.0])),
               V.Anchor("far", "load", V.Box([nx - 0.5, ny - 2.0, nz - 2.0], [nx + 1, ny + 1, nz + 1]),
                        loads={"d": [0, 0, -1]})]
    rng = np.random.default_rng(seed)
    active = rng.random(g.nelem) < 0.95
    prob = V.Prob(g, anchors, np.ones(g.nelem), active=active)
    return prob, np.zeros(g.nelem)


def test_scipy_tie_break_depends_on_source_order():
    """The reference's predecessor among equal-cost paths is an artifact of scipy's heap order: the same graph
    and the same source set, listed in reverse, give identical distances but a different predecessor tree."""
    from scipy.sparse.csgraph import dijkstra
    prob, f = _tie_problem()
    act, pos, W = V.path_graph(prob, f, 0.6)
    groups = V.anchor_groups(prob)
    src = np.unique(np.concatenate([pos[e] for n, ro, e in groups if ro == "fixed"]))
    src = src[src >= 0]
    assert src.size > 1
    d1, p1, _ = dijkstra(W, indices=src, min_only=True, return_predecessors=True)
    d2, p2, _ = dijkstra(W, indices=src[::-1], min_only=True, return_predecessors=True)
    assert np.array_equal(d1, d2)
    assert (p1 != p2).sum() > 0.05 * act.size


def test_smallest_index_predecessor_rule_bruteforce():
    """The defined tie rule (pred_rule='smallest') against a loop over each voxel's face neighbours; where the
    realising neighbour is unique it equals scipy's predecessor."""
    prob, f = _tie_problem(seed=1)
    f[:: 7] = 0.5                               # some non-uniform costs too
    need = 0.6
    groups = V.anchor_groups(prob)
    act, pos, dist, pred = V.shortest_paths(prob, groups, f, need, pred_rule="smallest")
    _, _, dist_h, pred_h = V.shortest_paths(prob, groups, f, need, pred_rule="heap")
    assert np.array_equal(dist, dist_h)
    g = prob.grid
    cost = np.maximum(0.0, need - f) + 1e-3
    src = set(np.unique(np.concatenate([pos[e] for n, ro, e in groups if ro == "fixed"])).tolist()) - {-1}
    ties = unique = 0
    for q, e in enumerate(act):
        i, j, k = (int(v) for v in g.elem_ijk(e))
        cands = []
        for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b_, c = i + di, j + dj, k + dk
            if 0 <= a < g.nx and 0 <= b_ < g.ny and 0 <= c < g.nz:
                v = a + g.nx * (b_ + g.ny * c)
                if pos[v] >= 0 and np.isfinite(dist[q]) and dist[pos[v]] + cost[e] == dist[q]:
                    cands.append(v)
        want = -1 if (q in src or not cands) else pos[min(cands)]
        assert pred[q] == want
        if q not in src and len(cands) == 1:
            unique += 1
            assert pred_h[q] == pred[q]
        ties += len(cands) > 1
    assert ties > 0.1 * act.size and unique > 0


def test_keep_load_paths_with_defined_rule_reconnects():
    prob, f = _tie_problem(shape=(20, 10, 8), seed=2)
    logs = []
    f2, level = V.keep_load_paths(prob, f, 0.5, log=logs.append, pred_rule="smallest")
    assert any("reconnected anchor 'far'" in s for s in logs)
    assert V.broken(prob, V.anchor_groups(prob), f2, min(level + 0.1, 0.95)) == []


# ---- iso level -----------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_choose_level_brackets_the_sorted_field(seed):
    rng = np.random.default_rng(seed)
    n = 5000
    f = rng.random(n)
    target = float(rng.uniform(0.1, 0.9) * n)
    mid = V.choose_level(None, f, None, target=target, raw=True)
    # count(f > t) <= target  <=>  t >= s[m], s sorted descending, m = floor(target)
    s = np.sort(f)[::-1]
    m = int(np.floor(target))
    star = float(np.clip(s[m], 0.05, 0.95))
    assert abs(mid - star) <= 0.9 * 2.0 ** -39


# ---- boundary / connected --------------------------------------------------------
def test_boundary_loads_sum_to_anchor_forces():
    g, anchors = _beam()
    prob = V.Prob(g, anchors, np.ones(g.nelem))
    fixed, F, info = V.boundary(prob, prob.active)
    # node layers x = 0 and x = 1: the second sits exactly h / 2 outside the box (selected by <=)
    assert info["wall"] == 2 * (g.ny + 1) * (g.nz + 1)
    assert fixed.size == 3 * info["wall"]
    np.testing.assert_allclose(F["down"].reshape(-1, 3).sum(0), [0, 0, -100.0], atol=1e-12)
    np.testing.assert_allclose(F["side"].reshape(-1, 3).sum(0), [0, 40.0, 0], atol=1e-12)
    far = V.Anchor("far", "fixed", V.Box([100, 100, 100], [101, 101, 101]))
    with pytest.raises(V.SpecError, match="touches no material"):
        V.boundary(V.Prob(g, anchors + [far], np.ones(g.nelem)), np.ones(g.nelem, bool))


def test_connected_drops_pieces_without_support():
    g, anchors = _beam()
    fill = np.ones(g.nelem)
    i, j, k = g.elem_ijk(np.arange(g.nelem))
    mask = (i != 12)                          # two pieces: the wall half and the tip half
    logs = []
    prob = V.Prob(g, anchors, fill, log=logs.append)
    out = V.connected(prob, mask)
    np.testing.assert_array_equal(out, mask & (i < 12))
    assert logs and "1 piece(s)" in logs[0]
    assert V.connected(prob, np.ones(g.nelem, bool)).all()
