"""GPU voxel passes (anneal_amd.problem / anneal_amd.export) against the CPU oracle.

Problem._connected / Problem.boundary (problem.py:276-326) and
export.output_field / choose_level / keep_load_paths (export.py:8-118).
Integer, mask and index results must be identical; signed distances are
bit-identical for axis boxes and spheres (the framed box and cylinder use a
BLAS product in the reference, so they are compared to 1e-12).
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle.fem_ref import Grid
from oracle import voxel_ref as V

pytestmark = pytest.mark.gpu


def _min_index_partition(g, mask):
    """ndimage labelling of mask as 'smallest voxel index of the component' (-1 outside)."""
    lab, n = ndimage.label(g.to3d(mask))
    lab = g.to1d(lab).astype(np.int64)
    first = np.full(n + 1, np.iinfo(np.int64).max)
    np.minimum.at(first, lab, np.arange(g.nelem))
    out = first[lab]
    out[lab == 0] = -1
    return out, n


def _shapes():
    fr = V.Frame([3.1, 2.2, 1.7], z=(0.3, -0.2, 0.9), x=(1, 0.4, 0))
    return {
        "box": V.Box([-1, -1, -1], [0.5, 11, 11]),
        "box2": V.Box([1.25, 0.5, 2.0], [4.75, 3.0, 6.5]),
        "sphere": V.Sphere([4.0, 3.0, 2.5], 2.3),
        "obox": V.Box(frame=fr, size=[3.0, 2.0, 1.5], z0=-0.75),
        "cyl": V.Cylinder([1.0, 1.0, 0.5], [6.0, 4.0, 5.0], 1.4),
        "pipe": V.Cylinder([2.0, 0.0, 3.0], [2.0, 8.0, 3.0], 2.0, 1.1),
        "union": V.Union([V.Sphere([1, 1, 1], 1.5), V.Box([5, 5, 5], [7, 6, 8])]),
    }


def test_region_sdf_matches_oracle():
    import torch
    from anneal_amd import _lib
    from anneal_amd.problem import RegionTables, shape_prims
    rng = np.random.default_rng(0)
    pts = np.concatenate([rng.uniform(-3, 12, size=(4000, 3)),
                          np.stack(np.meshgrid(*[np.arange(-2, 13) * 0.5] * 3, indexing="ij"), -1).reshape(-1, 3)])
    shapes = _shapes()
    dev = torch.device("cuda", 0)
    tables = RegionTables([(shape_prims(s), 0.0) for s in shapes.values()], dev)
    pd = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    out = torch.empty((len(shapes), len(pts)), dtype=torch.float64, device=dev)
    print(f"region sdf: {len(shapes)} regions x {len(pts)} points", flush=True)
    _lib.call("anneal_region_sdf", len(pts), _lib.ptr(pd), *tables.args()[:3], len(shapes), _lib.ptr(out),
              _lib.stream_ptr())
    got = out.cpu().numpy()
    for row, (name, s) in zip(got, shapes.items()):
        ref = s.sdf(pts)
        if name in ("box", "box2", "sphere"):
            assert np.array_equal(row, ref), (name, np.abs(row - ref).max())
        elif name == "union":
            assert np.array_equal(row, ref), name
        else:
            np.testing.assert_allclose(row, ref, rtol=0, atol=1e-12, err_msg=name)


def _anchor_set(h):
    s = _shapes()
    return [V.Anchor("wall", "fixed", s["box"], dofs="xz"),
            V.Anchor("pad", "load", s["box2"], loads={"a": [1.0, -2.0, 3.0], "b": [0.0, 0.0, -7.5]}),
            V.Anchor("ball", "load", s["sphere"], loads={"a": [0.5, 0.25, 0.0]}),
            V.Anchor("tilt", "fixed", s["obox"], dofs="y"),
            V.Anchor("rod", "load", s["cyl"], loads={"b": [2.0, 2.0, 2.0]}),
            V.Anchor("bore", "fixed", s["pipe"], bores=[([2.0, 0.0, 3.0], [2.0, 6.0, 3.0], 1.2)]),
            V.Anchor("two", "load", s["union"], loads={"a": [-1.0, 0.0, 0.0], "b": [0.0, 1.0, 0.0]})]


@pytest.mark.parametrize("shape,h,origin,frac", [
    ((12, 10, 10), 1.0, (0.0, 0.0, 0.0), 1.0),
    ((16, 12, 14), 0.5, (0.0, 0.0, 0.0), 0.7),
    ((9, 8, 11), 0.75, (-0.5, 0.25, 0.0), 0.85),
])
def test_node_selection_and_boundary_match_oracle(shape, h, origin, frac):
    from anneal_amd.problem import boundary
    g = Grid(shape, h, origin)
    rng = np.random.default_rng(int(h * 100))
    active = rng.random(g.nelem) < frac
    prob = V.Prob(g, _anchor_set(h), np.ones(g.nelem), active=active)
    print(f"boundary: {shape} h {h}", flush=True)
    try:
        ref = V.boundary(prob, active)
    except V.SpecError as e:
        with pytest.raises(Exception, match="touches no material"):
            boundary(prob, active)
        pytest.skip(f"oracle raised {e}")
    fixed, F, info = boundary(prob, active)
    assert info == ref[2]
    assert np.array_equal(fixed, ref[0])
    assert set(F) == set(ref[1])
    for c in F:
        assert np.array_equal(F[c], ref[1][c]), c


def test_boundary_reference_spec_and_missing_material():
    """The beam of the reference's tests/test_core.py:_spec: 40 x 10 x 10 voxels of 1 mm, wall and tip."""
    from anneal_amd.problem import boundary, SpecError
    g = Grid((40, 10, 10), 1.0, (0.0, 0.0, 0.0))
    wall = V.Anchor("wall", "fixed", V.Box([-1, -1, -1], [0.5, 11, 11]))
    tip = V.Anchor("tip", "load", V.Box([39.5, -1, -1], [41, 11, 11]), loads={"down": [0, 0, -100]})
    prob = V.Prob(g, [wall, tip], np.ones(g.nelem))
    fixed, F, info = boundary(prob, prob.active)
    ref = V.boundary(prob, prob.active)
    assert info == ref[2] == {"wall": 2 * 11 * 11, "tip": 2 * 11 * 11}
    assert np.array_equal(fixed, ref[0]) and np.array_equal(F["down"], ref[1]["down"])
    far = V.Anchor("tip", "load", V.Box([100, 100, 100], [101, 101, 101]), loads={"down": [0, 0, -100]})
    with pytest.raises(SpecError, match="touches no material"):
        boundary(V.Prob(g, [wall, far], np.ones(g.nelem)), np.ones(g.nelem, bool))


@pytest.mark.parametrize("shape,p,seed", [
    ((37, 23, 19), 0.2, 0), ((37, 23, 19), 0.3116, 1), ((37, 23, 19), 0.5, 2), ((37, 23, 19), 0.9, 3),
    ((1, 1, 50), 0.7, 4), ((64, 1, 1), 1.0, 5), ((192, 96, 96), 0.3116, 6),
])
def test_components_match_ndimage(shape, p, seed):
    import torch
    from anneal_amd.problem import components
    g = Grid(shape, 1.0, (0, 0, 0))
    m = np.random.default_rng(seed).random(g.nelem) < p
    print(f"components: {shape} p {p}", flush=True)
    lab, n = components(g, torch.from_numpy(m.astype(np.uint8)).cuda())
    ref, nref = _min_index_partition(g, m)
    assert n == nref
    assert np.array_equal(lab.cpu().numpy().astype(np.int64), ref)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_connected_matches_oracle(seed):
    from anneal_amd.problem import connected
    g = Grid((30, 10, 10), 1.0, (0, 0, 0))
    rng = np.random.default_rng(seed)
    i, j, k = g.elem_ijk(np.arange(g.nelem))
    mask = (rng.random(g.nelem) < 0.75) & (i != 15)
    logs, logs_ref = [], []
    anchors = [V.Anchor("wall", "fixed", V.Box([-1, -1, -1], [0.5, 11, 11])),
               V.Anchor("hole", "fixed", V.Sphere([25.0, 5.0, 5.0], 1.0), bores=[([25, 5, 0], [25, 5, 10], 1.5)]),
               V.Anchor("tip", "load", V.Box([29.5, -1, -1], [31, 11, 11]), loads={"d": [0, 0, -1]})]
    got = connected(V.Prob(g, anchors, np.ones(g.nelem), log=logs.append), mask)
    ref = V.connected(V.Prob(g, anchors, np.ones(g.nelem), log=logs_ref.append), mask)
    assert np.array_equal(got, ref)
    assert logs == logs_ref
    # a single piece comes back unchanged
    one = np.ones(g.nelem, bool)
    assert connected(V.Prob(g, anchors, np.ones(g.nelem)), one) is one


def test_output_field_exact():
    from anneal_amd.export import output_field
    g = Grid((20, 9, 7), 1.0, (0, 0, 0))
    rng = np.random.default_rng(4)
    n = g.nelem
    prob = V.Prob(g, [], rng.random(n), active=rng.random(n) < 0.8, frozen=rng.random(n) < 0.1,
                  solid_fill=np.where(rng.random(n) < 0.2, rng.random(n), 0.0),
                  void_fill=np.where(rng.random(n) < 0.2, rng.random(n), 0.0))
    x = rng.random(n)
    assert np.array_equal(output_field(prob, x), V.output_field(prob, x))


@pytest.mark.parametrize("n,seed,binary", [(5000, 0, False), (77777, 1, False), (4096, 2, True),
                                           (192 * 96 * 96, 3, False)])
def test_choose_level_matches_oracle(n, seed, binary):
    from anneal_amd.export import choose_level
    rng = np.random.default_rng(seed)
    g = Grid((n, 1, 1), 1.0, (0, 0, 0))
    fill = np.ones(n) if binary else rng.random(n)
    x = (rng.random(n) < 0.4).astype(float) if binary else rng.random(n)
    f = rng.random(n)
    f[:20] = [0.05, 0.95, 0.0, 1.0, 0.5] * 4          # values on the first trial thresholds and the ends
    prob = V.Prob(g, [], fill)
    print(f"choose_level: n {n}", flush=True)
    assert choose_level(prob, f, x) == V.choose_level(prob, f, x)
    # every trial decided at a level inside (0.3, 0.7): compare the unclamped bisection too
    f2 = 0.3 + 0.4 * rng.random(n)
    assert choose_level(prob, f2, x) == V.choose_level(prob, f2, x)


def _path_problem(shape, seed, blobs=False):
    g = Grid(shape, 1.0, (0, 0, 0))
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    anchors = [V.Anchor("wall", "fixed", V.Box([-1, -1, -1], [0.5, ny + 1, nz + 1])),
               V.Anchor("tip", "load", V.Box([nx - 0.5, -1, -1], [nx + 1, ny + 1, nz + 1]), loads={"d": [0, 0, -1]}),
               V.Anchor("top", "load", V.Box([nx / 2 - 1, ny - 0.5, nz / 2 - 1], [nx / 2 + 1, ny + 1, nz / 2 + 1]),
                        loads={"d": [0, -1, 0]})]
    fill = np.where(rng.random(g.nelem) < 0.9, 1.0, 0.6)
    active = rng.random(g.nelem) < 0.97
    prob = V.Prob(g, anchors, fill, active=active)
    f = 0.3 * rng.random(g.nelem)
    if blobs:
        i, j, k = g.elem_ijk(np.arange(g.nelem))
        f[(i < nx // 4) | (i > 3 * nx // 4)] = 1.0       # ties: many equal-cost paths inside the blobs
    return prob, f


@pytest.mark.parametrize("shape,seed", [((24, 8, 8), 0), ((30, 10, 12), 1), ((96, 48, 48), 2)])
def test_keep_load_paths_matches_oracle(shape, seed):
    """Random costs everywhere: the shortest path of every anchor is unique, so the repaired field is identical."""
    from anneal_amd.export import keep_load_paths
    prob, f = _path_problem(shape, seed)
    logs, logs_ref = [], []
    print(f"keep_load_paths: {shape}", flush=True)
    got, lvl = keep_load_paths(prob, f, 0.5, log=logs.append)
    ref, lvl_ref = V.keep_load_paths(prob, f, 0.5, log=logs_ref.append)
    assert lvl == lvl_ref
    assert logs == logs_ref
    assert np.array_equal(got, ref)


def test_keep_load_paths_distances_with_ties():
    """Equal-cost paths: distances are still bit-identical to scipy's Dijkstra; the repair reconnects."""
    from anneal_amd.export import keep_load_paths, path_arrays
    prob, f = _path_problem((40, 12, 12), 3, blobs=True)
    need = 0.6
    dist, pred, sweeps, st = path_arrays(prob, f, need)
    groups = V.anchor_groups(prob)
    act, pos, dref, _ = V.shortest_paths(prob, groups, f, need)
    d = dist.cpu().numpy()
    assert np.array_equal(d[act], dref)
    assert np.all(np.isinf(d[~prob.active]))
    # every predecessor is a neighbour that realises the distance
    p = pred.cpu().numpy()
    u = np.flatnonzero(p >= 0)
    cost = np.maximum(0.0, need - f) + 1e-3
    assert np.array_equal(d[p[u]] + cost[u], d[u])
    logs = []
    got, lvl = keep_load_paths(prob, f, 0.5, log=logs.append)
    assert V.broken(prob, groups, got, min(lvl + 0.1, 0.95)) == []
    assert not any("still thinly connected" in s for s in logs)


def test_keep_load_paths_on_uniform_voids_matches_defined_rule():
    """Uniform cost in the voids (ties almost everywhere): predecessors, repaired field and log lines are
    identical to the oracle's repair with the defined tie rule (smallest voxel index among the neighbours
    that realise the distance).  The reference's own tie choice is scipy's heap order, which changes with the
    order of the source list (tests/test_voxel_oracle.py), so ties are pinned to this rule instead."""
    from anneal_amd.export import keep_load_paths, path_arrays
    nx, ny, nz = 40, 16, 12
    g = Grid((nx, ny, nz), 1.0, (0, 0, 0))
    anchors = [V.Anchor("corner", "fixed", V.Box([-1, -1, -1], [0.5, 3.0, 3.0])),
               V.Anchor("far", "load", V.Box([nx - 0.5, ny - 3.0, nz - 3.0], [nx + 1, ny + 1, nz + 1]),
                        loads={"d": [0, 0, -1]}),
               V.Anchor("mid", "load", V.Box([nx / 2 - 1, ny - 0.5, -1], [nx / 2 + 1, ny + 1, 2.0]),
                        loads={"d": [0, -1, 0]})]
    rng = np.random.default_rng(4)
    active = rng.random(g.nelem) < 0.95
    prob = V.Prob(g, anchors, np.ones(g.nelem), active=active)
    f = np.zeros(g.nelem)
    f[rng.random(g.nelem) < 0.02] = 1.0        # scattered solid voxels, voids uniform
    need = 0.6
    dist, pred, _, _ = path_arrays(prob, f, need)
    groups = V.anchor_groups(prob)
    act, pos, dref, pref = V.shortest_paths(prob, groups, f, need, pred_rule="smallest")
    d, p = dist.cpu().numpy(), pred.cpu().numpy()
    assert np.array_equal(d[act], dref)
    want = np.where(pref >= 0, act[np.maximum(pref, 0)], -1)
    assert np.array_equal(p[act], want)
    logs, logs_ref = [], []
    got, lvl = keep_load_paths(prob, f, 0.5, log=logs.append)
    ref, lvl_ref = V.keep_load_paths(prob, f, 0.5, log=logs_ref.append, pred_rule="smallest")
    print("\n".join(logs), flush=True)
    assert lvl == lvl_ref and logs == logs_ref
    assert any("reconnected anchor" in s for s in logs)
    assert np.array_equal(got, ref)


def test_keep_load_paths_lowering_and_no_repair():
    from anneal_amd.export import keep_load_paths
    prob, f = _path_problem((24, 8, 8), 5)
    g = prob.grid
    i, j, k = g.elem_ijk(np.arange(g.nelem))
    f2 = f.copy()
    f2[(j == 4) & (k == 4)] = 0.55                      # a bar that connects once the level drops to 0.44
    f2[(i == 12) & (k == 4)] = 0.55
    for field, repair in ((f2, True), (f, False)):
        logs, logs_ref = [], []
        got, lvl = keep_load_paths(prob, field, 0.5, log=logs.append, repair=repair)
        ref, lvl_ref = V.keep_load_paths(prob, field, 0.5, log=logs_ref.append, repair=repair)
        assert (lvl, logs) == (lvl_ref, logs_ref)
        assert np.array_equal(got, ref)


def test_voxel_passes_at_192x96x96_match_oracle():
    """The full cantilever grid: boundary, connectivity, iso level and the repaired field, all identical."""
    from anneal_amd.export import choose_level, keep_load_paths, output_field
    from anneal_amd.problem import boundary, connected
    shape = (192, 96, 96)
    g = Grid(shape, 1.0, (0, 0, 0))
    nx, ny, nz = shape
    rng = np.random.default_rng(7)
    anchors = [V.Anchor("wall", "fixed", V.Box([-1, -1, -1], [0.5, ny + 1, nz + 1])),
               V.Anchor("tip", "load", V.Box([nx - 0.5, -1, -1], [nx + 1, ny + 1, nz + 1]), loads={"d": [0, 0, -1]}),
               V.Anchor("top", "load", V.Sphere([nx / 2, ny, nz / 2], 3.0), loads={"p": [0, -1, 0]})]
    active = rng.random(g.nelem) >= 0.03
    prob = V.Prob(g, anchors, np.where(active, 1.0, 0.0), active=active)
    print("voxel passes at 192x96x96", flush=True)
    fixed, F, info = boundary(prob, active)
    rf, rF, rinfo = V.boundary(prob, active)
    assert info == rinfo and np.array_equal(fixed, rf)
    assert all(np.array_equal(F[c], rF[c]) for c in rF)
    mask = rng.random(g.nelem) < 0.35
    assert np.array_equal(connected(prob, mask), V.connected(prob, mask))
    x = rng.random(g.nelem)
    f = output_field(prob, x)
    assert np.array_equal(f, V.output_field(prob, x))
    assert choose_level(prob, f, x) == V.choose_level(prob, f, x)
    low = 0.3 * rng.random(g.nelem)          # every anchor cut: both repaired along unique shortest paths
    logs, logs_ref = [], []
    got, lvl = keep_load_paths(prob, low, 0.5, log=logs.append)
    ref, lvl_ref = V.keep_load_paths(prob, low, 0.5, log=logs_ref.append)
    assert (lvl, logs) == (lvl_ref, logs_ref)
    assert np.array_equal(got, ref)
