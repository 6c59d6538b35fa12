"""CPU checks of the field-sampling oracle (anneal/fields.py:17-37) against loop restatements."""
import numpy as np
from scipy import ndimage

from oracle.fields_ref import (corner_values_arrays, corner_values_ref, dilate_brute, nearest_brute,
                               trilinear_brute)


def random_mesh(rng, shape, h, origin, nface, spill=1.5):
    lo = np.asarray(origin) - spill * h
    hi = np.asarray(origin) + (np.asarray(shape) + spill) * h
    tri = lo + (hi - lo) * rng.random((nface, 3, 3))
    n = rng.normal(size=(nface, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return tri, n


def special_points(shape, h, origin):
    """Corners on nodes, on the upper faces, and on the x.5 rounding ties of the lookup."""
    nx, ny, nz = shape
    pts = [np.asarray(origin) + h * np.array(p, float) for p in
           [(0, 0, 0), (nx, ny, nz), (nx, 0, nz), (0.5, 1.5, 2.5), (nx - 0.5, ny - 0.5, nz), (1, 1, 1),
            (nx + 3, -2, nz / 2), (nx / 2, ny / 2, nz / 2)]]
    return np.array(pts)


def test_map_coordinates_matches_loop_restatement():
    rng = np.random.default_rng(7)
    field = rng.normal(size=(5, 4, 3)).astype(np.float32)
    xs = np.concatenate([rng.uniform(-2, 6, size=(200, 3)), [[0, 0, 0], [4, 3, 2], [4.0, 0.5, 2.0], [-1, 5, 1]]])
    lin = ndimage.map_coordinates(field, xs.T, order=1, mode="nearest")
    brute = np.array([trilinear_brute(field, x) for x in xs], np.float32)
    assert np.array_equal(lin, brute)
    near = ndimage.map_coordinates(field, xs.T, order=0, mode="nearest")
    ties = np.array([[0.5, 1.5, 0.5], [2.5, 0.5, 1.5], [-0.5, 2.5, 1.5]])
    near_t = ndimage.map_coordinates(field, ties.T, order=0, mode="nearest")
    assert np.array_equal(near, np.array([nearest_brute(field, x) for x in xs]))
    assert np.array_equal(near_t, np.array([nearest_brute(field, x) for x in ties]))


def test_trilinear_reproduces_affine_fields():
    rng = np.random.default_rng(3)
    shape = (6, 4, 5)
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    i, j, k = np.meshgrid(*[np.arange(s + 1) for s in shape], indexing="ij")
    X = np.stack([i, j, k], -1).astype(float)
    u = X @ A.T + b                                      # u[i,j,k,:]
    u_nodes = u.transpose(2, 1, 0, 3).reshape(-1, 3)     # node index i + (nx+1)(j + (ny+1)k)
    tri, nrm = random_mesh(rng, shape, 1.0, (0, 0, 0), 50, spill=0.0)
    out = corner_values_arrays(tri, nrm, u_nodes.astype(np.float32), np.zeros(np.prod(shape), np.float32),
                               shape, 1.0, (0, 0, 0))
    exact = tri.reshape(-1, 3) @ A.T + b
    assert np.abs(out[:, :3] - exact).max() < 1e-5


def test_grey_dilation_window_matches_brute_force():
    rng = np.random.default_rng(5)
    vm3 = rng.random((5, 3, 4)).astype(np.float32)
    assert np.array_equal(ndimage.grey_dilation(vm3, size=(2, 2, 2)), dilate_brute(vm3))


def test_corner_values_npz_roundtrip(tmp_path):
    rng = np.random.default_rng(11)
    shape, h, origin = (6, 5, 4), 2.5, np.array([-3.0, 1.0, 0.5])
    nnode = 7 * 6 * 5
    path = tmp_path / "fields.npz"
    u = {c: rng.normal(size=(nnode, 3)).astype(np.float32) for c in ("b", "a")}
    vm = {c: rng.random(120).astype(np.float32) for c in ("b", "a")}
    np.savez_compressed(path, origin=origin, h=np.array([h]), shape=np.array(shape),
                        **{f"u::{c}": v for c, v in u.items()}, **{f"vm::{c}": v for c, v in vm.items()})
    tri, nrm = random_mesh(rng, shape, h, origin, 40)
    data, cases, case = corner_values_ref(tri, nrm, path, "zzz")
    assert cases == ["a", "b"] and case == "a"
    out = np.frombuffer(data, "<f4").reshape(-1, 4)
    pts = tri.reshape(-1, 3)
    u3 = u["a"].reshape(5, 6, 7, 3).transpose(2, 1, 0, 3)
    vm3 = dilate_brute(vm["a"].reshape(4, 5, 6).transpose(2, 1, 0))
    n3 = np.repeat(nrm, 3, axis=0)
    for q in range(pts.shape[0]):
        x = (pts[q] - origin) / h
        for d in range(3):
            assert out[q, d] == np.float32(trilinear_brute(u3[..., d], x))
        xi = (pts[q] - 0.6 * h * n3[q] - origin) / h - 0.5
        assert out[q, 3] == nearest_brute(vm3, xi)
