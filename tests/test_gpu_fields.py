"""GPU corner sampling (anneal_fields_corner_values) against the scipy oracle (fields.py:17-37)."""
import numpy as np
import pytest

from oracle.fields_ref import corner_values_arrays, corner_values_ref
from tests.test_fields_oracle import random_mesh, special_points

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,h,origin,nface", [
    ((6, 5, 4), 2.5, (-3.0, 1.0, 0.5), 300),
    ((48, 24, 24), 1.0, (0.0, 0.0, 0.0), 5000),
    ((1, 3, 2), 0.7, (10.0, -4.0, 2.0), 64),       # a one-voxel-thick grid
])
def test_corner_values_match_oracle_bitwise(shape, h, origin, nface):
    from anneal_amd.fields import sample_fields
    rng = np.random.default_rng(nface)
    nx, ny, nz = shape
    tri, nrm = random_mesh(rng, shape, h, origin, nface)
    sp = special_points(shape, h, origin)
    tri = np.concatenate([tri, sp[: (len(sp) // 3) * 3].reshape(-1, 3, 3)])
    nrm = np.concatenate([nrm, np.tile([[0.0, 0.0, 1.0]], (len(sp) // 3, 1))])
    u = rng.normal(size=((nx + 1) * (ny + 1) * (nz + 1), 3)).astype(np.float32)
    vm = rng.random(nx * ny * nz).astype(np.float32)
    print(f"corner values: shape {shape}, {tri.shape[0]} faces", flush=True)
    got = sample_fields(tri, nrm, u, vm, shape, h, origin)
    ref = corner_values_arrays(tri, nrm, u, vm, shape, h, origin)
    assert got.shape == ref.shape == (3 * tri.shape[0], 4)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_corner_values_file_api_and_empty_mesh(tmp_path):
    from anneal_amd.fields import corner_values, load

    class Mesh:
        def __init__(self, tri, nrm):
            self.triangles, self.face_normals = tri, nrm

    rng = np.random.default_rng(2)
    shape, h, origin = (8, 6, 5), 1.5, np.array([0.0, -2.0, 1.0])
    nnode = 9 * 7 * 6
    path = tmp_path / "fields.npz"
    np.savez_compressed(path, origin=origin, h=np.array([h]), shape=np.array(shape),
                        **{"u::up": rng.normal(size=(nnode, 3)).astype(np.float32),
                           "vm::up": rng.random(240).astype(np.float32),
                           "u::side": rng.normal(size=(nnode, 3)).astype(np.float32),
                           "vm::side": rng.random(240).astype(np.float32)})
    tri, nrm = random_mesh(rng, shape, h, origin, 100)
    got = corner_values(Mesh(tri, nrm), str(path), "up")
    ref = corner_values_ref(tri, nrm, str(path), "up")
    assert got == ref
    got = corner_values(Mesh(tri, nrm), str(path), "missing")
    assert got[1:] == (["side", "up"], "side") and got == corner_values_ref(tri, nrm, str(path), "missing")
    _, meta, cases = load(str(path))
    assert meta["shape"] == shape and meta["h"] == h and cases == ["side", "up"]
    empty = corner_values(Mesh(np.zeros((0, 3, 3)), np.zeros((0, 3))), str(path), "up")
    assert empty[0] == b""
