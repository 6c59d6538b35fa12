"""Host-side packing of anchor regions into the device primitive table (anneal_amd.problem), CPU only.

The table rows must hold exactly the values the reference's shapes compute
from their own attributes (kernel.py:62-63, 73, 99-104) so the device signed
distances see identical inputs; these tests restate the packing independently
and check the row a device kernel would read."""
import numpy as np
import pytest

from oracle import voxel_ref as V


def test_axis_box_row_is_centre_and_half_size():
    from anneal_amd.problem import KIND_BOX, shape_prims
    (kind, p), = shape_prims(V.Box([-1.0, 2.0, 0.5], [0.5, 11.0, 3.25]))
    assert kind == KIND_BOX
    np.testing.assert_array_equal(p, [-0.25, 6.5, 1.875, 0.75, 4.5, 1.375])


def test_framed_box_row():
    from anneal_amd.problem import KIND_OBOX, shape_prims
    fr = V.Frame([1.0, 2.0, 3.0], z=(0.3, -0.2, 0.9), x=(1, 0.4, 0))
    b = V.Box(frame=fr, size=[3.0, 2.0, 1.5], z0=-0.75)
    (kind, p), = shape_prims(b)
    assert kind == KIND_OBOX
    np.testing.assert_array_equal(p[:12], np.concatenate([fr.o, fr.x, fr.y, fr.z]))
    assert p[12] == -0.75 + 1.5 / 2          # z shift of the box centre in the frame
    np.testing.assert_array_equal(p[13:16], [1.5, 1.0, 0.75])


def test_sphere_cylinder_pipe_union_rows():
    from anneal_amd.problem import KIND_CYLINDER, KIND_SPHERE, shape_prims
    cyl = V.Cylinder([1.0, 1.0, 0.5], [6.0, 4.0, 5.0], 1.4)
    pipe = V.Cylinder([2.0, 0.0, 3.0], [2.0, 8.0, 3.0], 2.0, 1.1)
    rows = shape_prims(V.Union([V.Sphere([4.0, 3.0, 2.5], 2.3), cyl, pipe]))
    assert [k for k, _ in rows] == [KIND_SPHERE, KIND_CYLINDER, KIND_CYLINDER]
    np.testing.assert_array_equal(rows[0][1], [4.0, 3.0, 2.5, 2.3])
    np.testing.assert_array_equal(rows[1][1], [*cyl.a, *cyl.ax, cyl.L, 1.4, 0.0])
    np.testing.assert_array_equal(rows[2][1], [*pipe.a, *pipe.ax, pipe.L, 2.0, 1.1])
    assert np.linalg.norm(rows[1][1][3:6]) == pytest.approx(1.0)


def test_node_selector_thresholds_and_bore_walls():
    from anneal_amd.problem import KIND_CYLINDER, node_selector
    h = 0.75
    plain = V.Anchor("wall", "fixed", V.Box([0, 0, 0], [1, 1, 1]))
    prims, thr = node_selector(plain, h)
    assert thr == 0.5 * h and len(prims) == 1
    bored = V.Anchor("bore", "fixed", V.Sphere([0, 0, 0], 1.0), bores=[([0, 0, 0], [0, 0, 4.0], 1.2)])
    prims, thr = node_selector(bored, h)
    assert thr == 0.0
    (kind, p), = prims
    assert kind == KIND_CYLINDER
    # wall of Anchor.nodes (problem.py:108): radius r + 0.9 h outside, max(r - 0.9 h, 1e-3) inside
    np.testing.assert_array_equal(p, [0, 0, 0, 0, 0, 1.0, 4.0, 1.2 + 0.9 * h, max(1.2 - 0.9 * h, 1e-3)])


def test_unsupported_region_is_rejected():
    from anneal_amd.problem import shape_prims

    class Blob:
        pass
    with pytest.raises(TypeError):
        shape_prims(Blob())
