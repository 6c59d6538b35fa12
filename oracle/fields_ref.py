"""CPU oracle for display-mesh field sampling (test infrastructure, never product code).

Restates ``anneal/fields.py:10-37``: the npz layout written by
``anneal/run.py:101-110`` (``origin``, ``h``, ``shape``, ``u::case`` nnode x 3
float32, ``vm::case`` nelem float32) and ``corner_values``: trilinear
displacement at each triangle corner (scipy ``map_coordinates`` order 1, mode
"nearest", fields.py:26-28) and the 2x2x2 grey-dilated von Mises field looked
up with order 0 half a voxel inside the surface (fields.py:30-35), returned as
``[dx, dy, dz, vm] x 3 corners x faces`` little-endian float32 bytes.

Pinning: ``trilinear_brute`` / ``dilate_brute`` / ``nearest_brute`` below are
loop restatements of the interpolation, dilation and lookup from their
definitions (clamped coordinates, weights (1 - t, t), window [i, i+1] per axis,
round half up); the tests check the scipy path against them on random points
inside and outside the box, and check that trilinear sampling reproduces an
affine field exactly.
"""
import numpy as np
from scipy import ndimage


def load_npz(path):
    """fields.py:10-14: (npz, meta, sorted case names)."""
    z = np.load(path, allow_pickle=False)
    meta = {"origin": z["origin"], "h": float(z["h"][0]), "shape": tuple(int(v) for v in z["shape"])}
    cases = sorted(k.split("::", 1)[1] for k in z.files if k.startswith("u::"))
    return z, meta, cases


def corner_values_arrays(tri, normals, u_nodes, vm_elems, shape, h, origin):
    """fields.py:21-36 on arrays: returns float32 (3F, 4)."""
    nx, ny, nz = shape
    pts = np.asarray(tri, float).reshape(-1, 3)
    o = np.asarray(origin, float)
    u = np.asarray(u_nodes, np.float32).reshape(nz + 1, ny + 1, nx + 1, 3).transpose(2, 1, 0, 3)
    c = ((pts - o) / h).T
    disp = np.stack([ndimage.map_coordinates(u[..., k], c, order=1, mode="nearest") for k in range(3)], 1)
    vm3 = np.asarray(vm_elems, np.float32).reshape(nz, ny, nx).transpose(2, 1, 0)
    vm3 = ndimage.grey_dilation(vm3, size=(2, 2, 2)) if vm3.size else vm3
    n = np.repeat(np.asarray(normals, float), 3, axis=0)
    inside = pts - 0.6 * h * n
    ci = ((inside - o) / h - 0.5).T
    vm = ndimage.map_coordinates(vm3, ci, order=0, mode="nearest")
    return np.concatenate([disp, vm[:, None]], axis=1).astype("<f4")


def corner_values_ref(tri, normals, npz_path, case):
    """fields.py:17-37 with the mesh given as (triangles, face normals)."""
    z, meta, cases = load_npz(npz_path)
    if case not in cases:
        case = cases[0]
    out = corner_values_arrays(tri, normals, z[f"u::{case}"], z[f"vm::{case}"], meta["shape"], meta["h"],
                               meta["origin"])
    return out.tobytes(), cases, case


# ---- loop restatements used to pin the scipy path ---------------------------

def trilinear_brute(field, x):
    """field[i, j, k], point x (3,): clamp each coordinate to [0, n-1], then
    sum over the 8 corners of w_x w_y w_z field with weights (1 - t, t)."""
    idx, wts = [], []
    for d in range(3):
        n = field.shape[d]
        xd = min(max(float(x[d]), 0.0), n - 1.0)
        i0 = int(np.floor(xd))
        t = xd - i0
        idx.append((i0, min(i0 + 1, n - 1)))
        wts.append((1.0 - t, t))
    acc = 0.0
    for a in range(2):
        for b in range(2):
            for c in range(2):
                acc += float(field[idx[0][a], idx[1][b], idx[2][c]]) * wts[0][a] * wts[1][b] * wts[2][c]
    return acc


def dilate_brute(vm3):
    """2x2x2 grey dilation: out[i,j,k] = max over [i, i+1]^3 with the upper face reflected (= clamped)."""
    nx, ny, nz = vm3.shape
    out = np.empty_like(vm3)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                out[i, j, k] = vm3[i:min(i + 2, nx), j:min(j + 2, ny), k:min(k + 2, nz)].max()
    return out


def nearest_brute(field, x):
    """Order-0 lookup: clamp to [0, n-1], index floor(x + 0.5) (round half up)."""
    ijk = []
    for d in range(3):
        n = field.shape[d]
        xd = min(max(float(x[d]), 0.0), n - 1.0)
        ijk.append(int(np.floor(xd + 0.5)))
    return field[tuple(ijk)]
