"""Saved result fields sampled at display-mesh corners, on the GPU.

Drop-in for ``anneal/fields.py``: ``load(path)`` (fields.py:10-14) and
``corner_values(mesh, npz_path, case)`` (fields.py:17-37) keep their
signatures and return values.  The npz is read on the host (the layout of
``anneal/run.py:101-110``); the sampling runs in one HIP kernel
(``anneal_fields_corner_values``): trilinear displacement with coordinates
clamped to the node box, and the 2x2x2 grey-dilated von Mises field looked up
half a voxel inside the surface, folded into the same pass.
"""
import numpy as np

from . import _lib


def load(path):
    """fields.py:10-14: (npz, {"origin", "h", "shape"}, sorted case names)."""
    z = np.load(path, allow_pickle=False)
    meta = {"origin": z["origin"], "h": float(z["h"][0]), "shape": tuple(int(v) for v in z["shape"])}
    cases = sorted(k.split("::", 1)[1] for k in z.files if k.startswith("u::"))
    return z, meta, cases


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("anneal_amd needs a HIP device (MI355X); none is visible to torch")
    return torch


def sample_fields(triangles, face_normals, u_nodes, vm_elems, shape, h, origin):
    """[dx, dy, dz, vonMises] per triangle corner as float32 (3F, 4), computed on the device.

    triangles (F, 3, 3) and face_normals (F, 3) are the mesh's corner points
    and unit normals; u_nodes (nnode, 3) float32 and vm_elems (nelem,) float32
    are one load case of the field file.
    """
    torch = _torch()
    nx, ny, nz = (int(v) for v in shape)
    tri = np.ascontiguousarray(np.asarray(triangles, dtype=np.float64).reshape(-1, 3, 3))
    nrm = np.ascontiguousarray(np.asarray(face_normals, dtype=np.float64).reshape(-1, 3))
    nface = tri.shape[0]
    if nrm.shape[0] != nface:
        raise ValueError(f"{nface} triangles but {nrm.shape[0]} face normals")
    u = np.ascontiguousarray(np.asarray(u_nodes, dtype=np.float32).reshape(-1, 3))
    vm = np.ascontiguousarray(np.asarray(vm_elems, dtype=np.float32).ravel())
    if u.shape[0] != (nx + 1) * (ny + 1) * (nz + 1) or vm.size != nx * ny * nz:
        raise ValueError(f"field sizes {u.shape[0]} nodes / {vm.size} elements do not match shape {shape}")
    o = np.ascontiguousarray(np.asarray(origin, dtype=np.float64).ravel()[:3])
    dev = torch.device("cuda")
    out = torch.empty(3 * nface, 4, dtype=torch.float32, device=dev)
    if nface:
        u_d = torch.from_numpy(u).to(dev)
        vm_d = torch.from_numpy(vm).to(dev)
        tri_d = torch.from_numpy(tri).to(dev)
        nrm_d = torch.from_numpy(nrm).to(dev)
        _lib.call("anneal_fields_corner_values", _lib.ptr(u_d), _lib.ptr(vm_d), nx, ny, nz, float(h), _lib.ptr(o),
                  _lib.ptr(tri_d), _lib.ptr(nrm_d), nface, _lib.ptr(out), _lib.stream_ptr())
    return out.cpu().numpy()


def corner_values(mesh, npz_path, case):
    """fields.py:17-37: (float32 little-endian bytes, cases, case used).

    ``mesh`` is anything with ``triangles`` (F, 3, 3) and ``face_normals``
    (F, 3), e.g. a trimesh.Trimesh.  An unknown ``case`` falls back to the
    first one, as in the reference.
    """
    z, meta, cases = load(npz_path)
    if case not in cases:
        case = cases[0]
    out = sample_fields(mesh.triangles, mesh.face_normals, z[f"u::{case}"], z[f"vm::{case}"], meta["shape"],
                        meta["h"], meta["origin"])
    return out.astype("<f4").tobytes(), cases, case
