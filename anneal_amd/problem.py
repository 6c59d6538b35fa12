"""Device versions of the Problem steps that feed the SIMP loop (anneal/problem.py:276-326).

``connected(prob, mask)`` is ``Problem._connected`` and ``boundary(prob,
active)`` is ``Problem.boundary``: same arguments (``prob`` is any object with
the attributes those methods read -- ``grid``, ``anchors`` with ``name``,
``role``, ``dofs``, ``loads``, ``region`` and optional ``_bores``, ``cases``,
``log``), same return values, same warnings and the same SpecError.  The
anchor regions (anneal/kernel.py signed distances) are evaluated on the GPU
at every candidate node, the face-connected components are labelled on the
GPU, and the fixed-dof bits and nodal loads are formed there;
``boundary_device`` hands them over without leaving the device.
"""
import numpy as np

from . import _lib

KIND_BOX, KIND_OBOX, KIND_SPHERE, KIND_CYLINDER = 0, 1, 2, 3
NPARAM = 16
MAX_ANCHORS = 32


class SpecError(ValueError):
    """An anchor or the design space does not make a solvable problem (problem.py:24)."""


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("anneal_amd needs a HIP device (MI355X); none is visible to torch")
    return torch


def shape_prims(shape):
    """Signed-distance primitives of a region object (kernel.py Box / Sphere / Cylinder / Union).

    The parameters are formed with the same numpy expressions the shapes use
    (kernel.py:62-63, 73, 99-104) so the device sees identical values."""
    if hasattr(shape, "shapes"):
        out = []
        for s in shape.shapes:
            out += shape_prims(s)
        return out
    if hasattr(shape, "ax") and hasattr(shape, "L"):
        return [(KIND_CYLINDER, [*shape.a, *shape.ax, shape.L, shape.r, shape.r_in])]
    if hasattr(shape, "frame"):
        if shape.frame is None:
            c, half = (shape.lo + shape.hi) / 2, (shape.hi - shape.lo) / 2
            return [(KIND_BOX, [*c, *half])]
        fr = shape.frame
        zs = shape.z0 + shape.size[2] / 2
        return [(KIND_OBOX, [*fr.o, *fr.x, *fr.y, *fr.z, zs, *(shape.size / 2)])]
    if hasattr(shape, "c") and hasattr(shape, "r"):
        return [(KIND_SPHERE, [*shape.c, shape.r])]
    raise TypeError(f"unsupported region type {type(shape).__name__}")


def _wall_prims(bores, h):
    """Bore walls of Anchor.nodes (problem.py:106-111): pipes of radius r +- 0.9 h."""
    out = []
    for _, a, b, r in bores:
        a = np.asarray(a, float)
        b = np.asarray(b, float)
        L = np.linalg.norm(b - a)
        ax = (b - a) / L
        out.append((KIND_CYLINDER, [*a, *ax, L, float(r + 0.9 * h), float(max(r - 0.9 * h, 1e-3))]))
    return out


def node_selector(anchor, h):
    """(primitives, threshold) of Anchor.nodes: bore walls at <= 0, else the region at <= h / 2."""
    bores = getattr(anchor, "_bores", None)
    if bores:
        return _wall_prims(bores, h), 0.0
    return shape_prims(anchor.region), 0.5 * h


class RegionTables:
    """Device tables of anchors given as [(primitives, threshold)]."""

    def __init__(self, selectors, device):
        torch = _torch()
        if len(selectors) > MAX_ANCHORS:
            raise ValueError(f"at most {MAX_ANCHORS} anchors are supported on the device")
        kinds, prm, first, thr = [], [], [0], []
        for prims, t in selectors:
            for k, p in prims:
                row = np.zeros(NPARAM)
                row[:len(p)] = np.asarray(p, dtype=float)
                kinds.append(k)
                prm.append(row)
            first.append(len(kinds))
            thr.append(float(t))
        self.n = len(selectors)
        self.kind = torch.tensor(kinds or [0], dtype=torch.int32, device=device)
        self.prm = torch.tensor(np.array(prm) if prm else np.zeros((1, NPARAM)), dtype=torch.float64, device=device)
        self.first = torch.tensor(first, dtype=torch.int32, device=device)
        self.thr = torch.tensor(thr or [0.0], dtype=torch.float64, device=device)

    def args(self):
        return (_lib.ptr(self.kind), _lib.ptr(self.prm), _lib.ptr(self.first), _lib.ptr(self.thr), self.n)


def _stream():
    return _lib.stream_ptr()


def device_mask(mask, device):
    """uint8 device copy of a boolean voxel mask (numpy or torch)."""
    torch = _torch()
    if isinstance(mask, torch.Tensor):
        return mask.to(device=device, dtype=torch.uint8).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(mask).ravel(), dtype=np.uint8)).to(device)


def select(grid, tables, emask_dev, at_nodes):
    """Anchor bits per node (at_nodes) or per voxel centre, and per-anchor counts."""
    torch = _torch()
    n = grid.nnode if at_nodes else grid.nelem
    bits = torch.empty(n, dtype=torch.int32, device=emask_dev.device)
    counts = torch.zeros(MAX_ANCHORS, dtype=torch.int64, device=emask_dev.device)
    origin = np.ascontiguousarray(grid.origin, dtype=np.float64)
    _lib.call("anneal_region_select", grid.nx, grid.ny, grid.nz, float(grid.h), _lib.ptr(origin), *tables.args(),
              int(bool(at_nodes)), _lib.ptr(emask_dev), _lib.ptr(bits), _lib.ptr(counts), _stream())
    return bits, counts[:tables.n].cpu().numpy()


def components(grid, mask_dev):
    """Face-connected components: (label [nelem] int32, smallest voxel index per component; count)."""
    torch = _torch()
    lab = torch.empty(grid.nelem, dtype=torch.int32, device=mask_dev.device)
    work = torch.empty(4, dtype=torch.int64, device=mask_dev.device)
    n = _lib.c_int64()
    _lib.call("anneal_voxel_components", grid.nx, grid.ny, grid.nz, _lib.ptr(mask_dev), _lib.ptr(lab),
              _lib.ptr(work), _lib.ctypes.byref(n), _stream())
    return lab, int(n.value)


def link(grid, mask_dev, lab, ebits, seed, want_mask=True):
    """Components holding a voxel with ebits & seed: (count, OR of their ebits, restricted mask)."""
    torch = _torch()
    keep = torch.empty(grid.nelem, dtype=torch.uint8, device=mask_dev.device)
    out = torch.empty(grid.nelem, dtype=torch.uint8, device=mask_dev.device) if want_mask else None
    work = torch.empty(4, dtype=torch.int64, device=mask_dev.device)
    nk, linked = _lib.c_int64(), _lib.ctypes.c_uint32()
    _lib.call("anneal_voxel_link", grid.nx, grid.ny, grid.nz, _lib.ptr(mask_dev), _lib.ptr(lab), _lib.ptr(ebits),
              int(seed) & 0xFFFFFFFF, _lib.ptr(keep), _lib.ptr(out), _lib.ptr(work), _lib.ctypes.byref(nk),
              _lib.ctypes.byref(linked), _stream())
    return int(nk.value), int(linked.value), out


def _like(mask, out_u8):
    """Return a device uint8 mask as the caller's type (numpy bool or torch bool)."""
    torch = _torch()
    if isinstance(mask, torch.Tensor):
        return out_u8.to(device=mask.device, dtype=torch.bool)
    return out_u8.cpu().numpy().astype(bool)


def connected(prob, mask):
    """Problem._connected (problem.py:276-300): keep the face-connected pieces of ``mask`` that touch a
    fixed anchor.  With a single piece the mask is returned unchanged."""
    torch = _torch()
    g = prob.grid
    dev = torch.device("cuda", torch.cuda.current_device())
    m = device_mask(mask, dev)
    lab, n = components(g, m)
    if n <= 1:
        return mask
    fixed = [a for a in prob.anchors if a.role == "fixed"]
    seed = 0
    ebits = torch.zeros(g.nelem, dtype=torch.int32, device=dev)
    if fixed:
        tables = RegionTables([node_selector(a, g.h) for a in fixed], dev)
        nbits, _ = select(g, tables, m, at_nodes=True)
        _lib.call("anneal_gather_node_bits", g.nx, g.ny, g.nz, _lib.ptr(nbits), _lib.ptr(m), _lib.ptr(ebits),
                  _stream())
        seed = (1 << len(fixed)) - 1
    kept, _, out = link(g, m, lab, ebits, seed)
    dropped = n - kept
    if dropped:
        prob.log(f"  WARNING: {dropped} piece(s) of the design space touch no fixed anchor and were removed.")
    return _like(mask, out)


def _no_material_error(a, g):
    try:
        lo, hi = a.region.bounds()
        span = f" Its region spans {np.round(lo, 2).tolist()} to {np.round(hi, 2).tolist()} mm;"
    except Exception:
        span = ""
    top = g.origin + g.h * np.array(g.shape)
    return SpecError(
        f"Anchor '{a.name}' touches no material.{span} the design space spans "
        f"{np.round(g.origin, 1).tolist()} to {np.round(top, 1).tolist()} mm. "
        f"Move the region onto the part surface, or make it thicker than one voxel ({g.h} mm).")


def boundary_device(prob, active):
    """Problem.boundary on the device: (fixed_bits [nnode] uint8 with bit d = dof d fixed,
    F [ncase][3 nnode] float64 in the reference's dof order, info {anchor: node count})."""
    torch = _torch()
    g = prob.grid
    dev = torch.device("cuda", torch.cuda.current_device())
    m = device_mask(active, dev)
    anchors = list(prob.anchors)
    tables = RegionTables([node_selector(a, g.h) for a in anchors], dev)
    nbits, counts = select(g, tables, m, at_nodes=True)
    info = {}
    for a, c in zip(anchors, counts):
        if c == 0:
            raise _no_material_error(a, g)
        info[a.name] = int(c)
    cases = list(prob.cases)
    cidx = {c: i for i, c in enumerate(cases)}
    dofmask = np.zeros(max(1, len(anchors)), np.uint8)
    load = np.zeros((max(1, len(cases)), max(1, len(anchors)), 3))
    for ai, (a, cnt) in enumerate(zip(anchors, counts)):
        if a.role == "fixed":
            for d in a.dofs:
                dofmask[ai] |= np.uint8(1 << d)
        else:
            for c, v in a.loads.items():
                v = np.asarray(v, float)
                for d in range(3):
                    load[cidx[c], ai, d] = v[d] / int(cnt)
    dofmask_d = torch.from_numpy(dofmask).to(dev)
    load_d = torch.from_numpy(np.ascontiguousarray(load)).to(dev)
    fixed_bits = torch.empty(g.nnode, dtype=torch.uint8, device=dev)
    F = torch.empty((max(1, len(cases)), 3 * g.nnode), dtype=torch.float64, device=dev)
    _lib.call("anneal_boundary_fields", g.nx, g.ny, g.nz, _lib.ptr(nbits), len(anchors), _lib.ptr(dofmask_d),
              _lib.ptr(load_d), len(cases), _lib.ptr(fixed_bits), _lib.ptr(F), _stream())
    return fixed_bits, F[:len(cases)], info


def boundary(prob, active):
    """Problem.boundary (problem.py:302-326): (fixed dofs sorted int64, {case: F[3 nnode]}, info)."""
    fixed_bits, F, info = boundary_device(prob, active)
    bits = fixed_bits.cpu().numpy()
    nodes = np.flatnonzero(bits)
    parts = [3 * nodes[(bits[nodes] >> d) & 1 == 1] + d for d in range(3)]
    fixed = np.sort(np.concatenate(parts)).astype(np.int64) if nodes.size else np.zeros(0, int)
    Fh = F.cpu().numpy()
    return fixed, {c: Fh[i].copy() for i, c in enumerate(prob.cases)}, info
