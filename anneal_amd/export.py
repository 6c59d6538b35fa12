"""Device versions of the export steps after the SIMP loop (anneal/export.py:8-118).

``output_field``, ``choose_level`` and ``keep_load_paths`` keep the
reference's arguments, return values and log lines.  ``prob`` is any object
with the attributes they read (``grid``, ``anchors``, ``fill``, ``active``,
``frozen``, ``solid_fill``, ``void_fill``).  Fields may be numpy arrays (the
results come back as numpy) or device tensors (they stay on the device).

* output_field: one elementwise kernel.
* choose_level: the 40 bisection trials decided from 4 histogram passes.
* keep_load_paths: anchor groups from signed distances at voxel centres,
  connectivity of f > need by GPU union-find (one labelling per trial level),
  and the repair's shortest paths by GPU Bellman-Ford sweeps; the path is
  traced and the material raised on the device.  Distances equal scipy's
  Dijkstra bit for bit; where several shortest paths tie, the predecessor
  with the smallest voxel index is taken (scipy's heap order may pick another
  of the tied paths).
"""
import numpy as np

from . import _lib
from .problem import RegionTables, device_mask, select, components, link, shape_prims, _torch


def _dev():
    torch = _torch()
    return torch.device("cuda", torch.cuda.current_device())


def _f64(a, dev):
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())).to(dev)


def _back(like, t):
    torch = _torch()
    if isinstance(like, torch.Tensor):
        return t
    return t.cpu().numpy()


def output_field(prob, x):
    """export.output_field (export.py:8-14): optimized density combined with fill, solids, voids and frozen voxels."""
    torch = _torch()
    dev = _dev()
    # every argument tensor stays referenced until the kernel has been queued
    xd, fill = _f64(x, dev), _f64(prob.fill, dev)
    active, frozen = device_mask(prob.active, dev), device_mask(prob.frozen, dev)
    solid, void = _f64(prob.solid_fill, dev), _f64(prob.void_fill, dev)
    n = xd.numel()
    f = torch.empty(n, dtype=torch.float64, device=dev)
    _lib.call("anneal_output_field", n, _lib.ptr(xd), _lib.ptr(fill), _lib.ptr(active), _lib.ptr(solid),
              _lib.ptr(void), _lib.ptr(frozen), _lib.ptr(f), _lib.stream_ptr())
    return _back(x, f)


def choose_level(prob, f, x):
    """export.choose_level (export.py:17-28): iso level keeping sum(x fill) voxels, clamped to [0.3, 0.7]."""
    torch = _torch()
    dev = _dev()
    fd, xd, fill = _f64(f, dev), _f64(x, dev), _f64(prob.fill, dev)
    work = torch.empty(int(_lib.load().anneal_choose_level_work_bytes()), dtype=torch.uint8, device=dev)
    level, target = _lib.c_double(), _lib.c_double()
    _lib.call("anneal_choose_level", fd.numel(), _lib.ptr(fd), _lib.ptr(xd), _lib.ptr(fill), _lib.ptr(work),
              _lib.ctypes.byref(level), _lib.ctypes.byref(target), _lib.stream_ptr())
    return float(level.value)


class _PathState:
    """Device buffers of one keep_load_paths call."""

    def __init__(self, prob, f):
        torch = _torch()
        self.g = g = prob.grid
        self.dev = dev = _dev()
        self.f = _f64(f, dev).clone()
        self.fill = _f64(prob.fill, dev)
        self.active = device_mask(prob.active, dev)
        # anchor groups: voxel centres within 2.5 h of the region, active only (export.py:40-45)
        anchors = list(prob.anchors)
        tables = RegionTables([(shape_prims(a.region), 2.5 * g.h) for a in anchors], dev)
        self.ebits, counts = select(g, tables, self.active, at_nodes=False)
        self.groups = [(a.name, a.role, i) for i, (a, c) in enumerate(zip(anchors, counts)) if c > 0]
        self.fixed_bits = sum(1 << i for _, r, i in self.groups if r == "fixed")
        self.mask = torch.empty(g.nelem, dtype=torch.uint8, device=dev)

    def broken(self, need):
        """export.py:48-52: load groups sharing no component of f > need with a fixed group."""
        g = self.g
        _lib.call("anneal_voxel_threshold", g.nelem, _lib.ptr(self.f), float(need), _lib.ptr(self.mask),
                  _lib.stream_ptr())
        lab, _ = components(g, self.mask)
        _, linked, _ = link(g, self.mask, lab, self.ebits, self.fixed_bits, want_mask=False)
        return [(n, i) for n, r, i in self.groups if r == "load" and not (linked >> i) & 1]

    def count_above(self, t):
        torch = _torch()
        work = torch.empty(4, dtype=torch.int64, device=self.dev)
        c = _lib.c_int64()
        _lib.call("anneal_voxel_count_above", self.g.nelem, _lib.ptr(self.f), float(t), _lib.ptr(work),
                  _lib.ctypes.byref(c), _lib.stream_ptr())
        return int(c.value)


def shortest_paths(st, need):
    """Distances and predecessors of the repair graph (export.py:70-83) on the device."""
    torch = _torch()
    g, dev = st.g, st.dev
    dist = torch.empty(g.nelem, dtype=torch.float64, device=dev)
    pred = torch.empty(g.nelem, dtype=torch.int32, device=dev)
    stamp = torch.empty(g.nelem, dtype=torch.int32, device=dev)
    flag = torch.empty(2, dtype=torch.int32, device=dev)
    sweeps = _lib.ctypes.c_int()
    _lib.call("anneal_paths_sssp", g.nx, g.ny, g.nz, _lib.ptr(st.active), _lib.ptr(st.f), float(need),
              _lib.ptr(st.ebits), st.fixed_bits & 0xFFFFFFFF, _lib.ptr(dist), _lib.ptr(pred), _lib.ptr(stamp),
              _lib.ptr(flag), _lib.ctypes.byref(sweeps), _lib.stream_ptr())
    return dist, pred, int(sweeps.value)


def keep_load_paths(prob, f, level, log=print, margin=0.1, max_drop=0.15, repair=True):
    """export.keep_load_paths (export.py:31-118): make every load anchor reach a fixed anchor through
    material at least ``margin`` above the iso level -- first by lowering the level by up to
    ``max_drop``, then by adding material along the cheapest voxel path.  Returns (field, level)."""
    torch = _torch()
    st = _PathState(prob, f)
    g = st.g
    state = {"need": min(level + margin, 0.95)}
    level0 = level
    cut = st.broken(state["need"])
    while cut and level > level0 - max_drop + 1e-9:
        level = max(0.1, level - 0.02)
        state["need"] = min(level + margin, 0.95)
        cut = st.broken(state["need"])
        if level <= 0.1:
            break
    if level < level0:
        log(f"  iso level lowered {level0:.2f} -> {level:.2f} to keep load paths connected")
    if not cut:
        return _back(f, st.f), level
    if not repair:
        log(f"  WARNING: anchor(s) {[n for n, _ in cut]} are not connected to a support and path repair is off. "
            f"Turn repair on, or raise volumeFraction or iterations.")
        return _back(f, st.f), level
    need = state["need"]
    dist, pred, _ = shortest_paths(st, need)
    path = torch.empty(g.nelem, dtype=torch.int32, device=st.dev)
    work = torch.empty(4, dtype=torch.int64, device=st.dev)
    added = 0
    for name, gi in cut:
        start, length, dstart = _lib.c_int64(), _lib.c_int64(), _lib.c_double()
        _lib.call("anneal_paths_trace", g.nx, g.ny, g.nz, _lib.ptr(st.active), _lib.ptr(st.ebits), gi,
                  _lib.ptr(dist), _lib.ptr(pred), _lib.ptr(path), _lib.ptr(work), _lib.ctypes.byref(start),
                  _lib.ctypes.byref(length), _lib.ctypes.byref(dstart), _lib.stream_ptr())
        if length.value == 0:
            log(f"  WARNING: anchor '{name}' cannot be reached from any support inside the design space.")
            continue
        before = st.count_above(level)
        _lib.call("anneal_paths_raise", g.nx, g.ny, g.nz, _lib.ptr(path), length.value, _lib.ptr(st.active),
                  _lib.ptr(st.fill), float(need), _lib.ptr(st.f), _lib.stream_ptr())
        added += st.count_above(level) - before
        log(f"  reconnected anchor '{name}' to the supports along a {length.value}-voxel path "
            f"({added} voxels of material added)")
    still = [n for n, _ in st.broken(need)]
    if still:
        log(f"  WARNING: anchor(s) {still} are still thinly connected. Raise volumeFraction or iterations.")
    return _back(f, st.f), level


def path_arrays(prob, f, need):
    """(dist, pred) of the repair graph over all voxels (inf / -1 off the active set), for tests and benches."""
    st = _PathState(prob, f)
    dist, pred, sweeps = shortest_paths(st, need)
    return dist, pred, sweeps, st
