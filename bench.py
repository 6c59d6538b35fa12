"""SIMP iterations per second at 192x96x96 on one MI355X, plus the rooflines of its largest kernel and of K·u.

Workload (one "step" = one SIMP iteration of the GPU path):
  cantilever 192x96x96 hex8 voxels (1.77 M elements, 5.4 M dofs), x=0 face clamped,
  unit -z point load at node (192, 48, 48), E=1, nu=0.3, Emin=1e-9, p ramp to 3,
  hat filter rmin=3 voxels, volfrac 0.3, OC move 0.2, 100-iteration schedule,
  multigrid PCG to rtol 1e-6 (FP64 true residual), all on the GPU.
W warm-up iterations run first on a throw-away run (untimed); then a fresh run
is set up (untimed, reported as setup_s) and its first K iterations are timed.
The default K = 100 is the whole schedule, so continuation (p ramp, beta
doubling) and the harder late solves are all inside the timed region.

Also measured:
  * roofline: the SIMP step's largest kernel, the fine-level Chebyshev smoothing
    step of the V-cycle (the FP32 matrix-free K·u apply with the smoothing
    epilogue, 53 B/node), at the same grid, timed with HIP events on the
    stream it is launched on, over rotating buffers larger than the 256 MB
    MALL so the bytes come from HBM; roofline_ku: the same for the plain K·u
    apply (24 B/node, the north-star K·u figure);
  * cpu_baseline: the FP64 CPU oracle (a restatement of the reference's
    CSR + CG + filter + OC loop) on a bounded sample.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SHAPE = (192, 96, 96)
# FP32 apply tile shapes (TX, TY, ZC) in the library's variant order
TILE_SHAPES = ["16,16,8", "16,16,16", "32,8,8", "32,8,16", "32,16,8", "16,16,4", "64,4,16", "16,16,2", "8,8,2",
               "8,8,4", "16,8,2"]
HBM_PEAK_GBS = 8000.0


def cantilever(shape):
    import numpy as np
    from anneal_amd.fem import Grid
    g = Grid(shape, 1.0, (0, 0, 0))
    nx, ny, nz = shape
    n = np.arange(g.nnode)
    clamp = n[n % (nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    tip = nx + (nx + 1) * (ny // 2 + (ny + 1) * (nz // 2))
    F = np.zeros(3 * g.nnode)
    F[3 * tip + 2] = -1.0
    return g, np.ones(g.nelem), np.ones(g.nelem, bool), np.zeros(g.nelem, bool), fixed, F


LARGE_SHAPE = (384, 192, 192)
LARGE_SEED = 20260913


def large_part(shape=LARGE_SHAPE, seed=LARGE_SEED, keep_frac=0.05, solid_frac=0.05):
    """Large part with seeded keep-outs, keep-solid slabs and 3 load cases.

    x=0 face clamped.  Keep-out boxes (inactive voxels, density 0) are drawn
    inside 0.15 nx <= x < 0.85 nx until they cover keep_frac of the box;
    keep-solid slabs (frozen, density 1) are full y-z slabs at the clamp face and
    at the load face covering solid_frac together.  Three unit point loads sit on
    the x = nx face at seeded nodes (middle half of y and z) with seeded
    directions uniform on the sphere.  Returns grid, fill, active, frozen, fixed
    dofs, {case: F}, {case: weight}.
    """
    import numpy as np
    from anneal_amd.fem import Grid
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = shape
    g = Grid(shape, 1.0, (0, 0, 0))
    act3 = np.ones((nx, ny, nz), bool)
    target = keep_frac * act3.size
    x_lo, x_hi = int(0.15 * nx), int(0.85 * nx)
    while (~act3).sum() < target:
        size = rng.integers([nx // 16, ny // 8, nz // 8], [nx // 6, ny // 3, nz // 3])
        lo = rng.integers([x_lo, 0, 0], [x_hi - size[0], ny - size[1], nz - size[2]])
        act3[lo[0]:lo[0] + size[0], lo[1]:lo[1] + size[1], lo[2]:lo[2] + size[2]] = False
    frz3 = np.zeros_like(act3)
    t = max(1, int(round(solid_frac * nx / 2)))
    frz3[:t] = True
    frz3[nx - t:] = True
    frz3 &= act3
    to_flat = lambda a: a.transpose(2, 1, 0).ravel()          # element index i + nx (j + ny k)
    active, frozen = to_flat(act3), to_flat(frz3)
    n = np.arange(g.nnode)
    clamp = n[n % (nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    loads, weights = {}, {}
    for c in range(3):
        j, k = rng.integers(ny // 4, 3 * ny // 4 + 1), rng.integers(nz // 4, 3 * nz // 4 + 1)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        node = nx + (nx + 1) * (j + (ny + 1) * k)
        F = np.zeros(3 * g.nnode)
        F[3 * node:3 * node + 3] = d
        loads[f"case{c}"] = F
        weights[f"case{c}"] = 1.0
    return g, np.ones(g.nelem), active, frozen, fixed, loads, weights


BATCH_SEED = 3
BENCH_WARM = 3
BATCH_SHAPES = ((64, 64, 64), (96, 48, 48), (128, 64, 64))


def settle_host():
    """Collect what set-up left behind and freeze the survivors (gc.freeze) before a timed region.

    A full collection of the Python heap scans every tracked object (about 170 000 once torch is
    imported) and holds the host for 7-36 ms, during which no launch is issued and the device idles.
    Frozen objects are left out of later collections, so a collection inside the timed region scans only
    what the region allocated.  No device work changes."""
    import gc
    gc.collect()
    gc.freeze()


def batch_problems(seed=BATCH_SEED, n=16, ncase=4, iters=40, shapes=BATCH_SHAPES, r_mm=1.5):
    """Independent parts of uneven sizes, each with a block of load cases.

    Per problem (seeded): a grid drawn from ``shapes``; one of the six faces
    clamped; ``ncase`` unit point loads at nodes of the opposite face (middle
    60 % of its extent) with directions uniform on the sphere; volume fraction
    U(0.2, 0.4).  Returns a list of SimpRun argument dicts (BatchSimp input).
    """
    import numpy as np
    from anneal_amd.fem import Grid
    rng = np.random.Generator(np.random.PCG64(seed))
    probs = []
    for _ in range(n):
        shape = tuple(shapes[int(rng.integers(len(shapes)))])
        nx, ny, nz = shape
        dims = (nx, ny, nz)
        g = Grid(shape, 1.0, (0, 0, 0))
        nodes = np.arange(g.nnode)
        coord = (nodes % (nx + 1), (nodes // (nx + 1)) % (ny + 1), nodes // ((nx + 1) * (ny + 1)))
        axis, side = int(rng.integers(3)), int(rng.integers(2))
        clamp = nodes[coord[axis] == (0 if side == 0 else dims[axis])]
        fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
        far = dims[axis] if side == 0 else 0
        loads, weights = {}, {}
        others = [a for a in range(3) if a != axis]
        for c in range(ncase):
            pos = [0, 0, 0]
            pos[axis] = far
            for a in others:
                pos[a] = int(rng.integers(int(0.2 * dims[a]), int(0.8 * dims[a]) + 1))
            node = pos[0] + (nx + 1) * (pos[1] + (ny + 1) * pos[2])
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            F = np.zeros(3 * g.nnode)
            F[3 * node:3 * node + 3] = d
            loads[f"case{c}"] = F
            weights[f"case{c}"] = 1.0
        probs.append(dict(grid=g, fill=np.ones(g.nelem), active=np.ones(g.nelem, bool),
                          frozen=np.zeros(g.nelem, bool), fixed=fixed, loads=loads, weights=weights, E0=1.0, nu=0.3,
                          vf=float(rng.uniform(0.2, 0.4)), iters=iters, r_mm=r_mm, tol=0.0, emin=1e-9, rtol=1e-6))
    return probs


def batch_bench(steps, seq_steps, log):
    """Lock-step batch of 16 problems vs the same problems stepped one after another."""
    import torch
    from anneal_amd.batch import BatchSimp
    from anneal_amd.optimize import SimpRun
    probs = batch_problems()
    b = BatchSimp(probs)
    # three untimed iterations per leg: the first records the graphs and solves from zero, the second
    # allocates the buffers of the extrapolated start, the third is the first extrapolation (its
    # PyTorch kernel is loaded on first use, ≈ 25 ms); each leg's allocations are then cached
    for _ in range(BENCH_WARM):
        b.step()
    torch.cuda.synchronize()
    settle_host()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cg = sum(sum(r.fem.cg_iters) for r in b.runs)
    comp = [r.hist[-1]["compliance"] for r in b.runs]
    groups = [[list(b.runs[idx[0]].fem.grid.shape), len(idx)] for _, idx in b.solver.groups]
    b.close()
    del b
    # unbatched: each problem on its own, same iterations
    seq = [SimpRun(*[p[k] for k in ("grid", "fill", "active", "frozen", "fixed", "loads", "weights", "E0", "nu", "vf")],
                   **{k: v for k, v in p.items() if k not in ("grid", "fill", "active", "frozen", "fixed", "loads",
                                                               "weights", "E0", "nu", "vf")}, log=lambda m: None)
           for p in probs]
    for _ in range(BENCH_WARM):
        for r in seq:
            r.step()
    torch.cuda.synchronize()
    settle_host()
    t1 = time.perf_counter()
    for _ in range(seq_steps):
        for r in seq:
            r.step()
    torch.cuda.synchronize()
    dts = time.perf_counter() - t1
    cg_seq = sum(sum(r.fem.cg_iters) for r in seq)
    for r in seq:
        r.fem.close()
    n = len(probs)
    shapes = [p["grid"].shape for p in probs]
    log(f"bench: batch {n} problems x {steps} iterations in {dt:.3f} s ({dt / steps * 1e3:.2f} ms per batch "
        f"iteration); one at a time {dts / seq_steps * 1e3:.2f} ms per round")
    return {"value": round(n * steps / dt, 4), "unit": "problem SIMP iter/s", "ms_per_step": round(1e3 * dt / steps, 3),
            "sequential_ms_per_step": round(1e3 * dts / seq_steps, 3),
            "speedup_vs_one_at_a_time": round((dts / seq_steps) / (dt / steps), 4), "pcg_iters_total": int(cg),
            "pcg_iters_total_one_at_a_time": int(cg_seq), "groups": groups,
            "shapes": [list(s) for s in shapes], "load_cases": 4, "final_compliance": comp}


COPY_BYTES = 256 * 1024 * 1024


def traffic_probe():
    """Counter-run workload: the benchmark's apply on rotating buffers, then a plain copy of known size."""
    import numpy as np
    import torch
    from anneal_amd import _lib
    from anneal_amd.fem import FEM
    g, fill, active, frozen, fixed, F = cantilever(SHAPE)
    fem = FEM(g, active, fixed, 0.3, nrhs_max=1)
    rng = np.random.default_rng(0)
    fem.assemble_device(torch.from_numpy(1e-9 + rng.random(g.nelem) ** 3).cuda())
    bufs = [(torch.randn(1, 3, fem.ldn, device="cuda"), torch.empty(1, 3, fem.ldn, device="cuda")) for _ in range(8)]
    torch.cuda.synchronize()
    for variant in range(len(TILE_SHAPES)):        # every tile shape, 8 launches each on distinct buffers
        fem.autotune(variant)
        for i in range(8):
            u, y = bufs[i]
            _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(u), _lib.ptr(y), 1)
    # the fine level's Chebyshev step on the benchmark's tile shape: 8 launches on distinct (u, b, y) sets
    fem.autotune(2)
    bset = [torch.randn(1, 3, fem.ldn, device="cuda") for _ in range(8)]
    for i in range(8):
        u, y = bufs[i]
        _lib.call("anneal_fem_smooth_step_f32", fem._ctx, _lib.ptr(u), _lib.ptr(bset[i]), _lib.ptr(y), 1)
    a = torch.empty(COPY_BYTES // 4, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"probe: {len(TILE_SHAPES)} tile shapes x 8 applies, 8 Chebyshev steps (32x8x8), 4 copies of "
          f"{COPY_BYTES} bytes", file=sys.stderr, flush=True)


def unet_flops(dims, B, cin=8, cout=3, widths=(16, 32, 64, 64)):
    """Multiply-add FLOPs (2 per MAC) of one U-Net forward on B samples of `dims` nodes."""
    L = len(widths)
    d = [tuple(dims)]
    for _ in range(1, L):
        d.append(tuple(n // 2 + 1 for n in d[-1]))
    vox = [B * a * b * c for a, b, c in d]
    macs = 0
    for l in range(L):
        ci = cin if l == 0 else widths[l - 1]
        macs += vox[l] * 27 * (ci * widths[l] + widths[l] * widths[l])
    for l in range(L - 1):
        macs += vox[l] * 27 * ((widths[l + 1] + widths[l]) * widths[l] + widths[l] * widths[l])
    macs += vox[0] * widths[0] * cout
    return 2 * macs


def unet_probe(reps=20):
    """HIP-event timings of the BF16 MFMA U-Net at the learned stage's sizes (level 1 of 192x96x96 and
    128x64x64, and a training batch of 32 at 25x13x13) and the achieved MFMA rate."""
    import torch
    from anneal_amd.learned import UNet
    net = UNet(seed=0)
    out = {"peak_tflops_bf16": 2500.0, "cases": []}
    for dims, B in [((97, 49, 49), 1), ((65, 33, 33), 1), ((25, 13, 13), 32)]:
        V = dims[0] * dims[1] * dims[2]
        x = torch.randn(B, V, 8, device="cuda").to(torch.bfloat16)
        y = net.forward(x, dims)
        dout = torch.randn_like(y)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            net.forward(x, dims, y)
        e1.record()
        torch.cuda.synchronize()
        f_ms = e0.elapsed_time(e1) / reps
        e0.record()
        for _ in range(reps):
            net.forward(x, dims, y)
            net.zero_grad()
            net.backward(dout)
        e1.record()
        torch.cuda.synchronize()
        fb_ms = e0.elapsed_time(e1) / reps
        fl = unet_flops(dims, B)
        out["cases"].append({"dims": list(dims), "batch": B, "forward_ms": round(f_ms, 4),
                             "forward_backward_ms": round(fb_ms, 4), "forward_gflop": round(fl / 1e9, 3),
                             "forward_tflops": round(fl / (f_ms * 1e-3) / 1e12, 2),
                             "forward_backward_tflops": round(3 * fl / (fb_ms * 1e-3) / 1e12, 2)})
        print(f"probe: unet {dims} x{B}: forward {f_ms:.3f} ms, fwd+bwd {fb_ms:.3f} ms", file=sys.stderr, flush=True)
    return out


class _AxisBox:
    """Axis-aligned anchor region with the attributes the device passes read (kernel.py:56-65)."""

    def __init__(self, lo, hi):
        import numpy as np
        self.lo, self.hi, self.frame = np.asarray(lo, float), np.asarray(hi, float), None

    def bounds(self):
        return self.lo, self.hi


def voxel_problem(shape=SHAPE, seed=0):
    """A 192x96x96 part for the voxel passes: wall (fixed), tip and top (load) anchors, 3 % voids."""
    from types import SimpleNamespace
    import numpy as np
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    grid = SimpleNamespace(shape=shape, nx=nx, ny=ny, nz=nz, h=1.0, origin=np.zeros(3), nelem=nx * ny * nz,
                           nnode=(nx + 1) * (ny + 1) * (nz + 1))
    anchors = [SimpleNamespace(name="wall", role="fixed", dofs=[0, 1, 2], loads={},
                               region=_AxisBox([-1, -1, -1], [0.5, ny + 1, nz + 1])),
               SimpleNamespace(name="tip", role="load", dofs=[0, 1, 2], loads={"down": np.array([0, 0, -1.0])},
                               region=_AxisBox([nx - 0.5, -1, -1], [nx + 1, ny + 1, nz + 1])),
               SimpleNamespace(name="top", role="load", dofs=[0, 1, 2], loads={"push": np.array([0, -1.0, 0])},
                               region=_AxisBox([nx / 2 - 2, ny - 0.5, nz / 2 - 2], [nx / 2 + 2, ny + 1, nz / 2 + 2]))]
    n = grid.nelem
    active = rng.random(n) >= 0.03
    prob = SimpleNamespace(grid=grid, anchors=anchors, cases=["down", "push"], fill=np.where(active, 1.0, 0.0),
                           active=active, frozen=np.zeros(n, bool), solid_fill=np.zeros(n), void_fill=np.zeros(n),
                           log=lambda *a: None)
    fields = {"x": rng.random(n), "f_low": 0.3 * rng.random(n), "mask": rng.random(n) < 0.3116}
    return prob, fields


def _timed(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps * 1e3, r


def voxel_cpu_baseline(prob, fields):
    """The oracle (numpy / scipy restatement of problem.py and export.py) on the same inputs, one core."""
    from scipy import ndimage
    from oracle import voxel_ref as V
    from oracle.fem_ref import Grid as GridRef
    g = GridRef(prob.grid.shape, prob.grid.h, prob.grid.origin)
    anchors = [V.Anchor(a.name, a.role, V.Box(a.region.lo, a.region.hi), loads=a.loads) for a in prob.anchors]
    vp = V.Prob(g, anchors, prob.fill, active=prob.active)
    out = {}
    t = time.perf_counter()
    V.boundary(vp, vp.active)
    out["boundary"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    ndimage.label(g.to3d(fields["mask"]))
    out["components"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    V.connected(vp, fields["mask"])
    out["connected"] = (time.perf_counter() - t) * 1e3
    f = V.output_field(vp, fields["x"])
    t = time.perf_counter()
    V.choose_level(vp, f, fields["x"])
    out["choose_level"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    V.keep_load_paths(vp, fields["f_low"], 0.5, log=lambda *a: None)
    out["keep_load_paths"] = (time.perf_counter() - t) * 1e3
    return out


def export_pipeline(dev, iters=100, cpu=True):
    """The export steps on a real design: the 192x96x96 cantilever after its 100-iteration schedule, with a wall
    anchor (x = 0 face) and a load anchor around the tip load; output_field -> choose_level -> keep_load_paths
    on the device, and the same sequence in the oracle on one core."""
    import numpy as np
    import torch
    from types import SimpleNamespace
    from anneal_amd import export as X
    from anneal_amd.optimize import simp
    g, fill, active, frozen, fixed, F = cantilever(SHAPE)
    res = simp(g, fill, active, frozen, fixed, {"tip": F}, {"tip": 1.0}, 1.0, 0.3, 0.3, iters=iters, r_mm=3.0,
               tol=0.0, emin=1e-9, rtol=1e-6, log=lambda *a: None)
    nx, ny, nz = SHAPE
    anchors = [SimpleNamespace(name="wall", role="fixed", dofs=[0, 1, 2], loads={},
                               region=_AxisBox([-1, -1, -1], [0.5, ny + 1, nz + 1])),
               SimpleNamespace(name="tip", role="load", dofs=[0, 1, 2], loads={"tip": np.array([0, 0, -1.0])},
                               region=_AxisBox([nx - 1.5, ny / 2 - 1.5, nz / 2 - 1.5], [nx + 1, ny / 2 + 1.5,
                                                                                        nz / 2 + 1.5]))]
    n = g.nelem
    hp = SimpleNamespace(grid=g, anchors=anchors, fill=fill, active=active, frozen=frozen, solid_fill=np.zeros(n),
                         void_fill=np.zeros(n), log=lambda *a: None)
    dp = SimpleNamespace(**vars(hp))
    for k in ("fill", "solid_fill", "void_fill"):
        setattr(dp, k, torch.from_numpy(np.ascontiguousarray(getattr(hp, k), dtype=np.float64)).to(dev))
    for k in ("active", "frozen"):
        setattr(dp, k, torch.from_numpy(getattr(hp, k).astype(np.uint8)).to(dev))
    x = torch.from_numpy(np.asarray(res["x"], dtype=np.float64)).to(dev)

    def run():
        f = X.output_field(dp, x)
        lvl = X.choose_level(dp, f, x)
        f2, lvl2 = X.keep_load_paths(dp, f, lvl, log=lambda *a: None)
        return lvl, lvl2
    ms, (lvl, lvl2) = _timed(run, 5)
    out = {"level": lvl, "level_after_guard": lvl2, "device_ms": round(ms, 3)}
    if cpu:
        from oracle import voxel_ref as V
        xs = np.asarray(res["x"], dtype=np.float64)
        vp = V.Prob(g, [V.Anchor(a.name, a.role, V.Box(a.region.lo, a.region.hi), loads=a.loads) for a in anchors],
                    fill, active=active)
        t = time.perf_counter()
        f = V.output_field(vp, xs)
        lr = V.choose_level(vp, f, xs)
        V.keep_load_paths(vp, f, lr, log=lambda *a: None)
        out["cpu_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        out["same_level"] = lr == lvl
    print(f"probe: export pipeline on the cantilever design {out}", file=sys.stderr, flush=True)
    return out


def voxel_probe(reps=5, cpu=True):
    """Device timings of the voxel passes around the SIMP loop at 192x96x96 (wall clock: every pass ends in a
    host read of its counts), with the oracle's CPU time on the same inputs."""
    import numpy as np
    import torch
    from anneal_amd import problem as P, export as X
    prob, fields = voxel_problem()
    n = prob.grid.nelem
    dev = torch.device("cuda", 0)
    from types import SimpleNamespace
    host_prob = prob
    # device-resident copies of the part's fields: the timings exclude host-to-device copies
    prob = SimpleNamespace(**vars(host_prob))
    for k in ("fill", "solid_fill", "void_fill"):
        setattr(prob, k, torch.from_numpy(getattr(host_prob, k)).to(dev))
    for k in ("active", "frozen"):
        setattr(prob, k, torch.from_numpy(getattr(host_prob, k).astype(np.uint8)).to(dev))
    mask_d = torch.from_numpy(fields["mask"].astype(np.uint8)).to(dev)
    x_d = torch.from_numpy(fields["x"]).to(dev)
    f_d = X.output_field(prob, x_d)
    res = {}
    res["boundary"], _ = _timed(lambda: P.boundary_device(prob, prob.active), reps)
    res["components"], (lab, ncomp) = _timed(lambda: P.components(prob.grid, mask_d), reps)
    res["connected"], _ = _timed(lambda: P.connected(prob, mask_d), reps)
    res["output_field"], _ = _timed(lambda: X.output_field(prob, x_d), reps)
    res["choose_level"], level = _timed(lambda: X.choose_level(prob, f_d, x_d), reps)
    f_low = torch.from_numpy(fields["f_low"]).to(dev)
    res["keep_load_paths"], _ = _timed(lambda: X.keep_load_paths(prob, f_low, 0.5, log=lambda *a: None), 1)
    _, _, sweeps, _ = X.path_arrays(prob, f_low, 0.44)
    # algorithmic bytes per voxel: choose_level reads x, fill (sum) then f in 4 passes; output_field 6 inputs + 1 out
    gbs = {"choose_level": 48.0 * n / (res["choose_level"] * 1e-3) / 1e9,
           "output_field": 42.0 * n / (res["output_field"] * 1e-3) / 1e9,
           "components": 5.0 * n / (res["components"] * 1e-3) / 1e9}
    out = {"probe": "voxel passes", "grid": list(prob.grid.shape), "components": ncomp, "level": level,
           "sssp_sweeps": sweeps, "device_ms": {k: round(v, 3) for k, v in res.items()},
           "device_GBps": {k: round(v, 1) for k, v in gbs.items()}}
    design = export_pipeline(dev, cpu=cpu)
    out["cantilever_design"] = design
    if cpu:
        c = voxel_cpu_baseline(host_prob, fields)
        out["cpu_ms"] = {k: round(v, 1) for k, v in c.items()}
        out["cpu_cores"] = 1
    for k, v in res.items():
        print(f"probe: {k} {v:.3f} ms", file=sys.stderr, flush=True)
    return out


def measured_traffic(kernel, shape):
    """HBM bytes per launch of `kernel` (FETCH_SIZE and WRITE_SIZE counters, copy-calibrated) from the committed
    counter summary profiles/apply_192x96x96_traffic.json, or None."""
    if tuple(shape) != SHAPE:
        return None
    path = os.path.join(ROOT, "profiles", "apply_192x96x96_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def apply_roofline(fem, reps=20, nbuf=8):
    """Average duration of one FP32 K·u apply over rotating (u, y) buffers."""
    import torch
    from anneal_amd import _lib
    g = fem.grid
    tile = TILE_SHAPES[fem.tile()] + ",kApply"
    bufs = [(torch.randn(1, 3, fem.ldn, device="cuda", dtype=torch.float32),
             torch.empty(1, 3, fem.ldn, device="cuda", dtype=torch.float32)) for _ in range(nbuf)]
    for u, y in bufs:           # warm
        _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(u), _lib.ptr(y), 1)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    n = 0
    for _ in range(reps):
        for u, y in bufs:
            _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(u), _lib.ptr(y), 1)
            n += 1
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    nbytes = 24 * g.nnode + 4 * g.nelem          # u read, y write (12 B/node each), Ee (4 B/elem)
    achieved = nbytes / (ms * 1e-3) / 1e9
    kernel = f"k_apply<float,{tile}>"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel, (g.nx, g.ny, g.nz)),
            "kernel": kernel, "bytes_per_launch": nbytes, "us_per_launch": round(ms * 1e3, 2),
            "dof_applies_per_s": 3 * g.nnode / (ms * 1e-3)}


def chebyshev_roofline(fem, reps=20, nbuf=8):
    """Average duration of one fine-level Chebyshev smoothing step (the SIMP step's dominant kernel) over
    rotating (u, b, y) buffer sets, y read as the previous iterate and overwritten in place."""
    import torch
    from anneal_amd import _lib
    g = fem.grid
    tile = TILE_SHAPES[fem.tile()] + ",kCheb"
    sets = [tuple(torch.randn(1, 3, fem.ldn, device="cuda", dtype=torch.float32) for _ in range(3))
            for _ in range(nbuf)]
    for u, b, y in sets:         # warm
        _lib.call("anneal_fem_smooth_step_f32", fem._ctx, _lib.ptr(u), _lib.ptr(b), _lib.ptr(y), 1)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    n = 0
    for _ in range(reps):
        for u, b, y in sets:
            _lib.call("anneal_fem_smooth_step_f32", fem._ctx, _lib.ptr(u), _lib.ptr(b), _lib.ptr(y), 1)
            n += 1
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    # u, b, previous iterate y read (12 B/node each), D^-1 (4 B/node), free-dof mask (1 B/node), y written
    # (12 B/node); Ee (4 B/elem)
    nbytes = 53 * g.nnode + 4 * g.nelem
    achieved = nbytes / (ms * 1e-3) / 1e9
    kernel = f"k_apply<float,{tile}>"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel, (g.nx, g.ny, g.nz)),
            "kernel": kernel, "bytes_per_launch": nbytes, "us_per_launch": round(ms * 1e3, 2),
            "bytes_formula": "53 B/node (u, b, y in; D^-1 4; mask 1; y out) + 4 B/element (Ee)",
            "share_of_simp_kernel_time": "13 % (the largest kernel of the step; 3 launches per PCG iteration)"}


def learned_leg(g, fill, active, frozen, fixed, loads, weights, iters, steps, plain_ms, runs=200, train_steps=2000,
                mode=1, log=None):
    """Train the U-Net on the GPU (pairs from seeded 48x24x24 cantilever SIMP runs) and time the same SIMP
    schedule with the network as warm start (mode 1) or two-stage preconditioner (mode 2)."""
    import torch
    from anneal_amd import learned as Lm
    from anneal_amd.optimize import SimpRun
    t0 = time.perf_counter()
    X, T, dims = Lm.generate_pairs(runs, 50, seed=0)
    Xv, Tv, _ = Lm.generate_pairs(10, 50, seed=10_000)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = Lm.UNet(seed=0)
    t1 = time.perf_counter()
    Lm.train(net, X, T, dims, train_steps, batch=32, lr=2e-3, seed=0)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    val = Lm.evaluate(net, Xv, Tv, dims)
    del X, T
    if log:
        log(f"bench: learned: {Xv.shape[0]} validation pairs, relative squared error {val:.3f}; data {t_data:.1f} s, "
            f"training {t_train:.1f} s")
    run = SimpRun(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.3, iters=iters, r_mm=3.0, tol=0.0,
                  emin=1e-9, rtol=1e-6, log=lambda m: None, learned=net, learned_mode=mode)
    torch.cuda.synchronize()
    settle_host()
    t2 = time.perf_counter()
    for _ in range(steps):
        run.step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t2) / steps
    cg = run.fem.cg_iters
    return {"value": round(1e3 / ms, 4), "unit": "SIMP iter/s", "ms_per_step": round(ms, 3),
            "speedup_vs_plain": round(plain_ms / ms, 4), "mode": "warm start" if mode == 1 else "two-stage FPCG",
            "pcg_iters_total": int(sum(cg)), "final_compliance": run.hist[-1]["compliance"],
            "training": {"pairs": int(runs * 50), "grid": "48x24x24 cantilevers (network on 25x13x13 nodes)",
                         "steps": train_steps, "batch": 32, "data_s": round(t_data, 2), "train_s": round(t_train, 2),
                         "validation_rel_sq_error": round(val, 4)}}


def _host_cpu():
    """Model name of the host CPU (from /proc/cpuinfo) and the number of cores the OS reports."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _oracle_iterations(shape, iters, r_mm):
    """Setup seconds and per-iteration seconds of the FP64 oracle SIMP on the cantilever at `shape` (one thread)."""
    from threadpoolctl import threadpool_limits
    from oracle.optimize_ref import cantilever as cant_ref, simp_ref
    g, fill, active, frozen, fixed, F = cant_ref(shape)
    with threadpool_limits(limits=1):
        t0 = time.time()
        r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=iters, r_mm=r_mm,
                     tol=0.0, emin=1e-9, rtol=1e-6, log=lambda *a: None)
        wall = time.time() - t0
    secs = [h["seconds"] for h in r["history"]]
    return wall - sum(secs), secs


def cpu_baseline():
    """The FP64 oracle's SIMP iterations on one host core: measured at 96x48x48 (iterations 1-2, setup apart)
    with the benchmark's physics (rmin 3 voxels, Emin 1e-9, rtol 1e-6), and at 48x24x24 (iterations 1-4) for
    the growth of the cost with size; the 96x48x48 rate is scaled by the element count (x8) to 192x96x96, whose
    CSR assembly (about 40-50 GB, minutes of setup) does not fit a bench run."""
    import numpy as np
    setup48, s48 = _oracle_iterations((48, 24, 24), 4, 3.0)
    setup96, s96 = _oracle_iterations((96, 48, 48), 2, 3.0)
    it48, it96 = float(np.mean(s48)), float(np.mean(s96))
    exponent = float(np.log(it96 / it48) / np.log(8.0))    # cost ~ elements^exponent between the two sizes
    scale = (SHAPE[0] * SHAPE[1] * SHAPE[2]) / (96 * 48 * 48)
    model, ncpu = _host_cpu()
    return {"value": 1.0 / (it96 * scale), "unit": "SIMP iter/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (CSR refill + scipy CG with its own Galerkin V-cycle, scipy.ndimage filter, "
                      f"36-step OC), one thread, on the benchmark cantilever physics (rmin 3 voxels): 96x48x48 "
                      f"iterations 1-2 measured at {it96:.2f} s/iter (setup {setup96:.1f} s apart), scaled x{scale:.0f} "
                      f"by element count to 192x96x96; 48x24x24 iterations 1-4 at {it48:.2f} s/iter, so the cost grows "
                      f"as elements^{exponent:.2f} between the measured sizes",
            "s_per_iter_96x48x48": round(it96, 3), "s_per_iter_48x24x24": round(it48, 3),
            "setup_s_96x48x48": round(setup96, 2), "size_exponent": round(exponent, 3),
            "host_cpu": model, "host_cpu_count": ncpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed SIMP iterations (default: the whole schedule)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--learned-mode", type=int, default=1, choices=[1, 2])
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--tiles", default=None,
                    help="comma-separated FP32 apply tile shape ids per multigrid level, finest first, instead of "
                         "the autotuner's choice (a profiler's per-dispatch overhead hides the differences it times)")
    ap.add_argument("--probe", choices=["apply-traffic", "unet", "voxel"], default=None,
                    help="apply-traffic: FP32 applies on rotating buffers at 192x96x96 plus four 256 MB copies "
                         "(the workload of the FETCH_SIZE / WRITE_SIZE counter runs); unet: U-Net forward and "
                         "forward+backward timings at the learned stage's grid sizes; voxel: the boundary, "
                         "connectivity, iso-level and load-path passes at 192x96x96 next to the oracle's CPU time "
                         "(unet and voxel print one JSON line)")
    ap.add_argument("--workload", choices=["cantilever", "large", "batch", "cantilever128"], default="cantilever",
                    help="cantilever: 192x96x96, 100 iterations (the headline); cantilever128: the same at "
                         "128x64x64 (the learned-preconditioner size); large: 384x192x192 with keep-outs, "
                         "keep-solid slabs and 3 load cases, 60 iterations; batch: 16 parts x 4 load cases")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this path runs on a single MI355X (no multi-GPU variant is built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    from anneal_amd.optimize import SimpRun
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.probe == "apply-traffic":
        traffic_probe()
        return
    if args.probe == "unet":
        print(json.dumps(unet_probe()), flush=True)
        return
    if args.probe == "voxel":
        print(json.dumps(voxel_probe(cpu=not args.no_cpu)), flush=True)
        return
    if args.workload == "batch":
        steps = 20 if args.steps is None else args.steps
        res = batch_bench(steps, steps, log)
        out = {"metric": "problem-iterations/s, 16 uneven parts x 4 load cases, lock-step batch (plain MG-PCG @1e-6)",
               "value": res["value"], "unit": res["unit"], "n_gpus": 1, "steps": steps, "warmup": BENCH_WARM,
               "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32 operator/smoother, f64 dots/refinement/density", "data": "synthetic (no dataset)",
               "config": {"workload": f"batch seed {BATCH_SEED}: 16 parts from {list(BATCH_SHAPES)}, seeded clamp "
                                      f"faces, 4 point-load cases each, vf U(0.2,0.4), rmin 1.5, Emin 1e-9",
                          **{k: v for k, v in res.items() if k not in ("value", "unit", "ms_per_step")}},
               "roofline": None, "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return

    if args.workload in ("cantilever", "cantilever128"):
        shape, iters = (SHAPE if args.workload == "cantilever" else (128, 64, 64)), 100
        g, fill, active, frozen, fixed, F = cantilever(shape)
        loads, weights = {"c": F}, {"c": 1.0}
        workload = (f"cantilever {shape[0]}x{shape[1]}x{shape[2]} hex8, rmin 3, Emin 1e-9, MG-PCG rtol 1e-6, "
                    f"100-it schedule")
    else:
        shape, iters = LARGE_SHAPE, 60
        g, fill, active, frozen, fixed, loads, weights = large_part()
        workload = (f"large part 384x192x192 hex8, seed {LARGE_SEED}: 5% keep-out boxes, 5% keep-solid slabs, "
                    f"3 load cases (sum), rmin 3, Emin 1e-9, MG-PCG rtol 1e-6, 60-it schedule")
    steps = iters if args.steps is None else args.steps
    log = lambda m: print(m, file=sys.stderr, flush=True)

    def new_run(timers=None):
        r = SimpRun(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.3, iters=iters, r_mm=3.0,
                    tol=0.0, emin=1e-9, rtol=1e-6, log=log, timers=timers)
        if args.tiles:
            r.fem.set_level_tiles([int(t) for t in args.tiles.split(",")])
        return r
    print(f"bench: warm-up run on the {shape} {args.workload}", file=sys.stderr, flush=True)
    warm = new_run()
    for _ in range(args.warmup):
        warm.step()
    del warm
    torch.cuda.synchronize()
    t_setup = time.perf_counter()
    run = new_run()
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    settle_host()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(ahead=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / steps
    cg_timed = int(sum(run.fem.cg_iters))
    # the rest of the schedule, so that the whole continuation (p ramp, beta doubling) is measured too
    t1 = time.perf_counter()
    while run.it < iters:
        run.step(ahead=True)
    torch.cuda.synchronize()
    full_s = dt + time.perf_counter() - t1
    cg = run.fem.cg_iters
    h = run.hist
    print(f"bench: {steps} steps in {dt:.3f} s ({ms:.2f} ms/iter), setup {setup_s:.2f} s; whole {iters}-iteration "
          f"schedule {1e3 * full_s / iters:.2f} ms/iter, PCG iterations total {sum(cg)} (min {min(cg)}, max {max(cg)}), "
          f"final compliance {h[-1]['compliance']}", file=sys.stderr, flush=True)
    roof_ku = apply_roofline(run.fem, reps=args.apply_reps)
    print(f"bench: apply {roof_ku['us_per_launch']} us, {roof_ku['achieved']} GB/s", file=sys.stderr, flush=True)
    roof = chebyshev_roofline(run.fem, reps=args.apply_reps)
    print(f"bench: Chebyshev step {roof['us_per_launch']} us, {roof['achieved']} GB/s", file=sys.stderr, flush=True)
    lrn = None
    if not args.no_learned and args.workload in ("cantilever", "cantilever128"):
        lrn = learned_leg(g, fill, active, frozen, fixed, loads, weights, iters, steps, ms, mode=args.learned_mode,
                          log=log)
        print(f"bench: learned {lrn['ms_per_step']} ms/iter ({lrn['speedup_vs_plain']}x plain), PCG iterations "
              f"{lrn['pcg_iters_total']}", file=sys.stderr, flush=True)
    cpu = None if (args.no_cpu or args.workload != "cantilever") else cpu_baseline()
    sx, sy, sz = shape
    out = {
        "metric": f"SIMP iter/s at {sx}x{sy}x{sz} (plain MG-PCG @1e-6 tol); K\u00b7u HBM GB/s",
        "value": round(1e3 / ms, 4), "unit": "SIMP iter/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 operator/smoother, f64 dots/refinement/density",
        "data": "synthetic (no dataset)",
        "config": {"workload": workload, "solver": "plain MG-PCG", "iterations_timed": f"1-{steps}",
                   "setup_s": round(setup_s, 3), "pcg_iters_timed": cg_timed,
                   "full_schedule": {"iterations": iters, "iter_per_s": round(iters / full_s, 4),
                                     "ms_per_iter": round(1e3 * full_s / iters, 3)},
                   "pcg_iters_total": int(sum(cg)), "pcg_iters_max": int(max(cg)),
                   "level_tiles": {"ids": [lv[3] for lv in run.fem.level_info()[:-1]],
                                   "shapes": [TILE_SHAPES[lv[3]] for lv in run.fem.level_info()[:-1]]},
                   "oc_volume_evals": int(run.oc_evals), "oc_trials": int(sum(x["ocSteps"] for x in h)),
                   "final_compliance": h[-1]["compliance"], "final_volume_fraction": h[-1]["volumeFraction"],
                   "learned": lrn},
        "roofline": roof, "roofline_ku": roof_ku, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
