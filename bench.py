"""SIMP iterations per second at 192x96x96 on one MI355X, plus the K·u apply roofline.

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
  * roofline: the FP32 matrix-free K·u apply at the same grid, timed with HIP
    events on the stream it is launched on, over rotating buffers larger than
    the 256 MB MALL so the bytes come from HBM;
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


def apply_roofline(fem, reps=20, nbuf=8):
    """Average duration of one FP32 K·u apply over rotating (u, y) buffers."""
    import torch
    from anneal_amd import _lib
    g = fem.grid
    shapes = ["16,16,8", "16,16,16", "32,8,8", "32,8,16", "32,16,8", "16,16,4", "64,4,16", "16,16,2", "8,8,2",
              "8,8,4", "16,8,2"]
    tile = shapes[fem.tile()] + ",kApply"
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
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_apply<float,{tile}>", "bytes_per_launch": nbytes, "us_per_launch": round(ms * 1e3, 2),
            "dof_applies_per_s": 3 * g.nnode / (ms * 1e-3)}


def cpu_baseline(budget_s=20.0):
    """Oracle SIMP iterations at 48x24x24 on one host core, scaled by element count to 192x96x96."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle.optimize_ref import cantilever as cant_ref, simp_ref
    shape = (48, 24, 24)
    g, fill, active, frozen, fixed, F = cant_ref(shape)
    with threadpool_limits(limits=1):
        t0 = time.time()
        r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=4, r_mm=1.5,
                     tol=0.0, emin=1e-9, rtol=1e-6, log=lambda *a: None)
        wall = time.time() - t0
    secs = [h["seconds"] for h in r["history"][1:]]
    s_it = float(np.mean(secs))
    scale = (SHAPE[0] * SHAPE[1] * SHAPE[2]) / (shape[0] * shape[1] * shape[2])
    return {"value": 1.0 / (s_it * scale), "unit": "SIMP iter/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (CSR refill + scipy CG with its own Galerkin V-cycle, scipy.ndimage filter, "
                      f"36-step OC) on the same cantilever at 48x24x24, iterations 2-4 (mean {s_it:.2f} s/iter, "
                      f"{wall:.1f} s wall incl. setup), scaled x{scale:.0f} by element count to 192x96x96",
            "s_per_iter_48x24x24": s_it}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--apply-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this path runs on a single MI355X (no multi-GPU variant is built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    from anneal_amd.optimize import SimpRun

    g, fill, active, frozen, fixed, F = cantilever(SHAPE)
    log = lambda m: print(m, file=sys.stderr, flush=True)

    def new_run(timers=None):
        return SimpRun(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=100, r_mm=3.0,
                       tol=0.0, emin=1e-9, rtol=1e-6, log=log, timers=timers)
    print(f"bench: warm-up run on the {SHAPE} cantilever", file=sys.stderr, flush=True)
    warm = new_run()
    for _ in range(args.warmup):
        warm.step()
    del warm
    torch.cuda.synchronize()
    t_setup = time.perf_counter()
    run = new_run()
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps
    cg = run.fem.cg_iters
    h = run.hist
    print(f"bench: {args.steps} steps in {dt:.3f} s ({ms:.2f} ms/iter), setup {setup_s:.2f} s, "
          f"PCG iterations total {sum(cg)} (min {min(cg)}, max {max(cg)}), final compliance "
          f"{h[-1]['compliance'][0]:.6g}", file=sys.stderr, flush=True)
    roof = apply_roofline(run.fem, reps=args.apply_reps)
    print(f"bench: apply {roof['us_per_launch']} us, {roof['achieved']} GB/s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "SIMP iter/s at 192x96x96 (plain MG-PCG @1e-6 tol); K·u HBM GB/s",
        "value": round(1e3 / ms, 4), "unit": "SIMP iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 operator/smoother, f64 dots/refinement/density",
        "data": "synthetic cantilever (no dataset)",
        "config": {"workload": "cantilever 192x96x96 hex8, rmin 3, Emin 1e-9, MG-PCG rtol 1e-6, 100-it schedule",
                   "solver": "plain MG-PCG", "iterations_timed": f"1-{args.steps}", "setup_s": round(setup_s, 3),
                   "pcg_iters_total": int(sum(cg)), "pcg_iters_max": int(max(cg)),
                   "oc_volume_evals": int(run.oc_evals), "oc_trials": int(sum(x["ocSteps"] for x in h)),
                   "final_compliance": h[-1]["compliance"][0]},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
