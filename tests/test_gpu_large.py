"""Large part (384x192x192, 14.2 M elements, 43 M dofs): keep-outs, keep-solid slabs, 3 load cases.

Properties checked after two SIMP iterations: every case's solve meets the
FP64 true-residual tolerance (||F - K u|| <= rtol ||F||, K applied in FP64 on
the device), the OC design meets the volume target, keep-out voxels stay 0 and
keep-solid voxels 1, and the compliances are finite and positive.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_large_part_two_iterations():
    import torch
    import bench
    from anneal_amd import _lib
    from anneal_amd.optimize import SimpRun
    print("large part: building 384x192x192 workload", flush=True)
    g, fill, active, frozen, fixed, loads, weights = bench.large_part()
    run = SimpRun(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.3, iters=60, r_mm=3.0, tol=0.0,
                  emin=1e-9, rtol=1e-6, log=lambda m: print(m, flush=True))
    for _ in range(2):
        h = run.step()
    fem = run.fem
    y = torch.empty_like(run.U)
    _lib.call("anneal_fem_apply_f64", fem._ctx, _lib.ptr(run.U), _lib.ptr(y), run.ncase)
    nn = g.nnode                    # the apply leaves the padding of y (nodes nnode .. ldn) unwritten
    res = (run.F - y)[:, :, :nn].reshape(run.ncase, -1).norm(dim=1)
    rel = (res / run.F.reshape(run.ncase, -1).norm(dim=1)).cpu().numpy()
    print("true relative residuals", rel, "cg", h["cgIters"], flush=True)
    assert (rel <= 1.0e-6).all()
    assert all(np.isfinite(c) and c > 0 for c in h["compliance"])
    assert abs(h["volumeFraction"] - 0.3) < 1e-3
    x = run.result()["x"]
    assert np.all(x[~active] == 0.0)
    assert np.all(x[frozen] == 1.0)
    assert np.all((x >= 0) & (x <= 1))
