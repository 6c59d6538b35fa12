"""Write the oracle's golden vectors used by the GPU parity tests.

Run from the repository root:  python tests/golden/make_golden.py [name]
The fixtures are produced by the FP64 CPU oracle (oracle/), itself pinned by
the closed-form, identity and finite-difference checks in tests/test_oracle*.py.
Nothing here imports or runs the reference.

Fixtures:
  cantilever_16x8x8_it30, cantilever_48x24x24_it50: the benchmark cantilever
      (single case, cold start, all active, fill 1, tol 0).
  part_worst_24x12x12: a part with a keep-out box (inactive voxels), a
      keep-solid block (frozen), a partial-fill top layer, partly fixed nodes
      and 3 load cases combined with the p-norm ("worst", optimize.py:135-139).
  part_warm_24x12x12: the mass-search pass that follows it, as run.py:204-209
      calls simp: x0 = the previous design, a lower volume fraction,
      max(15, 2*iters//5) iterations and the default tol 0.01.
  cantilever_tol_16x8x8: the early stop (optimize.py:173-175) with tol 0.04 (at 0.01
      this cantilever runs its 60 iterations: the move limit keeps the change above it).
  many_cases_12x6x6: 10 load cases combined by weighted sum, default Emin 1e-6.
Each fixture stores the inputs (masks, fixed dofs, loads) with the trajectory.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.fem_ref import Grid  # noqa: E402
from oracle.optimize_ref import cantilever, simp_ref  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def _save(name, hist, r, rec, inputs, extra):
    np.savez_compressed(
        os.path.join(HERE, name),
        compliance=np.array([h["compliance"] for h in hist]),
        objective=np.array([h["objective"] for h in hist]),
        volume=np.array([h["volumeFraction"] for h in hist]),
        change=np.array([h["change"] for h in hist]),
        grey=np.array([h["grey"] for h in hist]),
        p=np.array([h["p"] for h in hist]), beta=np.array([h["beta"] for h in hist]),
        oc_steps=np.array([h["ocSteps"] for h in hist]),
        x_final=r["x"],
        x_iter10=rec[9]["xPhys"] if len(rec) >= 10 else rec[-1]["xPhys"],
        **inputs, **extra)
    print("wrote", name, "iterations", len(hist), "final objective", hist[-1]["objective"])


def cantilever_trajectory(shape, iters, r_mm, rtol, name):
    g, fill, active, frozen, fixed, F = cantilever(shape)
    rec = []
    r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=iters, r_mm=r_mm,
                 tol=0.0, emin=1e-9, rtol=rtol, record=rec, log=lambda *a: None)
    hist = r["history"]
    np.savez_compressed(
        os.path.join(HERE, name),
        shape=np.array(shape), iters=np.array(iters), r_mm=np.array(r_mm), rtol=np.array(rtol),
        compliance=np.array([h["compliance"][0] for h in hist]),
        volume=np.array([h["volumeFraction"] for h in hist]),
        change=np.array([h["change"] for h in hist]),
        grey=np.array([h["grey"] for h in hist]),
        p=np.array([h["p"] for h in hist]), beta=np.array([h["beta"] for h in hist]),
        oc_steps=np.array([h["ocSteps"] for h in hist]),
        x_final=r["x"],
        x_iter10=rec[9]["xPhys"] if len(rec) >= 10 else rec[-1]["xPhys"])
    print("wrote", name, "final compliance", hist[-1]["compliance"][0])


def part_problem(shape, seed, ncase):
    """A part with the features simp() sees in production (optimize.py:50-77, problem.py:301-326).

    Keep-out box (inactive), keep-solid block at the loaded end (frozen),
    partial-fill top layer (fill 0.6), the x = 0 face clamped plus a roller
    (z only) along the bottom edge of the far end, and ``ncase`` load cases:
    a unit force of seeded direction (uniform on the sphere) spread evenly over
    a seeded 2x2 node patch on the x = nx face, each with its own weight.
    """
    rng = np.random.default_rng(seed)
    g = Grid(shape, 1.0, (0, 0, 0))
    nx, ny, nz = shape
    i, j, k = g.elem_ijk(np.arange(g.nelem))
    keep_out = (i >= nx // 3) & (i < nx // 3 + max(2, nx // 4)) & (j < ny // 2) & (k >= nz // 3) & (k < 2 * nz // 3)
    active = ~keep_out
    frozen = (i >= nx - max(1, nx // 6)) & (j >= ny // 3) & (j < 2 * ny // 3) & (k >= nz // 3) & (k < 2 * nz // 3)
    frozen &= active
    fill = np.where(k == nz - 1, 0.6, 1.0)
    n = np.arange(g.nnode)
    ni, nj, nk = n % (nx + 1), (n // (nx + 1)) % (ny + 1), n // ((nx + 1) * (ny + 1))
    clamp = n[ni == 0]
    roller = n[(ni == nx) & (nk == 0)]
    fixed = np.sort(np.concatenate([(3 * clamp[:, None] + np.arange(3)).ravel(), 3 * roller + 2]))
    loads, weights = {}, {}
    for c in range(ncase):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        j0 = rng.integers(ny // 3, 2 * ny // 3)
        k0 = rng.integers(nz // 3, 2 * nz // 3)
        patch = n[(ni == nx) & (nj >= j0) & (nj <= j0 + 1) & (nk >= k0) & (nk <= k0 + 1)]
        F = np.zeros(3 * g.nnode)
        for comp in range(3):
            np.add.at(F, 3 * patch + comp, d[comp] / patch.size)
        loads[f"case{c}"] = F
        weights[f"case{c}"] = float(rng.uniform(0.5, 1.5))
    return g, fill, active, frozen, fixed, loads, weights


def _inputs(shape, fill, active, frozen, fixed, loads, weights):
    names = list(loads)
    return dict(shape=np.array(shape), fill=fill, active=active, frozen=frozen, fixed=np.asarray(fixed),
                loads=np.stack([loads[c] for c in names]), weights=np.array([weights[c] for c in names]))


def part_trajectories():
    shape = (24, 12, 12)
    g, fill, active, frozen, fixed, loads, weights = part_problem(shape, 7, 3)
    inputs = _inputs(shape, fill, active, frozen, fixed, loads, weights)
    kw = dict(iters=30, r_mm=2.2, combine="worst", tol=0.0, emin=1e-9, rtol=1e-11)
    rec = []
    r = simp_ref(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.3, record=rec, log=lambda *a: None,
                 **kw)
    extra = dict(E0=1.0, nu=0.3, vf=0.3, iters=30, r_mm=2.2, combine="worst", tol=0.0, emin=1e-9)
    _save("part_worst_24x12x12.npz", r["history"], r, rec, inputs, extra)
    # the next mass-search pass (run.py:202-209): warm start, lower target, 2/5 of the budget, default tol
    x0 = r["x"]
    iters = max(15, 30 * 2 // 5)
    rec = []
    r2 = simp_ref(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.25, iters=iters, r_mm=2.2,
                  combine="worst", x0=x0, emin=1e-9, rtol=1e-11, record=rec, log=lambda *a: None)
    extra = dict(E0=1.0, nu=0.3, vf=0.25, iters=iters, r_mm=2.2, combine="worst", tol=0.01, emin=1e-9, x0=x0)
    _save("part_warm_24x12x12.npz", r2["history"], r2, rec, inputs, extra)


def cantilever_tol():
    shape = (16, 8, 8)
    g, fill, active, frozen, fixed, F = cantilever(shape)
    rec = []
    r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=60, r_mm=1.5, tol=0.04,
                 emin=1e-9, rtol=1e-11, record=rec, log=lambda *a: None)
    inputs = _inputs(shape, fill, active, frozen, fixed, {"c": F}, {"c": 1.0})
    extra = dict(E0=1.0, nu=0.3, vf=0.3, iters=60, r_mm=1.5, combine="sum", tol=0.04, emin=1e-9)
    _save("cantilever_tol_16x8x8.npz", r["history"], r, rec, inputs, extra)


def many_cases():
    shape = (12, 6, 6)
    g, fill, active, frozen, fixed, loads, weights = part_problem(shape, 11, 10)
    rec = []
    r = simp_ref(g, fill, active, frozen, fixed, loads, weights, 2.0, 0.3, 0.35, iters=20, r_mm=1.5, combine="sum",
                 tol=0.0, rtol=1e-11, record=rec, log=lambda *a: None)
    inputs = _inputs(shape, fill, active, frozen, fixed, loads, weights)
    extra = dict(E0=2.0, nu=0.3, vf=0.35, iters=20, r_mm=1.5, combine="sum", tol=0.0, emin=1e-6)
    _save("many_cases_12x6x6.npz", r["history"], r, rec, inputs, extra)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("all", "small"):
        cantilever_trajectory((16, 8, 8), 30, 1.5, 1e-11, "cantilever_16x8x8_it30.npz")
    if which in ("all", "c1"):
        cantilever_trajectory((48, 24, 24), 50, 1.5, 1e-11, "cantilever_48x24x24_it50.npz")
    if which in ("all", "part"):
        part_trajectories()
    if which in ("all", "tol"):
        cantilever_tol()
    if which in ("all", "cases"):
        many_cases()
