"""GPU simp() against the oracle on the paths a production run takes.

run.py:202-209 calls simp() on parts with keep-outs (inactive voxels),
keep-solid regions (frozen), partial fill, several load cases combined by
weighted sum or by the p-norm ("worst", optimize.py:132-139), and warm-starts
every mass-search pass after the first from the previous design (x0,
optimize.py:72-77, 100-101) with the default early stop (optimize.py:173-175).
The oracle's FP64 trajectories of these runs are the golden fixtures written
by tests/golden/make_golden.py; the bar is the north-star one: compliance
(every case, every iteration) within 1e-6 relative, final densities within
1e-4 max-abs, the same continuation, the same OC trial counts and the same
stopping iteration.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _inputs(gold):
    shape = tuple(int(v) for v in gold["shape"])
    names = [f"case{c}" for c in range(gold["loads"].shape[0])] if gold["loads"].shape[0] > 1 else ["c"]
    loads = {n: gold["loads"][i] for i, n in enumerate(names)}
    weights = {n: float(gold["weights"][i]) for i, n in enumerate(names)}
    return shape, gold["fill"], gold["active"], gold["frozen"], gold["fixed"], loads, weights


def _check(r, gold, label):
    hist = r["history"]
    comp = np.array([h["compliance"] for h in hist])
    want = gold["compliance"].reshape(len(gold["p"]), -1)
    print(f"{label}: {len(hist)} iterations (oracle {len(want)})", flush=True)
    assert len(hist) == len(want), "stopped at a different iteration"
    rel = np.abs(comp - want) / np.abs(want)
    obj = np.array([h["objective"] for h in hist])
    rel_obj = np.abs(obj - gold["objective"]) / np.abs(gold["objective"])
    dx = np.abs(r["x"] - gold["x_final"]).max()
    vol = np.abs(np.array([h["volumeFraction"] for h in hist]) - gold["volume"]).max()
    print(f"  max rel compliance err {rel.max():.3e}, objective {rel_obj.max():.3e}, final density max-abs "
          f"{dx:.3e}, volume {vol:.3e}", flush=True)
    assert [h["p"] for h in hist] == gold["p"].tolist()
    assert [h["beta"] for h in hist] == gold["beta"].tolist()
    assert [h["ocSteps"] for h in hist] == gold["oc_steps"].tolist()
    assert rel.max() <= 1e-6
    assert rel_obj.max() <= 1e-6
    assert dx <= 1e-4
    assert vol <= 1e-6
    assert np.allclose([h["change"] for h in hist], gold["change"], atol=2e-4, rtol=0)


def _simp(gold, rtol=1e-10, **over):
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import simp
    shape, fill, active, frozen, fixed, loads, weights = _inputs(gold)
    kw = dict(iters=int(gold["iters"]), r_mm=float(gold["r_mm"]), combine=str(gold["combine"]),
              tol=float(gold["tol"]), emin=float(gold["emin"]), rtol=rtol, log=lambda *a: None)
    if "x0" in gold.files:
        kw["x0"] = gold["x0"]
    kw.update(over)
    return simp(Grid(shape, 1.0, (0, 0, 0)), fill, active, frozen, fixed, loads, weights, float(gold["E0"]),
                float(gold["nu"]), float(gold["vf"]), **kw)


@pytest.mark.parametrize("name", ["part_worst_24x12x12", "part_warm_24x12x12", "cantilever_tol_16x8x8",
                                  "many_cases_12x6x6"])
def test_simp_path_matches_oracle(name):
    """Keep-outs, keep-solids, fill 0.6, rollers, 3 cases 'worst' (part_worst); its warm-started next pass at a
    lower volume fraction (part_warm); the tol early stop (cantilever_tol); 10 cases by weighted sum at the
    default Emin (many_cases)."""
    path = os.path.join(GOLDEN, name + ".npz")
    gold = np.load(path)
    print(f"simp path {name}", flush=True)
    _check(_simp(gold), gold, name)


def test_more_load_cases_than_one_solve_block():
    """70 load cases (more than the 64 right-hand sides of one device solve) against the oracle."""
    from oracle.optimize_ref import simp_ref
    from oracle.fem_ref import Grid as GridRef
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import simp
    rng = np.random.default_rng(4)
    shape = (6, 3, 3)
    g = GridRef(shape, 1.0, (0, 0, 0))
    n = np.arange(g.nnode)
    clamp = n[n % (g.nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    far = n[n % (g.nx + 1) == g.nx]
    loads, weights = {}, {}
    for c in range(70):
        F = np.zeros(3 * g.nnode)
        F[3 * rng.choice(far) + rng.integers(3)] = rng.choice([-1.0, 1.0])
        loads[f"l{c}"] = F
        weights[f"l{c}"] = float(rng.uniform(0.5, 1.5))
    fill, active, frozen = np.ones(g.nelem), np.ones(g.nelem, bool), np.zeros(g.nelem, bool)
    kw = dict(iters=4, r_mm=1.5, tol=0.0, emin=1e-9, log=lambda *a: None)
    want = simp_ref(g, fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.4, rtol=1e-12, **kw)
    print("70 load cases on a 6x3x3 grid", flush=True)
    got = simp(Grid(shape, 1.0, (0, 0, 0)), fill, active, frozen, fixed, loads, weights, 1.0, 0.3, 0.4, rtol=1e-10,
               **kw)
    comp = np.array([h["compliance"] for h in got["history"]])
    ref = np.array([h["compliance"] for h in want["history"]])
    assert comp.shape == (4, 70)
    assert (np.abs(comp - ref) / np.abs(ref)).max() <= 1e-6
    assert np.abs(got["x"] - want["x"]).max() <= 1e-4


def test_operator_queued_ahead_changes_nothing():
    """run() queues each next iteration's operator before the previous iteration's last wait; the
    trajectory (every history value, the design) is bit-identical to stepping without it, through the
    continuation's beta changes, and stepping ahead leaves the operator of the next iteration."""
    from oracle.optimize_ref import cantilever
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import SimpRun
    shape = (24, 12, 12)
    g, fill, active, frozen, fixed, F = cantilever(shape)
    kw = dict(iters=30, r_mm=1.5, tol=0.0, emin=1e-9, rtol=1e-8, log=lambda *a: None)

    def new():
        return SimpRun(Grid(shape, 1.0, (0, 0, 0)), fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3,
                       0.3, **kw)
    print("cantilever 24x12x12, 30 iterations with and without the operator queued ahead", flush=True)
    a, b = new(), new()
    ra = a.run()
    for _ in range(30):
        b.step()
    rb = b.result()
    assert np.array_equal(ra["x"], rb["x"])
    keys = ("objective", "compliance", "p", "beta", "volumeFraction", "change", "grey", "cgIters", "ocSteps")
    for ha, hb in zip(ra["history"], rb["history"]):
        assert all(ha[k] == hb[k] for k in keys)
    assert {h["beta"] for h in ra["history"]} >= {0.0, 1.0, 2.0}
    a.fem.close()
    b.fem.close()
