"""Lock-step batch of independent problems (anneal_batch_*) against solving each problem alone."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _clamped(shape):
    from anneal_amd.fem import Grid
    g = Grid(shape, 1.0, (0, 0, 0))
    n = np.arange(g.nnode)
    clamp = n[n % (shape[0] + 1) == 0]
    return g, (3 * clamp[:, None] + np.arange(3)).ravel()


def test_batch_solve_matches_individual_solves():
    from anneal_amd.batch import BatchSolve
    from anneal_amd.fem import FEM
    rng = np.random.default_rng(0)
    specs = [((16, 8, 8), 2), ((12, 12, 6), 1), ((20, 10, 10), 3), ((8, 8, 8), 4)]
    fems, Ks, Fs, Us, Fred = [], [], [], [], []
    for shape, nr in specs:
        g, fixed = _clamped(shape)
        fem = FEM(g, np.ones(g.nelem, bool), fixed, 0.3, nrhs_max=4)
        K = fem.assemble(rng.uniform(0.01, 1.0, g.nelem))
        fem._activate(K)
        fem.rebuild()
        F = [fem.reduce_load(rng.normal(size=3 * g.nnode)) for _ in range(nr)]
        fems.append(fem)
        Ks.append(K)
        Fred.append(F)
        Fs.append(fem._to_device(F))
        Us.append(torch.zeros_like(Fs[-1]))
    print("batch solve of", [s for s, _ in specs], flush=True)
    b = BatchSolve(fems)
    its = b.solve(Fs, Us, [nr for _, nr in specs], 1e-9, 2000)
    print("batch iterations", its, flush=True)
    for fem, K, F, U, (shape, nr) in zip(fems, Ks, Fred, Us, specs):
        ub = fem._from_device(U, nr)
        ui = fem.solve(K, F, rtol=1e-9)
        for r in range(nr):
            assert np.linalg.norm(F[r] - K @ ub[r]) <= 1e-9 * np.linalg.norm(F[r]) * 1.0001
            assert np.linalg.norm(ub[r] - ui[r]) <= 1e-7 * np.linalg.norm(ui[r])
    # a second solve of the same batch replays the recorded graph; warm start from the solution
    its2 = b.solve(Fs, Us, [nr for _, nr in specs], 1e-9, 2000, use_x0=True)
    assert all(i == 0 for p in its2 for i in p)
    b.close()
    for f in fems:
        f.close()


def test_batch_graphs_match_direct_launches():
    """The batch with each member's iteration replayed from recorded graphs (graphs=1) gives the same bits
    as the same kernels launched one by one (the default), solve after solve."""
    from anneal_amd.batch import BatchSolve
    from anneal_amd.fem import FEM
    rng = np.random.default_rng(7)
    specs = [((16, 8, 8), 2), ((20, 10, 10), 3), ((8, 8, 8), 1)]
    fems, Fs = [], []
    for shape, nr in specs:
        g, fixed = _clamped(shape)
        fem = FEM(g, np.ones(g.nelem, bool), fixed, 0.3, nrhs_max=4)
        fem._activate(fem.assemble(rng.uniform(0.01, 1.0, g.nelem)))
        fem.rebuild()
        fems.append(fem)
        Fs.append(fem._to_device([fem.reduce_load(rng.normal(size=3 * g.nnode)) for _ in range(nr)]))
    nrs = [nr for _, nr in specs]
    b = BatchSolve(fems)
    results = {}
    for graphs in (0, 1, 0, 1):
        for f in fems:
            f.set_option("graphs", graphs)
        print("batch solve with graphs =", graphs, flush=True)
        Us = [torch.zeros_like(F) for F in Fs]
        its = b.solve(Fs, Us, nrs, 1e-9, 2000)
        torch.cuda.synchronize()
        got = (its, [U.cpu() for U in Us])
        if graphs in results:
            assert got[0] == results[graphs][0]
            assert all(torch.equal(x, y) for x, y in zip(got[1], results[graphs][1]))
        results[graphs] = got
    assert results[0][0] == results[1][0]
    assert all(torch.equal(x, y) for x, y in zip(results[0][1], results[1][1]))
    b.close()
    for f in fems:
        f.close()


def test_batch_solve_with_capped_inner_solves_and_iteration_budget():
    """Members whose inner PCG is capped at 1 and 2 iterations take many refinement passes in the batch
    (each problem's next iteration is launched before the host reads the current one, so a cap must stop
    the launches exactly); the batch still meets the tolerance, agrees with the members' own solves, and
    a budget below what the solves need ends them with the not-converged error."""
    from anneal_amd.batch import BatchSolve
    from anneal_amd.fem import FEM, SolverError
    rng = np.random.default_rng(5)
    specs = [((16, 8, 8), 2), ((12, 12, 6), 1), ((20, 10, 10), 3)]
    fems, Ks, Fs, Fred = [], [], [], []
    for shape, nr in specs:
        g, fixed = _clamped(shape)
        fem = FEM(g, np.ones(g.nelem, bool), fixed, 0.3, nrhs_max=4)
        K = fem.assemble(rng.uniform(0.01, 1.0, g.nelem))
        fem._activate(K)
        fem.rebuild()
        F = [fem.reduce_load(rng.normal(size=3 * g.nnode)) for _ in range(nr)]
        fems.append(fem)
        Ks.append(K)
        Fred.append(F)
        Fs.append(fem._to_device(F))
    fems[0].set_option("inner_max", 2)
    fems[2].set_option("inner_max", 1)
    nrs = [nr for _, nr in specs]
    print("batch solve with inner caps 2, -, 1 of", [s for s, _ in specs], flush=True)
    b = BatchSolve(fems)
    Us = [torch.zeros_like(F) for F in Fs]
    its = b.solve(Fs, Us, nrs, 1e-9, 2000)
    print("batch iterations", its, flush=True)
    for fem, K, F, U, nr in zip(fems, Ks, Fred, Us, nrs):
        ub = fem._from_device(U, nr)
        ui = fem.solve(K, F, rtol=1e-9)
        for r in range(nr):
            assert np.linalg.norm(F[r] - K @ ub[r]) <= 1e-9 * np.linalg.norm(F[r]) * 1.0001
            assert np.linalg.norm(ub[r] - ui[r]) <= 1e-7 * np.linalg.norm(ui[r])
    print("batch solve with a budget of 3 PCG iterations", flush=True)
    Us = [torch.zeros_like(F) for F in Fs]
    with pytest.raises(SolverError):
        b.solve(Fs, Us, nrs, 1e-9, 3)
    b.close()
    for f in fems:
        f.close()


def test_batch_simp_matches_individual_runs():
    import bench
    from anneal_amd.batch import BatchSimp
    from anneal_amd.optimize import SimpRun
    probs = bench.batch_problems(seed=11, n=3, ncase=2, iters=8, shapes=((16, 8, 8), (12, 12, 12), (20, 10, 8)))
    for p in probs:
        p["rtol"] = 1e-10
    print("batch SIMP of", [p["grid"].shape for p in probs], flush=True)
    b = BatchSimp(probs)
    for _ in range(8):
        b.step()
    keys = ("grid", "fill", "active", "frozen", "fixed", "loads", "weights", "E0", "nu", "vf")
    for p, rb in zip(probs, b.runs):
        ri = SimpRun(*[p[k] for k in keys], **{k: v for k, v in p.items() if k not in keys}, log=lambda m: None)
        for _ in range(8):
            ri.step()
        for hb, hi in zip(rb.hist, ri.hist):
            assert np.allclose(hb["compliance"], hi["compliance"], rtol=1e-7, atol=0)
            assert abs(hb["volumeFraction"] - hi["volumeFraction"]) < 1e-9
        assert np.abs(rb.result()["x"] - ri.result()["x"]).max() < 1e-6
        ri.fem.close()
    b.close()


def _group_problem(shape, seed, nr):
    """A problem of a group: its own clamp face, random voids and stiffness contrast, nr load cases."""
    from anneal_amd.fem import FEM, Grid
    rng = np.random.default_rng(seed)
    g = Grid(shape, 1.0, (0, 0, 0))
    active = rng.random(g.nelem) >= 0.1
    n = np.arange(g.nnode)
    axis = seed % 3
    coord = (n % (shape[0] + 1), (n // (shape[0] + 1)) % (shape[1] + 1), n // ((shape[0] + 1) * (shape[1] + 1)))[axis]
    fixed = (3 * n[coord == 0][:, None] + np.arange(3)).ravel()
    fem = FEM(g, active, fixed, 0.3, nrhs_max=4)
    Ee = np.where(rng.random(g.nelem) < 0.2, 1e-9, rng.uniform(0.05, 1.0, g.nelem))
    K = fem.assemble(Ee[fem.act])
    fem._activate(K)
    fem.rebuild()
    F = [fem.reduce_load(rng.normal(size=3 * g.nnode)) for _ in range(nr)]
    return fem, K, F


def test_group_solve_matches_individual_solves():
    """Five problems of one shape (different clamps, voids, stiffness) as one block of 15 right-hand sides."""
    from anneal_amd.batch import FemGroup
    shape, nr = (16, 8, 12), 3
    probs = [_group_problem(shape, s, nr) for s in range(5)]
    fems = [p[0] for p in probs]
    print("group solve:", len(fems), "x", shape, "x", nr, "cases", flush=True)
    grp = FemGroup(fems, len(fems) * nr)
    GF, GU = grp.buffers(nr)
    for j, (fem, K, F) in enumerate(probs):
        GF[j * nr:(j + 1) * nr].copy_(fem._to_device(F))
    its = grp.solve_device(GF, GU, len(fems) * nr, 1e-9, 2000, False)
    print("group iterations", its, flush=True)
    for j, (fem, K, F) in enumerate(probs):
        ub = fem._from_device(GU[j * nr:(j + 1) * nr].contiguous(), nr)
        ui = fem.solve(K, F, rtol=1e-9)
        for r in range(nr):
            assert np.linalg.norm(F[r] - K @ ub[r]) <= 1e-9 * np.linalg.norm(F[r]) * 1.0001
            assert np.linalg.norm(ub[r] - ui[r]) <= 1e-6 * np.linalg.norm(ui[r])
    # warm restart from the solution: nothing left to do
    its2 = grp.solve_device(GF, GU, len(fems) * nr, 1e-9, 2000, True)
    assert all(i == 0 for i in its2)
    grp.close()
    for f in fems:
        f.close()


def test_batch_with_groups_matches_ungrouped():
    """Groups and single problems in one lock-step batch give the solutions of the ungrouped batch."""
    from anneal_amd.batch import BatchSolve
    nr = 2
    specs = [((12, 8, 8), 10), ((12, 8, 8), 11), ((12, 8, 8), 12), ((10, 10, 6), 13), ((10, 10, 6), 14),
             ((14, 6, 6), 15)]
    probs = [_group_problem(shape, seed, nr) for shape, seed in specs]
    fems = [p[0] for p in probs]
    Fs = [p[0]._to_device(p[2]) for p in probs]
    out = {}
    for grouped in (True, False):
        Us = [torch.zeros_like(F) for F in Fs]
        b = BatchSolve(fems, group=grouped, group_rhs=nr)
        print("batch grouped" if grouped else "batch ungrouped", [k for k, _ in b.units], flush=True)
        if grouped:
            assert sorted(len(idx) for _, idx in b.groups) == [2, 3]
        b.solve(Fs, Us, [nr] * len(fems), 1e-9, 2000)
        out[grouped] = [fem._from_device(U, nr) for fem, U in zip(fems, Us)]
        b.close()
    for (fem, K, F), ug, uu in zip(probs, out[True], out[False]):
        for r in range(nr):
            assert np.linalg.norm(F[r] - K @ ug[r]) <= 1e-9 * np.linalg.norm(F[r]) * 1.0001
            assert np.linalg.norm(ug[r] - uu[r]) <= 1e-6 * np.linalg.norm(uu[r])
    for f in fems:
        f.close()


def test_batch_simp_of_one_shape_matches_individual_runs():
    """BatchSimp over three parts of one grid shape (solved as one group) against separate SimpRuns."""
    import bench
    from anneal_amd.batch import BatchSimp
    from anneal_amd.optimize import SimpRun
    probs = bench.batch_problems(seed=21, n=3, ncase=2, iters=6, shapes=((16, 8, 8),))
    for p in probs:
        p["rtol"] = 1e-10
    b = BatchSimp(probs)
    assert len(b.solver.groups) == 1
    for _ in range(6):
        b.step()
    keys = ("grid", "fill", "active", "frozen", "fixed", "loads", "weights", "E0", "nu", "vf")
    for p, rb in zip(probs, b.runs):
        ri = SimpRun(*[p[k] for k in keys], **{k: v for k, v in p.items() if k not in keys}, log=lambda m: None)
        for _ in range(6):
            ri.step()
        for hb, hi in zip(rb.hist, ri.hist):
            assert np.allclose(hb["compliance"], hi["compliance"], rtol=1e-7, atol=0)
            assert abs(hb["volumeFraction"] - hi["volumeFraction"]) < 1e-9
        assert np.abs(rb.result()["x"] - ri.result()["x"]).max() < 1e-6
        ri.fem.close()
    b.close()


def test_batch_simp_runs_stop_on_their_own():
    """Runs with their own iteration budgets and stopping tests leave the batch when simp() would stop them
    (optimize.py:105, 173-175): histories and designs equal those of separate runs."""
    import bench
    from anneal_amd.batch import BatchSimp
    from anneal_amd.optimize import SimpRun
    probs = bench.batch_problems(seed=5, n=3, ncase=2, iters=14, shapes=((12, 6, 6), (16, 8, 8)))
    for p in probs:
        p["rtol"] = 1e-10
    probs[0]["iters"] = 9                                    # a shorter budget
    probs[1].update(beta_max=1.0, tol=1.0)                   # converges at iteration 11 (p = 3 and beta = 1 from 9)
    b = BatchSimp(probs)
    res = b.run()
    assert [len(r["history"]) for r in res] == [9, 11, 14]
    assert not b.active()
    keys = ("grid", "fill", "active", "frozen", "fixed", "loads", "weights", "E0", "nu", "vf")
    for p, rb, want_len in zip(probs, res, (9, 11, 14)):
        ri = SimpRun(*[p[k] for k in keys], **{k: v for k, v in p.items() if k not in keys}, log=lambda m: None)
        hi = ri.run()
        assert len(hi["history"]) == want_len
        for hb, hs in zip(rb["history"], hi["history"]):
            assert np.allclose(hb["compliance"], hs["compliance"], rtol=1e-7, atol=0)
            assert hb["beta"] == hs["beta"] and hb["p"] == hs["p"]
        assert np.abs(rb["x"] - hi["x"]).max() < 1e-6
        ri.fem.close()
    b.close()
