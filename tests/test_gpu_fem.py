"""GPU parity of the matrix-free operator and the multigrid PCG solve against the FP64 oracle.

Tolerances: the FP64 apply must equal the assembled CSR product to 1e-12
relative; the FP32 apply to 2e-6 relative (single rounding of inputs and
stiffness scales); solves meet the requested rtol on the FP64 residual and
agree with the oracle's converged solution to the accuracy rtol implies.
"""
import numpy as np
import pytest

from gpu_util import aos_to_soa, oracle_fem, random_problem, soa_to_aos

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _product_fem(g, active, fixed):
    from anneal_amd.fem import FEM, Grid
    return FEM(Grid(g.shape, g.h, g.origin), active, fixed, 0.3)


@pytest.mark.parametrize("shape,inactive,fix,seed", [
    ((3, 2, 2), 0.0, "clamp", 0),
    ((7, 5, 4), 0.3, "random", 1),
    ((17, 16, 9), 0.2, "random", 2),
    ((33, 18, 17), 0.0, "clamp", 3),
    ((48, 24, 24), 0.1, "clamp", 4),
])
def test_apply_matches_csr(torch, shape, inactive, fix, seed):
    print(f"apply parity shape={shape} inactive={inactive} fix={fix} seed={seed}", flush=True)
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem(shape, seed, inactive, fix)
    ref = oracle_fem(g, active, fixed)
    K = ref.assemble(Ee[ref.act])
    fem = _product_fem(g, active, fixed)
    assert np.array_equal(fem.free, ref.free), "free dof set differs from fem.py:145-155"
    fem.assemble(Ee[fem.act])
    rng = np.random.default_rng(seed + 100)
    nrhs = 2
    us = [rng.standard_normal(ref.nfree) for _ in range(nrhs)]
    want = [K @ u for u in us]
    U = np.stack([aos_to_soa(ref.expand(u), g.nnode, fem.ldn) for u in us])
    U64 = torch.from_numpy(U).cuda()
    Y64 = torch.empty_like(U64)
    _lib.call("anneal_fem_apply_f64", fem._ctx, _lib.ptr(U64), _lib.ptr(Y64), nrhs)
    U32 = U64.float().contiguous()
    Y32 = torch.empty_like(U32)
    _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(U32), _lib.ptr(Y32), nrhs)
    torch.cuda.synchronize()
    for r in range(nrhs):
        y64 = soa_to_aos(Y64[r].cpu().numpy(), g.nnode)
        y32 = soa_to_aos(Y32[r].double().cpu().numpy(), g.nnode)
        scale = np.abs(want[r]).max()
        assert np.abs(y64[ref.free] - want[r]).max() <= 1e-12 * scale
        assert np.abs(y32[ref.free] - want[r]).max() <= 2e-6 * scale
        notfree = np.setdiff1d(np.arange(3 * g.nnode), ref.free)
        assert np.all(y64[notfree] == 0) and np.all(y32[notfree] == 0)


def test_apply_is_deterministic(torch):
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem((40, 20, 20), 7, 0.1)
    fem = _product_fem(g, active, fixed)
    fem.assemble(Ee[fem.act])
    U = torch.randn(1, 3, fem.ldn, dtype=torch.float32, device="cuda")
    a, b = torch.empty_like(U), torch.empty_like(U)
    _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(U), _lib.ptr(a), 1)
    _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(U), _lib.ptr(b), 1)
    torch.cuda.synchronize()
    nn = g.nnode                      # entries past nnode are padding the kernel never writes
    assert torch.equal(a[..., :nn], b[..., :nn])


@pytest.mark.parametrize("shape,inactive,seed,rtol", [
    ((12, 6, 6), 0.0, 0, 1e-10),
    ((20, 10, 9), 0.15, 1, 1e-8),
    ((48, 24, 24), 0.0, 2, 1e-8),
])
def test_solve_matches_oracle(torch, shape, inactive, seed, rtol):
    print(f"solve parity shape={shape} inactive={inactive} rtol={rtol}", flush=True)
    g, active, fixed, Ee = random_problem(shape, seed, inactive)
    ref = oracle_fem(g, active, fixed)
    K = ref.assemble(Ee[ref.act])
    rng = np.random.default_rng(seed)
    Fs = [rng.standard_normal(ref.nfree), rng.standard_normal(ref.nfree)]
    fem = _product_fem(g, active, fixed)
    Kd = fem.assemble(Ee[fem.act])
    us = fem.solve(Kd, Fs, rtol=rtol)
    assert len(fem.cg_iters) == 2
    for f, u in zip(Fs, us):
        res = np.linalg.norm(f - K @ u) / np.linalg.norm(f)
        assert res <= rtol * 1.0001, res
    uref = ref.solve(K, Fs, rtol=1e-12)
    for u, v in zip(us, uref):
        # energy-norm error is bounded by the residual: compare compliances
        c, cref = Fs[0] @ u, Fs[0] @ v
        assert abs(c - cref) <= 1e2 * rtol * abs(cref) + 1e-300


@pytest.mark.parametrize("shape,inactive", [((4, 2, 2), 0.0), ((8, 4, 4), 0.15)])
def test_solve_dense_bruteforce(torch, shape, inactive):
    """Single-level grids (135 and 675 dofs: one and three column blocks of the Gauss-Jordan
    inverse, the last pivot block partial): against numpy.linalg.solve of the assembled
    matrix.  The dense inverse is the whole preconditioner here, so PCG needs only the few
    iterations that FP32 vectors take to reach 1e-12; a wrong inverse needs many more."""
    g, active, fixed, Ee = random_problem(shape, 3, inactive, "clamp")
    ref = oracle_fem(g, active, fixed)
    A = ref.assemble(Ee[ref.act]).toarray()
    f = np.random.default_rng(1).standard_normal(ref.nfree)
    exact = np.linalg.solve(A, f)
    fem = _product_fem(g, active, fixed)
    assert fem.levels == 1
    u = fem.solve(fem.assemble(Ee[fem.act]), [f], rtol=1e-12)[0]
    print(f"dense {shape}: {fem.cg_iters[-1]} PCG iterations", flush=True)
    assert np.abs(u - exact).max() <= 1e-8 * np.abs(exact).max()
    assert fem.cg_iters[-1] <= 8


def test_warm_start_and_errors(torch):
    from anneal_amd.fem import SolverError
    g, active, fixed, Ee = random_problem((16, 8, 8), 5)
    fem = _product_fem(g, active, fixed)
    K = fem.assemble(Ee[fem.act])
    f = np.random.default_rng(2).standard_normal(fem.nfree)
    u = fem.solve(K, [f], rtol=1e-8)[0]
    n_cold = fem.cg_iters[-1]
    # a start that already meets the tolerance: the first pass (set up on the device before the host
    # has read the residual) ends at its first step's wait with no step counted, and u comes back as it went in
    u2 = fem.solve(K, [f], x0_list=[u], rtol=1e-8)[0]
    assert fem.cg_iters[-1] == 0 < n_cold
    assert np.array_equal(u2, u)
    with pytest.raises(SolverError, match="did not converge in 1 iterations"):
        fem.solve(K, [f], rtol=1e-12, maxiter=1)


def test_floating_structure_raises(torch):
    """No fixed dofs: the system is singular; the reference raises SolverError (fem.py:225-229)."""
    from anneal_amd.fem import SolverError
    g, active, _, Ee = random_problem((6, 3, 3), 6)
    fem = _product_fem(g, active, np.zeros(0, int))
    f = np.random.default_rng(3).standard_normal(fem.nfree)
    with pytest.raises(SolverError):
        fem.solve(fem.assemble(Ee[fem.act]), [f], rtol=1e-8, maxiter=200)


def test_element_energy(torch):
    g, active, fixed, Ee = random_problem((9, 7, 5), 8, 0.2)
    ref = oracle_fem(g, active, fixed)
    fem = _product_fem(g, active, fixed)
    u = np.random.default_rng(4).standard_normal(3 * g.nnode)
    want = ref.element_energy(u)
    got = fem.element_energy(u)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_graph_replay_matches_eager(torch):
    """The captured V-cycle graph gives bit-identical solves to eager launches, replay after replay."""
    g, active, fixed, Ee = random_problem((40, 20, 20), 9, 0.05)
    fem = _product_fem(g, active, fixed)
    K = fem.assemble(Ee[fem.act])
    rng = np.random.default_rng(5)
    fs = [rng.standard_normal(fem.nfree) for _ in range(3)]
    fem.set_option("graphs", 0)
    eager = [fem.solve(K, [f], rtol=1e-8)[0] for f in fs]
    fem.set_option("graphs", 1)
    for rep in range(2):
        for f, want in zip(fs, eager):
            got = fem.solve(K, [f], rtol=1e-8)[0]
            torch.cuda.synchronize()
            assert np.array_equal(got, want), f"replay {rep} differs"
    # a new stiffness field read through the same graph
    K2 = fem.assemble(Ee[fem.act] * rng.uniform(0.5, 2.0, fem.act.size))
    fem.set_option("graphs", 0)
    e2 = fem.solve(K2, [fs[0]], rtol=1e-8)[0]
    fem.set_option("graphs", 1)
    assert np.array_equal(fem.solve(K2, [fs[0]], rtol=1e-8)[0], e2)


def test_autotune_shapes_agree(torch):
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem((30, 17, 12), 10, 0.1)
    fem = _product_fem(g, active, fixed)
    fem.assemble(Ee[fem.act])
    U = torch.randn(1, 3, fem.ldn, dtype=torch.float32, device="cuda")
    outs = []
    for v in range(11):
        fem.autotune(v)
        Y = torch.zeros_like(U)
        _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(U), _lib.ptr(Y), 1)
        torch.cuda.synchronize()
        outs.append(Y[..., :g.nnode].clone())
    for o in outs[1:]:
        assert torch.allclose(o, outs[0], rtol=2e-6, atol=1e-6 * outs[0].abs().max().item())
    chosen, us = fem.autotune(-1)
    assert 0 <= chosen < 11 and len(us) == 11


def test_pinned_level_tiles(torch):
    """Tile shapes pinned per level survive the first operator and solve to the same answer."""
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem((40, 20, 20), 11, 0.05)
    fem = _product_fem(g, active, fixed)
    nlev = len(fem.level_info()) - 1  # the coarsest level is a dense solve
    pins = [7, 8, 10, 9][:nlev]
    fem.set_level_tiles(pins)
    K = fem.assemble(Ee[fem.act])
    assert [lv[3] for lv in fem.level_info()[:nlev]] == pins
    f = np.random.default_rng(3).standard_normal(fem.nfree)
    u1 = fem.solve(K, [f], rtol=1e-10)[0]
    fem.set_level_tiles([2] * nlev)
    assert [lv[3] for lv in fem.level_info()[:nlev]] == [2] * nlev
    u2 = fem.solve(K, [f], rtol=1e-10)[0]
    # both meet rtol 1e-10 on the FP64 residual: their compliances agree to ~1e-10
    assert abs(f @ u1 - f @ u2) <= 1e-8 * abs(f @ u1)
    with pytest.raises(_lib.AnnealLibError):
        fem.set_level_tiles([11])
    with pytest.raises(_lib.AnnealLibError):
        _lib.call("anneal_fem_set_level_tile", fem._ctx, nlev + 1, 0)


def test_slow_solves_raise_the_deep_smoothing_degree(torch):
    """A solve above deep_escalate_iters PCG iterations per rhs switches levels >= 2 to the escalation degree."""
    g, active, fixed, Ee = random_problem((96, 48, 48), 12, 0.05)
    a = _product_fem(g, active, fixed)
    b = _product_fem(g, active, fixed)
    nlev = len(a.level_info()) - 1
    assert nlev >= 3  # a smoothed level below level 1 exists
    for fem in (a, b):  # same tile shapes: same reduction partitions, comparable bits
        fem.set_level_tiles([2, 10, 8, 8][:nlev])
    a.set_option("deep_escalate_iters", 1)      # any solve escalates
    a.set_option("deep_escalate_degree", 7)
    b.set_option("smoother_degree_deep", 7)     # the escalated configuration from the start
    Ka, Kb = a.assemble(Ee[a.act]), b.assemble(Ee[b.act])
    f = np.random.default_rng(4).standard_normal(a.nfree)
    a.solve(Ka, [f], rtol=1e-8)                 # degree 3 on every level, then the switch
    ua = a.solve(Ka, [f], rtol=1e-8)[0]
    ub = b.solve(Kb, [f], rtol=1e-8)[0]
    torch.cuda.synchronize()
    assert a.cg_iters[-1] == b.cg_iters[-1]
    assert np.array_equal(ua, ub)
    # off by default for fast solves: a context left at the defaults keeps degree 3
    c = _product_fem(g, active, fixed)
    c.set_option("deep_escalate_iters", 0)
    uc = c.solve(c.assemble(Ee[c.act]), [f], rtol=1e-8)[0]
    assert abs(f @ uc - f @ ua) <= 1e-6 * abs(f @ ua)


@pytest.mark.parametrize("shape,inactive", [((40, 20, 20), 0.0), ((33, 17, 11), 0.2)])
def test_vcycle_is_symmetric_positive_definite(torch, shape, inactive):
    """The preconditioner must be SPD for PCG: <r1, M r2> = <r2, M r1> and <r, M r> > 0.

    Exercises restriction (P^T), prolongation (P), the Chebyshev smoother and the
    coarse inverse together: a restriction that is not the transpose of the
    prolongation breaks the symmetry at once.
    """
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem(shape, 11, inactive)
    fem = _product_fem(g, active, fixed)
    fem.assemble(Ee[fem.act])
    _lib.call("anneal_fem_rebuild", fem._ctx)
    mask = np.zeros((3, fem.ldn))
    mask[:, :g.nnode] = ((fem.dof_mask[None, :] >> np.arange(3)[:, None]) & 1)
    rng = np.random.default_rng(3)
    rs = [torch.from_numpy((rng.standard_normal((3, fem.ldn)) * mask)[None]).float().cuda() for _ in range(3)]
    zs = []
    for r in rs:
        z = torch.zeros_like(r)
        _lib.call("anneal_fem_vcycle", fem._ctx, _lib.ptr(r), _lib.ptr(z), 1)
        zs.append(z)
    torch.cuda.synchronize()
    d = lambda a, b: float((a.double() * b.double()).sum())
    for i in range(3):
        assert d(rs[i], zs[i]) > 0
        for j in range(i + 1, 3):
            a, b = d(rs[i], zs[j]), d(rs[j], zs[i])
            assert abs(a - b) <= 1e-5 * (abs(a) + abs(b)), (a, b)


@pytest.mark.parametrize("scalar_E", [True, False])
def test_stress_and_von_mises_match_oracle(scalar_E):
    """FEM.stress / von_mises on the device vs the FP64 restatement of fem.py:244-263."""
    from anneal_amd.fem import FEM
    g, active, fixed, Ee = random_problem((10, 6, 5), 21, inactive_frac=0.2)
    fem = FEM(g, active, fixed, 0.3)
    ref = oracle_fem(g, active, fixed)
    rng = np.random.default_rng(4)
    u = rng.normal(size=3 * g.nnode)
    E = 210e3 if scalar_E else rng.uniform(1.0, 3.0, fem.act.size)
    sig, sig_ref = fem.stress(u, E), ref.stress(u, E)
    vm, vm_ref = fem.von_mises(u, E), ref.von_mises(u, E)
    assert sig.shape == sig_ref.shape and vm.shape == vm_ref.shape
    assert np.abs(sig - sig_ref).max() <= 1e-12 * np.abs(sig_ref).max()
    assert np.abs(vm - vm_ref).max() <= 1e-12 * np.abs(vm_ref).max()
    d = rng.normal(size=3)
    assert np.allclose(FEM.normal_stress(sig, d), ref.normal_stress(sig_ref, d), rtol=1e-12, atol=0)
    fem.close()


def test_chebyshev_step_fixed_point(torch):
    """anneal_fem_smooth_step_f32 (the V-cycle's fine-level Chebyshev step, in place): with b = K u and the
    previous iterate equal to u, y = M (u + c1 (u - u) + c2 D^-1 (b - K u)) returns u on the free dofs."""
    from anneal_amd import _lib
    g, active, fixed, Ee = random_problem((40, 20, 16), 12, 0.1)
    fem = _product_fem(g, active, fixed)
    fem.assemble(Ee[fem.act])
    mask = torch.from_numpy(((fem.dof_mask[:, None] >> np.arange(3)) & 1).T.astype(np.float32)).cuda()
    U = torch.zeros(1, 3, fem.ldn, dtype=torch.float32, device="cuda")
    U[0, :, :g.nnode] = torch.randn(3, g.nnode, device="cuda") * mask
    B = torch.zeros_like(U)
    _lib.call("anneal_fem_apply_f32", fem._ctx, _lib.ptr(U), _lib.ptr(B), 1)
    Y = U.clone()
    _lib.call("anneal_fem_smooth_step_f32", fem._ctx, _lib.ptr(U), _lib.ptr(B), _lib.ptr(Y), 1)
    torch.cuda.synchronize()
    d = (Y - U)[0, :, :g.nnode].abs().max().item()
    assert d <= 1e-5 * U.abs().max().item(), d
