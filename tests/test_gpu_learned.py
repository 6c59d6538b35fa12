"""Learned flexible-PCG stage: glue parity against the oracle composition, and convergence.

The stage is z = P N(R r): full-weighting restriction to multigrid level 1
(R = P^T, masked by the coarse dof mask), the U-Net on that node grid (input
channels: cbrt of the mean adjacent coarse Ee / e_ref, r scaled to unit RMS,
fixed flags, active flag), output scaled back by |R r| / sqrt(V) / e_ref and
masked, then trilinear prolongation masked by the fine dof mask.
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def clamp_problem(shape, seed):
    from anneal_amd.fem import Grid
    nx, ny, nz = shape
    g = Grid(shape, 1.0, (0, 0, 0))
    n = np.arange(g.nnode)
    clamp = n[n % (nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    rng = np.random.default_rng(seed)
    Ee = rng.uniform(0.05, 1.0, g.nelem)
    return g, fixed, Ee


def test_learned_stage_matches_oracle_composition():
    from anneal_amd.fem import FEM
    from anneal_amd.learned import UNet, CIN, COUT, WIDTHS, he_init
    shape = (16, 8, 8)
    nx, ny, nz = shape
    g, fixed, Ee = clamp_problem(shape, 3)
    fem = FEM(g, np.ones(g.nelem, bool), fixed, 0.3, nrhs_max=2)
    fem.assemble(Ee)
    params = he_init(seed=4, out_scale=1.0)
    net = UNet(params=params)
    e_ref = 0.7
    fem.set_learned(net, level=1, e_ref=e_ref, mode=2)
    rng = np.random.default_rng(5)
    ld = fem.ldn
    nn = g.nnode
    r = np.zeros((2, 3, ld), np.float32)
    r[:, :, :nn] = rng.normal(size=(2, 3, nn))
    rv = r[:, :, :nn]
    rv[:, :, np.arange(nn) % (nx + 1) == 0] = 0.0         # fixed dofs carry no residual
    rd = torch.tensor(r, device="cuda")
    zd = torch.zeros_like(rd)
    fem.learned_stage(rd, zd, 2)
    z = zd.cpu().numpy()

    # oracle composition (float64 transfers, bf16-emulating network)
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    fine_nodes = (nz + 1, ny + 1, nx + 1)
    coarse_nodes = (cz + 1, cy + 1, cx + 1)
    rt = torch.tensor(r[:, :, :nn].reshape(2, 3, *fine_nodes), dtype=torch.float64)
    b = U.restrict(rt).numpy()                             # [2][3][cz+1][cy+1][cx+1]
    cmask = np.ones(coarse_nodes, bool)
    cmask[:, :, 0] = False                                 # coarse nodes over the clamped face
    b[:, :, ~cmask] = 0.0
    Ef = Ee.astype(np.float32).reshape(nz, ny, nx)
    Ec = Ef.reshape(cz, 2, cy, 2, cx, 2).astype(np.float64).sum(axis=(1, 3, 5)) * 0.125
    Ecp = np.zeros((cz + 2, cy + 2, cx + 2))
    cnt = np.zeros_like(Ecp)
    Ecp[1:-1, 1:-1, 1:-1] = Ec
    cnt[1:-1, 1:-1, 1:-1] = 1.0
    S = sum(Ecp[a:a + cz + 1, bb:bb + cy + 1, c:c + cx + 1] for a in (0, 1) for bb in (0, 1) for c in (0, 1))
    C = sum(cnt[a:a + cz + 1, bb:bb + cy + 1, c:c + cx + 1] for a in (0, 1) for bb in (0, 1) for c in (0, 1))
    rho = np.cbrt(S / C / e_ref)
    V = (cx + 1) * (cy + 1) * (cz + 1)
    nrm2 = (b.reshape(2, -1) ** 2).sum(axis=1)
    x = np.zeros((2, V, CIN))
    for s in range(2):
        x[s, :, 0] = rho.ravel()
        x[s, :, 1:4] = (b[s].reshape(3, -1) * np.sqrt(V / nrm2[s])).T
        x[s, :, 4:7] = (~cmask).ravel()[:, None]
        x[s, :, 7] = 1.0
    xb = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).float()
    out = U.forward(U.unflatten(params, CIN, COUT, list(WIDTHS)), xb, 2, (cx + 1, cy + 1, cz + 1), list(WIDTHS),
                    bf16=True).detach().double().numpy()
    zc = np.zeros((2, 3) + coarse_nodes)
    for s in range(2):
        zc[s] = (out[s] * np.sqrt(nrm2[s] / V) / e_ref).T.reshape(3, *coarse_nodes)
    zc[:, :, ~cmask] = 0.0
    zf = U.prolong(torch.tensor(zc), fine_nodes).numpy()
    fmask = np.ones(fine_nodes, bool)
    fmask[:, :, 0] = False
    zf[:, :, ~fmask] = 0.0
    zg = z[:, :, :nn].reshape(2, 3, *fine_nodes)
    err = np.linalg.norm(zg - zf) / np.linalg.norm(zf)
    print("learned stage rel err", err, flush=True)
    assert err < 2e-2
    assert np.all(z[:, :, nn:] == 0.0)
    fem.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_flexible_pcg_with_network_stage_converges(mode):
    from anneal_amd.fem import FEM
    from anneal_amd.learned import UNet
    from anneal_amd import _lib
    shape = (32, 16, 16)
    g, fixed, Ee = clamp_problem(shape, 7)
    fem = FEM(g, np.ones(g.nelem, bool), fixed, 0.3, nrhs_max=2)
    K = fem.assemble(Ee)
    rng = np.random.default_rng(8)
    F = [fem.reduce_load(rng.normal(size=3 * g.nnode)) for _ in range(2)]
    u_plain = fem.solve(K, F, rtol=1e-8)
    it_plain = list(fem.cg_iters)
    net = UNet(seed=2)                      # untrained: the stage must not break convergence
    fem.set_learned(net, level=1, e_ref=1.0, mode=mode)
    u_l = fem.solve(K, F, rtol=1e-8)
    it_l = fem.cg_iters[len(it_plain):]
    print(f"mode {mode}: plain iterations {it_plain}, learned {it_l}", flush=True)
    for a, b_, f in zip(u_plain, u_l, F):
        assert np.linalg.norm(a - b_) <= 1e-6 * np.linalg.norm(a)
        res = f - K @ b_
        assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f) * 1.0001
    fem.close()


def test_trained_network_gain():
    """What training buys: energy error of the network's warm start and PCG iterations, trained vs not.

    Trained on pairs from seeded 48x24x24 cantilever SIMP runs, evaluated on a
    held-out run's designs (cold solves from x = 0 to rtol 1e-8).  The warm start
    is z1 = gamma P N(R F) with the energy-optimal gamma, so its relative energy
    error is sqrt(1 - <F, z1>^2 / (<z1, K z1> <F, u>)); a network that learned
    nothing scores about 1.
    """
    from anneal_amd import learned as Lm
    from anneal_amd.optimize import SimpRun
    X, T, dims = Lm.generate_pairs(150, 40, seed=0, kind="load")   # 6000 pairs: fewer overfit the 0.9 M parameters
    Xv, Tv, _ = Lm.generate_pairs(6, 40, seed=5000, kind="load")
    trained = Lm.UNet(seed=0)
    untrained = Lm.UNet(seed=0)
    Lm.train(trained, X, T, dims, 1500, batch=32, lr=2e-3, seed=0)
    lv_t, lv_u = Lm.evaluate(trained, Xv, Tv, dims), Lm.evaluate(untrained, Xv, Tv, dims)
    print(f"validation relative squared error: trained {lv_t:.3f}, untrained {lv_u:.3f}", flush=True)
    assert lv_t < 0.9 * lv_u

    # held-out designs: snapshots of a SIMP run with another seed
    g, fill, active, frozen, fixed, F, vf = Lm.training_problem(7777)
    designs = []
    run = SimpRun(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, vf, iters=30, r_mm=1.5, tol=0.0,
                  emin=1e-9, rtol=1e-6, log=lambda m: None)
    for it in range(30):
        run.step()
        if it in (0, 9, 19, 29):
            designs.append(run.Ee.clone())
    fem = run.fem
    ld = fem.ldn
    res = {}
    for name, net in (("plain", None), ("untrained", untrained), ("trained", trained)):
        iters, ratios = [], []
        for Ee in designs:
            fem.assemble_device(Ee)
            fem.rebuild()
            fem.set_learned(None)
            U = torch.zeros_like(run.F)
            its = fem.solve_device(run.F, U, 1, 1e-10, 2000, use_x0=False)
            if net is not None:
                fem.set_learned(net, level=1, e_ref=1.0, mode=1)
                r32 = run.F.float().contiguous()
                z = torch.zeros_like(r32)
                fem.learned_stage(r32, z, 1)
                z64 = z.double()
                Kz = torch.empty_like(z64)
                from anneal_amd import _lib
                _lib.call("anneal_fem_apply_f64", fem._ctx, _lib.ptr(z64), _lib.ptr(Kz), 1)
                fz, zkz, fu = float((run.F * z64).sum()), float((z64 * Kz).sum()), float((run.F * U).sum())
                ratios.append(np.sqrt(max(0.0, 1.0 - fz * fz / (zkz * fu))))
                U2 = torch.zeros_like(run.F)
                its = fem.solve_device(run.F, U2, 1, 1e-8, 2000, use_x0=False)
                assert float((U2 - U).norm() / U.norm()) < 1e-6
            else:
                U2 = torch.zeros_like(run.F)
                its = fem.solve_device(run.F, U2, 1, 1e-8, 2000, use_x0=False)
            iters.append(its[0])
        res[name] = (iters, ratios)
        print(f"{name}: cold-solve PCG iterations {iters}, warm-start energy error {np.round(ratios, 3)}", flush=True)
    fem.set_learned(None)
    assert np.mean(res["trained"][1]) < np.mean(res["untrained"][1])
