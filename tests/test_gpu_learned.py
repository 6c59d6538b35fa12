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
