"""GPU parity of the density side of SIMP (filter, projection, sensitivities, OC, whole loop).

Everything here is FP64 on the device; the bar is 1e-12 relative for single
kernels against the oracle, and for whole SIMP runs the north-star bar:
per-iteration compliance within 1e-6 relative and final densities within
1e-4 max-abs of the oracle's FP64 trajectory (golden fixtures).
"""
import os

import numpy as np
import pytest

from oracle.fem_ref import FEMRef, Grid as GridRef
from oracle.optimize_ref import FilterRef, cantilever, heaviside, oc_update

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _setup(shape, seed, r, inactive=0.2, frozen_frac=0.1):
    rng = np.random.default_rng(seed)
    g = GridRef(shape, 1.0, (0, 0, 0))
    active = rng.random(g.nelem) >= inactive
    active[0] = True
    frozen = (rng.random(g.nelem) < frozen_frac) & active
    return g, active, frozen, rng


@pytest.mark.parametrize("shape,r,seed", [((9, 7, 5), 1.5, 0), ((40, 11, 13), 3.0, 1), ((33, 17, 10), 2.2, 2),
                                          ((32, 8, 12), 1.6, 3), ((35, 13, 11), 4.0, 4), ((41, 19, 12), 5.5, 5),
                                          ((30, 17, 14), 7.2, 6), ((6, 5, 4), 9.0, 7)])
def test_filter_forward_backward(torch, shape, r, seed):
    """Radii 1.5-6 run the ring-sum kernel (R = 1..5), 7.2 and 9 the gather kernels (R = 7, 8; the
    second wider than the grid)."""
    print(f"filter parity shape={shape} r={r}", flush=True)
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import Filter
    g, active, _, rng = _setup(shape, seed, r)
    ref = FilterRef(g, active, r)
    f = Filter(Grid(shape, 1.0, (0, 0, 0)), active, r)
    x = rng.random(ref.act.size)
    d = rng.standard_normal(ref.act.size)
    for got, want in ((f.forward(x), ref.forward(x)), (f.backward(d), ref.backward(d))):
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("shape,r,seed", [((9, 7, 5), 1.5, 0), ((40, 11, 13), 3.0, 1), ((35, 13, 11), 4.0, 4),
                                          ((30, 17, 14), 7.2, 6)])
def test_backward_pair_equals_two_backward_passes(torch, shape, r, seed):
    """The paired backward pass of the sensitivities gives the bits of two single passes and the design
    statistics of the pair (count, sum of -y1, sum of y2 over active non-frozen voxels) to FP64 rounding;
    radii 1.5-4 run the two-channel ring kernel, 7.2 the gather passes."""
    print(f"backward pair shape={shape} r={r}", flush=True)
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import DeviceFilter
    g, active, frozen, rng = _setup(shape, seed, r)
    df = DeviceFilter(Grid(shape, 1.0, (0, 0, 0)), active, r, frozen)
    n = g.nelem
    d1 = torch.from_numpy(rng.standard_normal(n)).cuda()
    d2 = torch.from_numpy(rng.random(n)).cuda()
    y1, y2 = df.backward(d1), df.backward(d2)
    p1, p2 = torch.empty_like(d1), torch.empty_like(d2)
    st = torch.full((6,), 7.0, dtype=torch.float64, device="cuda")
    df.backward_pair(d1, p1, d2, p2, st)
    assert torch.equal(p1, y1) and torch.equal(p2, y2)
    des = active & ~frozen
    a1, a2, st = y1.cpu().numpy(), y2.cpu().numpy(), st.cpu().numpy()
    assert st[0] == 0.0 and st[1] == 0.0 and st[2] == 0.0
    assert st[3] == des.sum()
    assert abs(st[4] - (-a1[des]).sum()) <= 1e-12 * np.abs(a1[des]).sum()
    assert abs(st[5] - a2[des].sum()) <= 1e-12 * np.abs(a2[des]).sum()
    df.backward_pair(d1, p1, d2, p2)                # without statistics
    assert torch.equal(p1, y1) and torch.equal(p2, y2)


@pytest.mark.parametrize("beta,r", [(0.0, 1.5), (1.0, 1.5), (4.0, 1.5), (2.0, 4.0), (4.0, 5.5), (2.0, 7.2)])
def test_physical_projection(torch, beta, r):
    from anneal_amd.optimize import DeviceFilter
    from anneal_amd.fem import Grid
    shape = (20, 9, 7)
    g, active, frozen, rng = _setup(shape, 5, r)
    ref = FilterRef(g, active, r)
    x_act = rng.random(ref.act.size)
    phi_act = rng.uniform(0.2, 1.0, ref.act.size)
    xp, dxp = heaviside(ref.forward(x_act), beta)
    xp = xp.copy()
    xp[frozen[ref.act]] = 1.0
    df = DeviceFilter(Grid(shape, 1.0, (0, 0, 0)), active, r, frozen)

    def full(v):
        a = np.zeros(g.nelem)
        a[ref.act] = v
        return torch.from_numpy(a).cuda()
    X, PHI = full(x_act), full(phi_act)
    XP, DXP = torch.zeros_like(X), torch.zeros_like(X)
    vol = torch.zeros(1, dtype=torch.float64, device="cuda")
    df.physical(X, beta, PHI, XP, DXP, vol)
    got_xp = XP.cpu().numpy()[ref.act]
    got_dxp = DXP.cpu().numpy()[ref.act]
    assert np.abs(got_xp - xp).max() <= 1e-13
    assert np.abs(got_dxp - dxp).max() <= 1e-12 * max(1.0, np.abs(dxp).max())
    assert abs(vol.item() - (phi_act * xp).sum()) <= 1e-12 * (phi_act * xp).sum()


@pytest.mark.parametrize("beta,target,r", [(0.0, "inside", 1.5), (2.0, "inside", 1.5), (0.0, "unreachable", 1.5),
                                           (2.0, "inside", 4.0), (0.0, "inside", 5.5), (2.0, "inside", 7.2),
                                           (0.0, "unreachable", 7.2)])
def test_oc_bisection(torch, beta, target, r):
    """Device OC update vs the oracle's 36-trial bisection (optimize.py:144-161): same trial
    count and the same last trial design, from fewer volume evaluations."""
    from anneal_amd.optimize import DeviceFilter
    from anneal_amd.fem import Grid
    shape = (24, 12, 10)
    g, active, frozen, rng = _setup(shape, 9, r)
    ref = FilterRef(g, active, r)
    act = ref.act
    frz, des = frozen[act], ~frozen[act]
    x = rng.uniform(0.05, 0.9, act.size)
    x[frz] = 1.0
    dcx = -rng.random(act.size) * 10 ** rng.uniform(-2, 1, act.size)
    dvx = rng.uniform(0.1, 1.0, act.size)
    phi = rng.uniform(0.5, 1.0, act.size)

    def physical(xd):
        xp, _ = heaviside(ref.forward(xd), beta)
        xp = xp.copy()
        xp[frz] = 1.0
        return xp
    scale = max(float(np.mean(-dcx[des]) / max(np.mean(dvx[des]), 1e-30)), 1e-300)
    mv = 0.2 if beta <= 1 else 0.1

    def vol(lam):
        xn = np.clip(x * np.sqrt(np.maximum(-dcx, 0) / np.maximum(dvx * lam, 1e-300)), np.maximum(0, x - mv),
                     np.minimum(1, x + mv))
        xn[frz] = 1.0
        return (phi * physical(xn)).sum()
    if target == "inside":       # a target the bracket contains, as in a real OC step
        vt = 0.5 * (vol(scale * 1e-8) + vol(scale * 1e8))
    else:                        # more volume than any trial reaches: every trial decides "lo"
        vt = 0.3 * phi.sum()
        assert vol(scale * 1e8) > vt
    want, steps_ref = oc_update(x, dcx, dvx, phi, frz, des, vt, 0.2, beta, physical)
    df = DeviceFilter(Grid(shape, 1.0, (0, 0, 0)), active, r, frozen)

    def full(v):
        a = np.zeros(g.nelem)
        a[act] = v
        return torch.from_numpy(a).cuda()
    XN = torch.zeros(g.nelem, dtype=torch.float64, device="cuda")
    X, DCX, DVX, PHI = full(x), full(dcx), full(dvx), full(phi)
    # hints far off, near and absent, with the default and with narrow or wide first brackets: only the
    # number of evaluations may change, never the trials or the design
    for hint, width in ((0.0, 0.0), (scale * 0.37, 0.0), (scale * 1e6, 0.0), (scale * 0.37, 1e-4),
                        (scale * 0.37, 0.3)):
        info = {}
        steps = df.oc_bisect(X, DCX, DVX, PHI, beta, 0.2, scale * 1e-8, scale * 1e8, vt, XN, hint=hint, info=info,
                             width=width)
        got = XN.cpu().numpy()[act]
        print(f"oc beta={beta} {target} hint={hint / scale:.3g}*scale width={width}: {steps} trials from "
              f"{info['evals']} volume evaluations", flush=True)
        assert steps == steps_ref
        assert np.abs(got - want).max() <= 1e-10
        assert info["evals"] < steps_ref
    # a warm start at the previous root needs only a handful of evaluations
    info = {}
    df.oc_bisect(X, DCX, DVX, PHI, beta, 0.2, scale * 1e-8, scale * 1e8, vt, XN, hint=info_lam(df, X, DCX, DVX,
                 PHI, beta, scale, vt, XN), info=info)
    if target == "inside":
        assert info["evals"] <= 12
    # the split calls with two contexts in flight at once (as a batch of runs uses them): same trials
    # and design, the second context started from the warm hint
    df2 = DeviceFilter(Grid(shape, 1.0, (0, 0, 0)), active, r, frozen)
    XN2 = torch.zeros_like(XN)
    XN.zero_()
    df.oc_begin(X, DCX, DVX, PHI, beta, 0.2, scale * 1e-8, scale * 1e8, vt, XN)
    df2.oc_begin(X, DCX, DVX, PHI, beta, 0.2, scale * 1e-8, scale * 1e8, vt, XN2, hint=scale * 0.37)
    pending = [df, df2]
    while pending:
        pending = [d for d in pending if not d.oc_poll()]
    infos = [{}, {}]
    assert df.oc_end(infos[0]) == steps_ref and df2.oc_end(infos[1]) == steps_ref
    for got in (XN, XN2):
        assert np.abs(got.cpu().numpy()[act] - want).max() <= 1e-10
    df2.close()


def info_lam(df, X, DCX, DVX, PHI, beta, scale, vt, XN):
    info = {}
    df.oc_bisect(X, DCX, DVX, PHI, beta, 0.2, scale * 1e-8, scale * 1e8, vt, XN, info=info)
    return info["lam_last"] * 1.003


def test_sensitivity_kernel(torch):
    from anneal_amd import _lib
    from anneal_amd.fem import FEM, Grid
    shape = (10, 6, 5)
    g, active, frozen, rng = _setup(shape, 11, 1.5)
    n = np.arange(g.nnode)
    clamp = n[n % (g.nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    ref = FEMRef(g, active, fixed, 0.3)
    fem = FEM(Grid(shape, 1.0, (0, 0, 0)), active, fixed, 0.3, nrhs_max=2)
    act = ref.act
    xphys = rng.uniform(0.01, 1, act.size)
    dxp = rng.uniform(0.5, 2, act.size)
    phi = rng.uniform(0.5, 1, act.size)
    us = [rng.standard_normal(3 * g.nnode) for _ in range(2)]
    coef = np.array([0.7, 1.3])
    p, emin, E0 = 2.5, 1e-9, 3.0
    dce = [-p * xphys ** (p - 1) * (1 - emin) * E0 * phi * ref.element_energy(u) for u in us]
    dc = np.minimum(coef[0] * dce[0] + coef[1] * dce[1], 0.0)

    def full(v):
        a = np.zeros(g.nelem)
        a[act] = v
        return torch.from_numpy(a).cuda()
    U = fem._to_device(us, reduced=False)
    flags = torch.from_numpy(np.asarray(active, np.uint8)).cuda()
    DC, DV = torch.zeros(g.nelem, dtype=torch.float64, device="cuda"), torch.zeros(g.nelem, dtype=torch.float64,
                                                                                     device="cuda")
    # keep every device buffer referenced until the kernel has run
    X, DXP, PHI, COEF = full(xphys), full(dxp), full(phi), torch.from_numpy(coef).cuda()
    _lib.call("anneal_fem_sensitivity", fem._ctx, _lib.ptr(U), 2, _lib.ptr(flags), _lib.ptr(X), _lib.ptr(DXP),
              _lib.ptr(PHI), _lib.ptr(COEF), p, emin, E0, _lib.ptr(DC), _lib.ptr(DV))
    torch.cuda.synchronize()
    assert np.abs(DC.cpu().numpy()[act] - dc * dxp).max() <= 1e-12 * np.abs(dc * dxp).max()
    assert np.abs(DV.cpu().numpy()[act] - phi * dxp).max() == 0.0


def _run_golden(torch, name, rtol):
    if not os.path.exists(os.path.join(GOLDEN, name)):
        pytest.skip(f"golden fixture {name} not generated (tests/golden/make_golden.py)")
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import simp
    gold = np.load(os.path.join(GOLDEN, name))
    shape = tuple(int(v) for v in gold["shape"])
    print(f"simp parity {shape} iters={int(gold['iters'])}", flush=True)
    g, fill, active, frozen, fixed, F = cantilever(shape)
    r = simp(Grid(shape, 1.0, (0, 0, 0)), fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3,
             iters=int(gold["iters"]), r_mm=float(gold["r_mm"]), tol=0.0, emin=1e-9, rtol=rtol,
             log=lambda *a: None)
    comp = np.array([h["compliance"][0] for h in r["history"]])
    rel = np.abs(comp - gold["compliance"]) / np.abs(gold["compliance"])
    dx = np.abs(r["x"] - gold["x_final"]).max()
    print(f"  max rel compliance err {rel.max():.3e}, final density max-abs err {dx:.3e}", flush=True)
    assert [h["p"] for h in r["history"]] == gold["p"].tolist()
    assert [h["beta"] for h in r["history"]] == gold["beta"].tolist()
    assert rel.max() <= 1e-6
    assert dx <= 1e-4
    vol = np.array([h["volumeFraction"] for h in r["history"]])
    assert np.abs(vol - gold["volume"]).max() <= 1e-6
    return r


def test_simp_small_matches_golden(torch):
    _run_golden(torch, "cantilever_16x8x8_it30.npz", 1e-8)


def test_simp_config1_matches_golden(torch):
    """48x24x24 cantilever, 50 iterations, Emin 1e-9, CG rtol 1e-8 (the benchmark's CPU workload)."""
    _run_golden(torch, "cantilever_48x24x24_it50.npz", 1e-8)


def test_v6_simp_beats_uniform(torch):
    """validate.py:124-129: > 20 % stiffer than uniform at equal mass, |vol err| < 1e-3."""
    from anneal_amd.optimize import benchmark_cantilever
    r = benchmark_cantilever()
    print(f"V6 stiffer {r['stiffer_pct']:.1f}% vol_err {r['vol_err']:.1e}", flush=True)
    assert r["stiffer_pct"] > 20 and abs(r["vol_err"]) < 1e-3


@pytest.mark.parametrize("r_mm", [4.0, 7.2])
def test_simp_with_wide_filter_matches_oracle(torch, r_mm):
    """simp() with filter radii the reference's run.py produces for thin-walled materials at fine voxel sizes
    (run.py:197-199: r = max(minMember / 2, 1.5 h)): R = 3 (ring kernel) and R = 7 (gather kernels)."""
    from oracle.optimize_ref import simp_ref
    from anneal_amd.fem import Grid
    from anneal_amd.optimize import simp
    shape = (20, 10, 10)
    g, fill, active, frozen, fixed, F = cantilever(shape)
    kw = dict(iters=12, r_mm=r_mm, tol=0.0, emin=1e-9, log=lambda *a: None)
    want = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, rtol=1e-11, **kw)
    print(f"simp with r_mm={r_mm} on {shape}", flush=True)
    got = simp(Grid(shape, 1.0, (0, 0, 0)), fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3,
               rtol=1e-10, **kw)
    c_got = np.array([h["compliance"][0] for h in got["history"]])
    c_want = np.array([h["compliance"][0] for h in want["history"]])
    assert (np.abs(c_got - c_want) / c_want).max() <= 1e-6
    assert np.abs(got["x"] - want["x"]).max() <= 1e-4
    assert [h["ocSteps"] for h in got["history"]] == [h["ocSteps"] for h in want["history"]]
