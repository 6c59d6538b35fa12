"""FP64 CPU oracle for the SIMP loop (test infrastructure, never product code).

Restates the reference's ``anneal/optimize.py`` (density filter, Heaviside
projection, penalty/beta continuation, sensitivities, OC bisection) on top of
``oracle.fem_ref``.  The reference hard-codes ``EMIN = 1e-6`` (optimize.py:18)
and the solve tolerance ``rtol=1e-6`` (optimize.py:122); both are parameters
here with those defaults so that the benchmark workloads (Emin 1e-9, rtol 1e-8)
can be stated exactly.

Pinning: the filter is checked against a brute-force neighbourhood sum and
against adjointness; the projection against its closed form and finite
differences; the OC update against the bisection invariants (final volume
within tolerance of the target, move limits, frozen = 1); the whole loop by
V6 (validate.py:124-129: > 20 % stiffer than uniform, |vol err| < 1e-3) and
the continuation schedule tables; ``element_sensitivity`` and ``combine_cases``
by central differences of the objective through filter, projection, stiffness
law and solve, for "sum" and the p-norm "worst" (tests/test_oracle_sensitivity.py).
"""
import time

import numpy as np
from scipy import ndimage

from .fem_ref import FEMRef, Grid

EMIN = 1e-6        # optimize.py:18
PNORM = 8.0        # optimize.py:19
ETA = 0.5          # optimize.py:20
BETA_MAX = 4.0     # optimize.py:21


def cone_kernel(r_vox):
    """max(0, r - |d|) on [-R, R]^3 with R = max(ceil(r) - 1, 1) (optimize.py:28-31)."""
    R = max(int(np.ceil(r_vox)) - 1, 1)
    ax = np.arange(-R, R + 1, dtype=float)
    d = np.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)
    return np.maximum(0.0, r_vox - d)


class FilterRef:
    """Linear hat filter normalised over active voxels (optimize.py:24-47)."""

    def __init__(self, grid, active, r_vox):
        self.g = grid
        self.active3 = grid.to3d(active).astype(float)
        self.k = cone_kernel(r_vox)
        Hs = ndimage.convolve(self.active3, self.k, mode="constant")
        Hs[Hs == 0] = 1.0
        self.Hs = Hs
        self.act = np.flatnonzero(active)

    def _full(self, v):
        f = np.zeros(self.g.nelem)
        f[self.act] = v
        return self.g.to3d(f)

    def forward(self, x_act):
        y = ndimage.convolve(self._full(x_act) * self.active3, self.k, mode="constant") / self.Hs
        return self.g.to1d(y)[self.act]

    def backward(self, d_act):
        y = ndimage.convolve(self._full(d_act) / self.Hs, self.k, mode="constant") * self.active3
        return self.g.to1d(y)[self.act]


def heaviside(xt, beta, eta=ETA):
    """Projection and its derivative; identity for beta <= 0 (optimize.py:81-87)."""
    if beta <= 0:
        return xt, np.ones_like(xt)
    a, b = np.tanh(beta * eta), np.tanh(beta * (1.0 - eta))
    t = np.tanh(beta * (xt - eta))
    return (a + t) / (a + b), beta * (1.0 - t * t) / (a + b)


def penalty_schedule(it, iters, p_final, warm):
    """Penalty continuation (optimize.py:109-110)."""
    hold = max(1, min(3, int(0.4 * iters / max((p_final - 1.0) / 0.25, 1))))
    return p_final if warm else min(p_final, 1.0 + 0.25 * ((it - 1) // hold))


def element_sensitivity(fem, u_red, xPhys, p, E0, phi, emin=EMIN):
    """d(f.u)/d(xPhys) of one load case: -p xPhys^(p-1) (1-emin) E0 phi u_e^T ke u_e (optimize.py:118, 126-129)."""
    return -p * xPhys ** (p - 1) * (1 - emin) * E0 * phi * fem.element_energy(fem.expand(u_red))


def combine_cases(cs, dce, w, c0, combine):
    """Objective and clamped combined sensitivity over the load cases (optimize.py:130-140).

    "sum" (or one case): obj = w.cs, dc = sum w_i dce_i.  "worst": r = w cs / c0,
    obj = ||r||_8, dc = sum (r_i/obj)^7 w_i / c0_i dce_i.  Then dc = min(dc, 0).
    """
    if combine == "sum" or len(cs) == 1:
        obj = float(w @ cs)
        dc = sum(wi * d for wi, d in zip(w, dce))
    else:
        r = w * cs / c0
        obj = float((r ** PNORM).sum() ** (1 / PNORM))
        coef = (r / obj) ** (PNORM - 1) * w / c0
        dc = sum(ci * d for ci, d in zip(coef, dce))
    return obj, np.minimum(dc, 0.0)


def oc_update(x, dcx, dvx, phi, frz, des, vtarget, move, beta, physical):
    """Geometric bisection on the OC multiplier, returns the last trial design (optimize.py:144-161)."""
    if des.any():
        scale = max(float(np.mean(-dcx[des]) / max(np.mean(dvx[des]), 1e-30)), 1e-300)
    else:
        scale = 1.0
    lo, hi = scale * 1e-8, scale * 1e8
    mv = move if beta <= 1 else move / 2
    xd = x.copy()
    steps = 0
    for _ in range(100):
        lam = np.sqrt(lo * hi)
        ratio = np.sqrt(np.maximum(-dcx, 0.0) / np.maximum(dvx * lam, 1e-300))
        xn = np.clip(x * ratio, np.maximum(0.0, x - mv), np.minimum(1.0, x + mv))
        xn[frz] = 1.0
        vol = (phi * physical(xn)).sum()
        if vol > vtarget:
            lo = lam
        else:
            hi = lam
        xd = xn
        steps += 1
        if hi / lo < 1 + 1e-9:
            break
    return xd, steps


def simp_ref(grid, fill, active, frozen, fixed, loads, weights, E0, nu, vf,
             iters=80, r_mm=None, p_final=3.0, move=0.2, combine="sum", log=print, tol=0.01,
             x0=None, beta_max=BETA_MAX, emin=EMIN, rtol=1e-6, record=None):
    """The SIMP driver (optimize.py:50-180).  ``record`` (a list) receives per-iteration arrays."""
    t_start = time.time()
    fem = FEMRef(grid, active, fixed, nu, log=log)
    act = fem.act
    phi = np.asarray(fill)[act]
    frz = np.asarray(frozen)[act]
    des = ~frz
    h = grid.h
    r_vox = max((r_mm or 1.5 * h) / h, 1.5)
    filt = FilterRef(grid, active, r_vox)
    cases = list(loads)
    Fr = [fem.reduce_load(loads[c]) for c in cases]
    w = np.array([weights[c] for c in cases])
    if all(np.allclose(f, 0) for f in Fr):
        raise ValueError("All loads are zero after mapping onto the grid. Check the load anchors' regions.")
    Vtot = phi.sum()
    Vtarget = vf * Vtot
    Vfrozen = phi[frz].sum()
    x = np.ones(act.size)
    x[des] = np.clip((Vtarget - Vfrozen) / max(phi[des].sum(), 1e-12), 0.01, 1.0)
    warm = x0 is not None
    if warm:
        xs = np.clip(np.asarray(x0)[act], 0.0, 1.0)
        cur = (phi[des] * xs[des]).sum()
        if cur > 0:
            x[des] = np.clip(xs[des] * (Vtarget - Vfrozen) / cur, 0.001, 1.0)
    beta = [beta_max / 2 if warm else 0.0]

    def physical(xd, with_grad=False):
        xp, dxp = heaviside(filt.forward(xd), beta[0])
        xp = xp.copy()
        xp[frz] = 1.0
        return (xp, dxp) if with_grad else xp

    hist = []
    U = [None] * len(cases)
    c0 = None
    base_iters = None
    xPhys = physical(x)
    beta_step = max(5, int(0.12 * iters))
    beta_since = 0
    for it in range(1, iters + 1):
        t0 = time.time()
        p = penalty_schedule(it, iters, p_final, warm)
        if p >= p_final and not warm and beta_max > 0:        # optimize.py:111-116
            if beta[0] == 0:
                beta[0], beta_since = 1.0, 0
            elif beta_since >= beta_step and beta[0] < beta_max:
                beta[0], beta_since = min(beta_max, 2 * beta[0]), 0
            beta_since += 1
        xPhys, dxp = physical(x, with_grad=True)
        Ee = E0 * phi * (emin + xPhys ** p * (1 - emin))
        K = fem.assemble(Ee)
        last = fem.cg_iters[-1] if fem.cg_iters else None
        rebuild = base_iters is None or (last is not None and last > 1.6 * base_iters)
        U = fem.solve(K, Fr, x0_list=None if U[0] is None else U, rebuild=rebuild, rtol=rtol)
        if rebuild:
            base_iters = fem.cg_iters[-1]
        cs = np.array([f @ u for f, u in zip(Fr, U)])
        dce = [element_sensitivity(fem, u, xPhys, p, E0, phi, emin) for u in U]
        if c0 is None:
            c0 = cs.copy()
        obj, dc = combine_cases(cs, dce, w, c0, combine)
        dcx = filt.backward(dc * dxp)
        dvx = filt.backward(phi * dxp)
        xd, steps = oc_update(x, dcx, dvx, phi, frz, des, Vtarget, move, beta[0], physical)
        change = float(np.abs(xd - x)[des].max()) if des.any() else 0.0
        x = xd
        xPhys = physical(x)
        grey = float(((xPhys > 0.1) & (xPhys < 0.9))[des].mean()) if des.any() else 0.0
        hist.append({"iter": it, "objective": obj, "compliance": cs.tolist(), "p": p, "beta": beta[0],
                     "volumeFraction": float((phi * xPhys).sum() / Vtot), "change": change,
                     "grey": grey, "cgIters": fem.cg_iters[-len(cases):], "amgRebuilt": bool(rebuild),
                     "seconds": time.time() - t0, "ocSteps": steps})
        if record is not None:
            record.append({"x": x.copy(), "xPhys": xPhys.copy()})
        log(f"  it {it:3d}  p {p:.2f}  b {beta[0]:.0f}  obj {obj:.4g}  vol {hist[-1]['volumeFraction']:.3f}  "
            f"change {change:.3f}  grey {grey:.2f}  cg {fem.cg_iters[-1]}{' amg' if rebuild else ''}  "
            f"{hist[-1]['seconds']:.1f}s")
        if p >= p_final and beta[0] >= beta_max and it > (5 if warm else 10) and change < tol:
            log(f"  converged at iteration {it}")
            break
    full = np.zeros(grid.nelem)
    full[act] = xPhys
    return {"x": full, "history": hist, "fem_builds": fem.amg_builds, "rMM": r_vox * h,
            "seconds": time.time() - t_start, "cases": cases}


def cantilever(shape, h=1.0, load=-1.0):
    """The benchmark cantilever: x=0 face clamped, unit -z point load at (nx, ny/2, nz/2).

    Inputs for simp() in the argument order of optimize.py:50-51.
    """
    g = Grid(shape, h, (0, 0, 0))
    nx, ny, nz = g.shape
    n = np.arange(g.nnode)
    ijk = np.stack([n % (nx + 1), (n // (nx + 1)) % (ny + 1), n // ((nx + 1) * (ny + 1))], axis=1)
    clamp = n[ijk[:, 0] == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    tip = nx + (nx + 1) * (ny // 2 + (ny + 1) * (nz // 2))
    F = np.zeros(3 * g.nnode)
    F[3 * tip + 2] = load
    fill = np.ones(g.nelem)
    active = np.ones(g.nelem, bool)
    frozen = np.zeros(g.nelem, bool)
    return g, fill, active, frozen, fixed, F


def benchmark_cantilever_ref(log=lambda *a: None, emin=EMIN):
    """V6 (optimize.py:183-210): SIMP design vs a uniform one at equal mass and penalty 3."""
    g = Grid((32, 8, 12), 1.0, (0, 0, 0))
    fill = np.ones(g.nelem)
    active = np.ones(g.nelem, bool)
    frozen = np.zeros(g.nelem, bool)
    n = np.arange(g.nnode)
    p = g.node_coords(n)
    clamp = n[p[:, 0] < 1e-9]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    tip = n[(abs(p[:, 0] - 32) < 1e-9) & (abs(p[:, 2] - 6) < 1e-9)]
    F = np.zeros(3 * g.nnode)
    F[3 * tip + 2] = -1.0 / tip.size
    vf = 0.3
    r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1000.0, 0.3, vf,
                 iters=40, r_mm=1.6, log=log, emin=emin)
    fem = FEMRef(g, active, fixed, 0.3)
    Fr = fem.reduce_load(F)

    def comp(xp):
        u = fem.solve(fem.assemble(1000.0 * (emin + xp ** 3 * (1 - emin))), [Fr], rebuild=True)[0]
        return float(Fr @ u)
    c_opt, c_uni = comp(r["x"]), comp(np.full(g.nelem, vf))
    return {"stiffer_pct": 100 * (1 - c_opt / c_uni), "vol_err": r["history"][-1]["volumeFraction"] - vf}
