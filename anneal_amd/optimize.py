"""SIMP topology optimisation on the GPU (drop-in for the reference's anneal/optimize.py).

Same names, signatures, history keys and log lines as the reference
(optimize.py:18-210).  The loop keeps every per-voxel array on the device; each
iteration returns only scalars to the host:

* ``physical``: one fused filter + Heaviside projection + frozen kernel;
* stiffness law -> multigrid PCG for all load cases as one block;
* one fused sensitivity kernel (element energies, penalty chain rule,
  load-case combination, clamp), then two adjoint filter passes;
* the OC bisection runs on the device: each trial is one fused kernel
  (trial design -> filter -> projection -> FP64 volume partials) followed by
  a one-block kernel that sums the partials in a fixed order and moves the
  bracket, so the 36 trials need no host round trip.

The density side is FP64 throughout (the reference's arithmetic); only the
elasticity operator inside PCG runs in FP32, with FP64 refinement.
"""
import time

import numpy as np

from . import _lib
from .fem import FEM, Grid, _torch

EMIN = 1e-6        # optimize.py:18
PNORM = 8.0        # optimize.py:19
ETA = 0.5          # optimize.py:20 (the projection kernel uses eta = 0.5)
BETA_MAX = 4.0     # optimize.py:21


def _flags(grid, active, frozen=None):
    f = np.asarray(active).astype(bool).ravel().astype(np.uint8)
    if frozen is not None:
        f |= (np.asarray(frozen).astype(bool).ravel() & (f == 1)).astype(np.uint8) << 1
    return np.ascontiguousarray(f)


class DeviceFilter:
    """The hat filter as a device context over full-grid float64 CUDA tensors."""

    def __init__(self, grid, active, r_vox, frozen=None):
        torch = _torch()
        self.g = grid
        self.r_vox = float(r_vox)
        self.flags_host = _flags(grid, active, frozen)
        self.stream = torch.cuda.current_stream()
        ctx = _lib.c_void_p()
        _lib.call("anneal_filter_create", _lib.ctypes.byref(ctx), grid.nx, grid.ny, grid.nz, self.r_vox,
                  _lib.ptr(self.flags_host), _lib.c_void_p(self.stream.cuda_stream))
        self._ctx = ctx
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.flags = torch.from_numpy(self.flags_host).to(self.device)

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_filter_destroy", self._ctx)
            self._ctx = _lib.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x, out=None):
        out = _torch().empty_like(x) if out is None else out
        _lib.call("anneal_filter_forward", self._ctx, _lib.ptr(x), _lib.ptr(out))
        return out

    def backward(self, d, out=None):
        out = _torch().empty_like(d) if out is None else out
        _lib.call("anneal_filter_backward", self._ctx, _lib.ptr(d), _lib.ptr(out))
        return out

    def physical(self, x, beta, phi, xphys, dxp, vol=None):
        _lib.call("anneal_filter_physical", self._ctx, _lib.ptr(x), float(beta), _lib.ptr(phi), _lib.ptr(xphys),
                  _lib.ptr(dxp), _lib.ptr(vol))

    def oc_bisect(self, x, dcx, dvx, phi, beta, move, lo, hi, vtarget, xnew):
        steps = _lib.ctypes.c_int()
        _lib.call("anneal_oc_bisect", self._ctx, _lib.ptr(x), _lib.ptr(dcx), _lib.ptr(dvx), _lib.ptr(phi),
                  float(beta), float(move), float(lo), float(hi), float(vtarget), _lib.ptr(xnew),
                  _lib.ctypes.byref(steps))
        return steps.value


class Filter:
    """Linear hat filter normalised over active voxels (optimize.py:24-47), on the GPU.

    Host interface of the reference: ``forward``/``backward`` take and return
    float64 arrays over the active voxels.
    """

    def __init__(self, grid, active, r_vox):
        self.g = grid
        self.act = np.flatnonzero(np.asarray(active).ravel())
        self.dev = DeviceFilter(grid, active, r_vox)
        R = max(int(np.ceil(r_vox)) - 1, 1)
        ax = np.arange(-R, R + 1)
        X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
        self.k = np.maximum(0.0, r_vox - np.sqrt(X ** 2 + Y ** 2 + Z ** 2))

    def _full(self, v_act):
        torch = _torch()
        f = np.zeros(self.g.nelem)
        f[self.act] = v_act
        return torch.from_numpy(f).to(self.dev.device)

    def forward(self, x_act):
        return self.dev.forward(self._full(x_act)).cpu().numpy()[self.act]

    def backward(self, d_act):
        return self.dev.backward(self._full(d_act)).cpu().numpy()[self.act]


def _schedule_p(it, iters, p_final, warm):
    """Penalty continuation (optimize.py:108-110)."""
    hold = max(1, min(3, int(0.4 * iters / max((p_final - 1.0) / 0.25, 1))))
    return p_final if warm else min(p_final, 1.0 + 0.25 * ((it - 1) // hold))


def simp(grid, fill, active, frozen, fixed, loads, weights, E0, nu, vf,
         iters=80, r_mm=None, p_final=3.0, move=0.2, combine="sum", log=print, tol=0.01, x0=None,
         beta_max=BETA_MAX, *, emin=None, rtol=1e-6, maxiter=2000, fem=None, timers=None):
    """SIMP loop of optimize.py:50-180 with device-resident state.

    Extra keyword-only arguments (defaults reproduce the reference): ``emin``
    (module ``EMIN`` when None, optimize.py:18), ``rtol`` of the inner solves
    (optimize.py:122 uses 1e-6), ``maxiter`` (fem.py:206), an existing ``fem``
    context to reuse, and ``timers`` (a dict that receives per-phase seconds).
    """
    torch = _torch()
    t_start = time.time()
    emin = EMIN if emin is None else float(emin)
    fill = np.asarray(fill, float).ravel()
    active = np.asarray(active).astype(bool).ravel()
    frozen = np.asarray(frozen).astype(bool).ravel()
    cases = list(loads)
    if fem is None:
        fem = FEM(grid, active, fixed, nu, log=log, nrhs_max=max(1, len(cases)))
    act = fem.act
    dev = fem.device
    phi_h = fill[act]
    frz_h = frozen[act]
    des_h = ~frz_h
    h = grid.h
    r_vox = max((r_mm or 1.5 * h) / h, 1.5)
    filt = DeviceFilter(grid, active, r_vox, frozen)
    Fr = [fem.reduce_load(loads[c]) for c in cases]
    w = np.array([weights[c] for c in cases], dtype=float)
    if all(np.allclose(f, 0) for f in Fr):
        raise ValueError("All loads are zero after mapping onto the grid. Check the load anchors' regions.")

    Vtot = phi_h.sum()
    Vtarget = vf * Vtot
    Vfrozen = phi_h[frz_h].sum()
    x_h = np.ones(act.size)
    x_h[des_h] = np.clip((Vtarget - Vfrozen) / max(phi_h[des_h].sum(), 1e-12), 0.01, 1.0)
    warm = x0 is not None
    if warm:
        xs = np.clip(np.asarray(x0, float).ravel()[act], 0.0, 1.0)
        cur = (phi_h[des_h] * xs[des_h]).sum()
        if cur > 0:
            x_h[des_h] = np.clip(xs[des_h] * (Vtarget - Vfrozen) / cur, 0.001, 1.0)

    nel = grid.nelem

    def full_dev(v_act):
        f = np.zeros(nel)
        f[act] = v_act
        return torch.from_numpy(f).to(dev)

    f64 = dict(dtype=torch.float64, device=dev)
    x = full_dev(x_h)
    phi = full_dev(phi_h)
    xphys, dxp = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
    Ee = torch.zeros(nel, **f64)
    dcdx, dvdx = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
    dcx, dvx = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
    xnew = torch.zeros(nel, **f64)
    stats = torch.zeros(8, **f64)
    work = torch.zeros(4096, **f64)
    ncase = len(cases)
    F = fem._to_device(Fr)
    U = torch.zeros_like(F)
    cs_dev = torch.zeros(ncase, **f64)
    coef_dev = torch.zeros(ncase, **f64)
    stream = _lib.c_void_p(torch.cuda.current_stream().cuda_stream)
    flags = filt.flags
    T = timers if timers is not None else {}
    for k in ("physical", "solve", "sensitivity", "oc", "stats"):
        T.setdefault(k, 0.0)

    def tick():
        torch.cuda.synchronize()
        return time.perf_counter()

    beta = beta_max / 2 if warm else 0.0
    hist = []
    c0 = None
    base_iters = None
    have_u = False
    filt.physical(x, beta, phi, xphys, dxp)
    beta_step = max(5, int(0.12 * iters))
    beta_since = 0
    p = 1.0
    for it in range(1, iters + 1):
        t0 = time.time()
        p = _schedule_p(it, iters, p_final, warm)
        if p >= p_final and not warm and beta_max > 0:         # optimize.py:111-116
            if beta == 0:
                beta, beta_since = 1.0, 0
            elif beta_since >= beta_step and beta < beta_max:
                beta, beta_since = min(beta_max, 2 * beta), 0
            beta_since += 1
        ta = tick()
        filt.physical(x, beta, phi, xphys, dxp)
        _lib.call("anneal_stiffness_law", _lib.ptr(xphys), _lib.ptr(phi), _lib.ptr(flags), float(E0), emin,
                  float(p), nel, _lib.ptr(Ee), stream)
        K = fem.assemble_device(Ee)
        tb = tick()
        T["physical"] += tb - ta
        last = fem.cg_iters[-1] if fem.cg_iters else None
        rebuild = base_iters is None or (last is not None and last > 1.6 * base_iters)
        if rebuild:
            fem.amg_builds += 1
        its = fem.solve_device(F, U, ncase, rtol, maxiter, use_x0=have_u)
        fem.cg_iters.extend(its)
        have_u = True
        if rebuild:
            base_iters = fem.cg_iters[-1]
        _lib.call("anneal_dot_f64", _lib.ptr(F), _lib.ptr(U), grid.nnode, fem.ldn, ncase, _lib.ptr(cs_dev),
                  _lib.ptr(work), stream)
        cs = cs_dev.cpu().numpy()
        tc = tick()
        T["solve"] += tc - tb
        if c0 is None:
            c0 = cs.copy()
        if combine == "sum" or ncase == 1:
            obj = float(w @ cs)
            coef = w
        else:
            r = w * cs / c0
            obj = float((r ** PNORM).sum() ** (1 / PNORM))
            coef = (r / obj) ** (PNORM - 1) * w / c0
        coef_dev.copy_(torch.from_numpy(np.asarray(coef, float)))
        _lib.call("anneal_fem_sensitivity", fem._ctx, _lib.ptr(U), ncase, _lib.ptr(flags), _lib.ptr(xphys),
                  _lib.ptr(dxp), _lib.ptr(phi), _lib.ptr(coef_dev), float(p), emin, float(E0), _lib.ptr(dcdx),
                  _lib.ptr(dvdx))
        filt.backward(dcdx, dcx)
        filt.backward(dvdx, dvx)
        td = tick()
        T["sensitivity"] += td - tc
        # OC (optimize.py:144-163)
        _lib.call("anneal_design_stats", None, None, None, _lib.ptr(phi), _lib.ptr(dcx), _lib.ptr(dvx),
                  _lib.ptr(flags), nel, _lib.ptr(stats), _lib.ptr(work), stream)
        st = stats.cpu().numpy()
        ndes = st[3]
        if ndes > 0:
            scale = max(float((st[4] / ndes) / max(st[5] / ndes, 1e-30)), 1e-300)
        else:
            scale = 1.0
        steps = filt.oc_bisect(x, dcx, dvx, phi, beta, move, scale * 1e-8, scale * 1e8, Vtarget, xnew)
        te = tick()
        T["oc"] += te - td
        filt.physical(xnew, beta, phi, xphys, dxp)
        _lib.call("anneal_design_stats", _lib.ptr(x), _lib.ptr(xnew), _lib.ptr(xphys), _lib.ptr(phi), None, None,
                  _lib.ptr(flags), nel, _lib.ptr(stats), _lib.ptr(work), stream)
        st = stats.cpu().numpy()
        x, xnew = xnew, x
        change = float(st[2]) if ndes > 0 else 0.0
        grey = float(st[1] / ndes) if ndes > 0 else 0.0
        volfrac = float(st[0] / Vtot)
        T["stats"] += tick() - te
        hist.append({"iter": it, "objective": obj, "compliance": cs.tolist(), "p": p, "beta": beta,
                     "volumeFraction": volfrac, "change": change, "grey": grey,
                     "cgIters": fem.cg_iters[-ncase:], "amgRebuilt": bool(rebuild),
                     "seconds": time.time() - t0, "ocSteps": steps})
        log(f"  it {it:3d}  p {p:.2f}  b {beta:.0f}  obj {obj:.4g}  vol {volfrac:.3f}  "
            f"change {change:.3f}  grey {grey:.2f}  cg {fem.cg_iters[-1]}{' amg' if rebuild else ''}  "
            f"{hist[-1]['seconds']:.1f}s")
        if p >= p_final and beta >= beta_max and it > (5 if warm else 10) and change < tol:
            log(f"  converged at iteration {it}")
            break

    full = xphys.cpu().numpy().copy()
    full[~active] = 0.0
    return {"x": full, "history": hist, "fem_builds": fem.amg_builds, "rMM": r_vox * h,
            "seconds": time.time() - t_start, "cases": cases}


def benchmark_cantilever(log=lambda *a: None):
    """V6 (optimize.py:183-210): the optimised cantilever must beat a uniform design."""
    g = Grid((32, 8, 12), 1.0, (0, 0, 0))
    fill = np.ones(g.nelem)
    active = np.ones(g.nelem, bool)
    frozen = np.zeros(g.nelem, bool)
    n = np.arange(g.nnode)
    pxyz = g.node_coords(n)
    clamp = n[pxyz[:, 0] < 1e-9]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    tip = n[(abs(pxyz[:, 0] - 32) < 1e-9) & (abs(pxyz[:, 2] - 6) < 1e-9)]
    F = np.zeros(3 * g.nnode)
    F[3 * tip + 2] = -1.0 / tip.size
    vf = 0.3
    r = simp(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1000.0, 0.3, vf, iters=40, r_mm=1.6, log=log)
    fem = FEM(g, active, fixed, 0.3)
    Fr = fem.reduce_load(F)

    def comp(xp):
        u = fem.solve(fem.assemble(1000.0 * (EMIN + xp ** 3 * (1 - EMIN))), [Fr], rebuild=True)[0]
        return float(Fr @ u)
    c_opt = comp(r["x"])
    c_uni = comp(np.full(g.nelem, vf))
    return {"stiffer_pct": 100 * (1 - c_opt / c_uni), "vol_err": r["history"][-1]["volumeFraction"] - vf}
