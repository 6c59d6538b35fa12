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
MAX_BLOCK = 64          # right-hand sides per device solve (anneal_fem_create's nrhs_max limit)
STATS_WORK = 3072       # doubles of scratch for anneal_design_stats
DOT_WORK_PER_RHS = 512  # doubles of scratch per right-hand side for anneal_dot_f64


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

    def backward_pair(self, d1, out1, d2, out2, stats=None):
        """backward(d1) and backward(d2) in one launch (same bits); ``stats`` (6 doubles on the device, may be
        None) receives the design count, sum(-out1) and sum(out2) over design voxels at [3], [4], [5]."""
        _lib.call("anneal_filter_backward_pair", self._ctx, _lib.ptr(d1), _lib.ptr(out1), _lib.ptr(d2),
                  _lib.ptr(out2), _lib.ptr(stats) if stats is not None else None)

    def physical(self, x, beta, phi, xphys, dxp, vol=None):
        _lib.call("anneal_filter_physical", self._ctx, _lib.ptr(x), float(beta), _lib.ptr(phi), _lib.ptr(xphys),
                  _lib.ptr(dxp), _lib.ptr(vol))

    def oc_bisect(self, x, dcx, dvx, phi, beta, move, lo, hi, vtarget, xnew, hint=0.0, info=None, width=0.0):
        """Replayed OC bisection; returns the number of trials.  ``info`` (a dict) gets evals and lam_last.
        ``hint``: guess of the root (the first multiplier evaluated); ``width``: log step of the second
        evaluation when the context has no slope from an earlier bisection (<= 0: 0.005)."""
        steps, evals, lam = _lib.ctypes.c_int(), _lib.ctypes.c_int(), _lib.c_double()
        _lib.call("anneal_oc_bisect", self._ctx, _lib.ptr(x), _lib.ptr(dcx), _lib.ptr(dvx), _lib.ptr(phi),
                  float(beta), float(move), float(lo), float(hi), float(vtarget), float(hint), float(width),
                  _lib.ptr(xnew),
                  _lib.ctypes.byref(steps), _lib.ctypes.byref(evals), _lib.ctypes.byref(lam))
        if info is not None:
            info["evals"], info["lam_last"] = evals.value, lam.value
        return steps.value

    def oc_begin(self, x, dcx, dvx, phi, beta, move, lo, hi, vtarget, xnew, hint=0.0, width=0.0):
        """First part of :meth:`oc_bisect`: queue the first evaluations and return."""
        _lib.call("anneal_oc_begin", self._ctx, _lib.ptr(x), _lib.ptr(dcx), _lib.ptr(dvx), _lib.ptr(phi),
                  float(beta), float(move), float(lo), float(hi), float(vtarget), float(hint), float(width),
                  _lib.ptr(xnew))

    def oc_begin_stats(self, x, dcx, dvx, phi, beta, move, stats, vtarget, xnew, ratio=0.0, width=0.0):
        """:meth:`oc_begin` with the bracket formed on the device from the design statistics ``stats``
        (optimize.py:146-148); the first evaluation at ``ratio`` times the bracket scale."""
        _lib.call("anneal_oc_begin_stats", self._ctx, _lib.ptr(x), _lib.ptr(dcx), _lib.ptr(dvx), _lib.ptr(phi),
                  float(beta), float(move), _lib.ptr(stats), float(ratio), float(vtarget), float(width),
                  _lib.ptr(xnew))

    def oc_poll(self):
        """Wait for the last queued evaluations; True once the bisection has stopped (else queue more)."""
        done = _lib.ctypes.c_int()
        _lib.call("anneal_oc_poll", self._ctx, _lib.ctypes.byref(done))
        return bool(done.value)

    def oc_end(self, info=None):
        """Queue the final design; returns the number of trials (as :meth:`oc_bisect`)."""
        steps, evals, lam = _lib.ctypes.c_int(), _lib.ctypes.c_int(), _lib.c_double()
        _lib.call("anneal_oc_end", self._ctx, _lib.ctypes.byref(steps), _lib.ctypes.byref(evals),
                  _lib.ctypes.byref(lam))
        if info is not None:
            info["evals"], info["lam_last"] = evals.value, lam.value
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


class SimpRun:
    """One SIMP optimisation (optimize.py:50-180) with device-resident state, stepped one iteration at a time.

    ``simp()`` is ``SimpRun(...).run()``; benchmarks call :meth:`step` to time
    individual iterations.  Arguments are those of ``simp``.
    """

    def __init__(self, grid, fill, active, frozen, fixed, loads, weights, E0, nu, vf,
                 iters=80, r_mm=None, p_final=3.0, move=0.2, combine="sum", log=print, tol=0.01, x0=None,
                 beta_max=BETA_MAX, *, emin=None, rtol=1e-6, maxiter=2000, fem=None, timers=None, learned=None,
                 learned_mode=2, learned_level=1, on_solve=None, extrapolate=True):
        torch = _torch()
        self.t_start = time.time()
        self.torch = torch
        self.grid = grid
        self.emin = EMIN if emin is None else float(emin)
        self.E0, self.vf, self.iters, self.p_final, self.move = float(E0), vf, iters, p_final, move
        self.combine, self.log, self.tol, self.beta_max = combine, log, tol, beta_max
        self.rtol, self.maxiter = rtol, maxiter
        fill = np.asarray(fill, float).ravel()
        active = np.asarray(active).astype(bool).ravel()
        frozen = np.asarray(frozen).astype(bool).ravel()
        self.active = active
        self.cases = list(loads)
        self.ncase = len(self.cases)
        if fem is None:
            # one solve takes at most 64 right-hand sides; more load cases are solved in blocks
            fem = FEM(grid, active, fixed, nu, log=log, nrhs_max=min(MAX_BLOCK, max(1, self.ncase)))
        self.fem = fem
        if learned is not None:
            fem.set_learned(learned, level=learned_level, e_ref=float(E0), mode=learned_mode)
        self.on_solve = on_solve          # called as on_solve(run) after every solve (e.g. to collect training pairs)
        act = fem.act
        dev = fem.device
        phi_h = fill[act]
        frz_h = frozen[act]
        des_h = ~frz_h
        h = grid.h
        self.r_vox = max((r_mm or 1.5 * h) / h, 1.5)
        self.filt = DeviceFilter(grid, active, self.r_vox, frozen)
        Fr = [fem.reduce_load(loads[c]) for c in self.cases]
        self.w = np.array([weights[c] for c in self.cases], dtype=float)
        if all(np.allclose(f, 0) for f in Fr):
            raise ValueError("All loads are zero after mapping onto the grid. Check the load anchors' regions.")
        self.Vtot = phi_h.sum()
        self.Vtarget = vf * self.Vtot
        Vfrozen = phi_h[frz_h].sum()
        x_h = np.ones(act.size)
        x_h[des_h] = np.clip((self.Vtarget - Vfrozen) / max(phi_h[des_h].sum(), 1e-12), 0.01, 1.0)
        self.warm = x0 is not None
        if self.warm:       # optimize.py:72-77
            xs = np.clip(np.asarray(x0, float).ravel()[act], 0.0, 1.0)
            cur = (phi_h[des_h] * xs[des_h]).sum()
            if cur > 0:
                x_h[des_h] = np.clip(xs[des_h] * (self.Vtarget - Vfrozen) / cur, 0.001, 1.0)
        nel = self.nel = grid.nelem

        def full_dev(v_act):
            f = np.zeros(nel)
            f[act] = v_act
            return torch.from_numpy(f).to(dev)

        f64 = dict(dtype=torch.float64, device=dev)
        self.x = full_dev(x_h)
        self.phi = full_dev(phi_h)
        self.xphys, self.dxp = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
        self.Ee = torch.zeros(nel, **f64)
        self.dcdx, self.dvdx = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
        self.dcx, self.dvx = torch.zeros(nel, **f64), torch.zeros(nel, **f64)
        self.xnew = torch.zeros(nel, **f64)
        self.stats = torch.zeros(8, **f64)
        # scratch of anneal_design_stats (3072 doubles) and anneal_dot_f64 (512 per load case)
        self.work = torch.zeros(max(STATS_WORK, DOT_WORK_PER_RHS * self.ncase), **f64)
        self.F = fem._to_device(Fr)
        self.U = torch.zeros_like(self.F)
        self.extrapolate = bool(extrapolate)
        self._u_prev = None
        self.cs_dev = torch.zeros(self.ncase, **f64)
        self.coef_dev = torch.zeros(self.ncase, **f64)
        # the weighted sum (and a single case) combines the sensitivities with the constant
        # weights: they are uploaded once, and the compliances are read with the end-of-step
        # statistics instead of in a host wait of their own
        self._fixed_coef = combine == "sum" or self.ncase == 1
        if self._fixed_coef:
            self.coef_dev.copy_(torch.from_numpy(self.w.copy()))
        self._torch_stream = torch.cuda.current_stream()
        self.stream = _lib.c_void_p(self._torch_stream.cuda_stream)
        # pinned read-back buffers of the per-iteration statistics and compliances
        self.stats_host = torch.zeros(8, dtype=torch.float64, pin_memory=True)
        self.cs_host = torch.zeros(self.ncase, dtype=torch.float64, pin_memory=True)
        self.timing = timers is not None      # per-phase timing synchronises the device five times per step
        self.T = timers if timers is not None else {}
        for k in ("physical", "solve", "sensitivity", "oc", "stats"):
            self.T.setdefault(k, 0.0)
        self.beta = beta_max / 2 if self.warm else 0.0      # optimize.py:79, 100-101
        self.hist = []
        self.c0 = None
        self.base_iters = None
        self.have_u = False
        self.filt.physical(self.x, self.beta, self.phi, self.xphys, self.dxp)
        # (design buffer, beta) that xphys / dxp currently hold: an iteration whose beta did not change
        # starts from the physical density the previous OC step already formed
        self._phys_key = (self.x.data_ptr(), self.beta)
        self.beta_step = max(5, int(0.12 * iters))
        self.beta_since = 0
        self.it = 0
        self.converged = False
        self.oc_ratio = 0.0
        self.oc_ratio_prev = 0.0
        self.oc_extrapolate = True
        self.oc_evals = 0
        self.keep_start = False   # True: keep each solve's starting point in U_start (training pairs)
        self._queued = None       # ((it, p, beta, beta_since), t0) of an operator queued ahead
        self._ahead = False
        self.U_start = None

    def _tick(self):
        if self.timing:
            self.torch.cuda.synchronize()
        return time.perf_counter()

    def step(self, ahead=False):
        """One SIMP iteration (optimize.py:105-175); returns its history entry.

        ``ahead``: queue the next iteration's operator (stiffness law and multigrid set-up) before this
        iteration's last host wait when the next iteration is certain to follow; the FEM operator then
        belongs to that next iteration when step returns (``run`` steps this way)."""
        self._ahead = ahead
        self._pre_solve()
        its = []
        nb = self.fem.nrhs_max
        for s in range(0, self.ncase, nb):          # load cases in blocks of at most nrhs_max
            e = min(self.ncase, s + nb)
            its += self.fem.solve_device(self.F[s:e], self.U[s:e], e - s, self.rtol, self.maxiter,
                                         use_x0=self.have_u)
        return self._post_solve(its)

    def _advance(self):
        """The next iteration's number, penalty and projection sharpness (optimize.py:105-116), not applied."""
        it = self.it + 1
        p = _schedule_p(it, self.iters, self.p_final, self.warm)
        beta, since = self.beta, self.beta_since
        if p >= self.p_final and not self.warm and self.beta_max > 0:        # optimize.py:111-116
            if beta == 0:
                beta, since = 1.0, 0
            elif since >= self.beta_step and beta < self.beta_max:
                beta, since = min(self.beta_max, 2 * beta), 0
            since += 1
        return it, p, beta, since

    def _operator(self, x, p, beta):
        """Physical density (unless xphys already holds it), stiffness law and operator (optimize.py:117-118)."""
        filt, fem = self.filt, self.fem
        if self._phys_key != (x.data_ptr(), beta):
            filt.physical(x, beta, self.phi, self.xphys, self.dxp)
            self._phys_key = (x.data_ptr(), beta)
        _lib.call("anneal_stiffness_law", _lib.ptr(self.xphys), _lib.ptr(self.phi), _lib.ptr(filt.flags), self.E0,
                  self.emin, float(p), self.nel, _lib.ptr(self.Ee), self.stream)
        fem.assemble_device(self.Ee)

    def _pre_solve(self):
        """Continuation, physical density, stiffness law and the rebuild decision (optimize.py:105-124)."""
        torch, fem, T = self.torch, self.fem, self.T
        queued, self._queued = self._queued, None
        if queued is not None:      # the operator was queued before the previous iteration's last wait
            (it, p, beta, since), self._t0 = queued
        else:
            it, p, beta, since = self._advance()
            self._t0 = time.time()
        self.it, self._p, self.beta, self.beta_since = it, p, beta, since
        ta = self._tick()
        if queued is None:
            self._operator(self.x, p, beta)
        tb = self._tick()
        T["physical"] += tb - ta
        last = fem.cg_iters[-1] if fem.cg_iters else None
        self._rebuild = self.base_iters is None or (last is not None and last > 1.6 * self.base_iters)
        if self._rebuild:
            fem.rebuild()
        self._extrapolate()
        if self.keep_start:
            self.U_start = self.U.clone() if self.have_u else torch.zeros_like(self.U)
        self._tb = tb

    def _extrapolate(self):
        """Start the solve from 2 u_{k-1} - u_{k-2} instead of u_{k-1}.

        The displacement follows the design smoothly from one SIMP iteration to
        the next, so the linear extrapolation of the last two solutions lies
        closer to the new solution (on the 192x96x96 cantilever schedule it cuts
        the PCG iterations by a fifth).  The solve still runs to the same FP64
        true-residual tolerance, so only the starting point changes."""
        if not self.extrapolate or not self.have_u:
            return
        if self._u_prev is None:          # second solve: remember u_1, start from it
            self._u_prev = self.U.clone()
            return
        # u_{k-2} <- u_{k-2} + 2 (u_{k-1} - u_{k-2}) = 2 u_{k-1} - u_{k-2} in one pass, then the
        # buffers swap roles: U is the start of this solve, _u_prev keeps u_{k-1}
        self._u_prev.lerp_(self.U, 2.0)
        self.U, self._u_prev = self._u_prev, self.U

    def _post_solve(self, its):
        """Compliance, sensitivities, OC update and the history entry (optimize.py:125-175)."""
        self._sensitivities(its)
        self._oc_begin()
        while not self.filt.oc_poll():
            pass
        self._oc_end()
        nxt = self._next_operator_ahead()
        if nxt is None:
            self._sync()
        else:
            # the next iteration's operator (its design and sharpness are known now) runs on the device
            # while the host waits for the statistics and writes the history entry
            done = self.torch.cuda.Event()
            done.record(self._torch_stream)
            t0 = time.time()
            self._operator(self.xnew, nxt[1], nxt[2])
            self._queued = (nxt, t0)
            done.synchronize()
        return self._record()

    def _next_operator_ahead(self):
        """The next iteration's schedule when its operator can be queued before this iteration's last
        wait: a next iteration follows for certain (the schedule is not done and the convergence test
        of optimize.py:173-175 cannot stop the loop here) and its projection sharpness is unchanged,
        so the physical density is already in place (nothing that result() returns is touched)."""
        if not self._ahead or self.timing or self.it >= self.iters:
            return None
        if (self.tol is not None and self.tol > 0 and self._p >= self.p_final and self.beta >= self.beta_max
                and self.it > (5 if self.warm else 10)):
            return None
        nxt = self._advance()
        return nxt if nxt[2] == self.beta else None

    # _post_solve in phases that only queue device work, separated by host waits (``_sync``,
    # ``filt.oc_poll``), so that a batch of runs can share each wait

    def _sync(self):
        self._torch_stream.synchronize()

    def _sensitivities(self, its):
        """Compliance, sensitivities and the statistics that scale the OC bracket (optimize.py:125-145)."""
        torch, fem, filt, T = self.torch, self.fem, self.filt, self.T
        nel, flags, stream = self.nel, filt.flags, self.stream
        fem.cg_iters.extend(its)
        self.have_u = True
        if self._rebuild:
            self.base_iters = fem.cg_iters[-1]
        if self.on_solve is not None:
            self.on_solve(self)
        assert self.work.numel() >= DOT_WORK_PER_RHS * self.ncase
        _lib.call("anneal_dot_f64", _lib.ptr(self.F), _lib.ptr(self.U), self.grid.nnode, fem.ldn, self.ncase,
                  _lib.ptr(self.cs_dev), _lib.ptr(self.work), stream)
        tc = self._tick()
        T["solve"] += tc - self._tb
        w = self.w
        if not self._fixed_coef:                                               # optimize.py:132-139
            cs = self.cs_dev.cpu().numpy()
            if self.c0 is None:
                self.c0 = cs.copy()
            r = w * cs / self.c0
            self._obj = float((r ** PNORM).sum() ** (1 / PNORM))
            self._cs = cs
            coef = (r / self._obj) ** (PNORM - 1) * w / self.c0
            self.coef_dev.copy_(torch.from_numpy(np.asarray(coef, float)))
        _lib.call("anneal_fem_sensitivity", fem._ctx, _lib.ptr(self.U), self.ncase, _lib.ptr(flags),
                  _lib.ptr(self.xphys), _lib.ptr(self.dxp), _lib.ptr(self.phi), _lib.ptr(self.coef_dev),
                  float(self._p), self.emin, self.E0, _lib.ptr(self.dcdx), _lib.ptr(self.dvdx))
        # both backward passes and the statistics that scale the OC bracket in one launch
        filt.backward_pair(self.dcdx, self.dcx, self.dvdx, self.dvx, self.stats)
        self._td = self._tick()
        T["sensitivity"] += self._td - tc
        self.stats_host.copy_(self.stats, non_blocking=True)

    def _oc_begin(self):
        """The first OC trials, bracketed on the device from the statistics of :meth:`_sensitivities`
        (no host wait): lo, hi = 1e-8, 1e8 times scale = mean(-dcx) / mean(dvx) over design voxels."""
        # where to look for the root: the previous roots relative to their scales, extrapolated
        # geometrically from the last two (the multiplier drifts smoothly); only the choice of
        # multipliers to evaluate depends on it, never the trial decisions or the result
        ratio = self.oc_ratio
        if self.oc_extrapolate and self.oc_ratio and self.oc_ratio_prev:
            ratio = self.oc_ratio * (self.oc_ratio / self.oc_ratio_prev)
        self.filt.oc_begin_stats(self.x, self.dcx, self.dvx, self.phi, self.beta, self.move, self.stats,
                                 self.Vtarget, self.xnew, ratio=ratio or 0.0)

    def _oc_end(self):
        """The OC design (optimize.py:144-163), its physical density and the statistics of the history entry."""
        filt, flags, nel, stream = self.filt, self.filt.flags, self.nel, self.stream
        info = self._oc_info = {}
        self._oc_steps = filt.oc_end(info)
        # the statistics that scaled the bracket have reached the host before the bisection's
        # state read-back (same stream); the next read-back below overwrites them
        st = self.stats_host.numpy()
        ndes = self._ndes = st[3]
        self._scale = max(float((st[4] / ndes) / max(st[5] / ndes, 1e-30)), 1e-300) if ndes > 0 else 1.0
        self.oc_ratio_prev, self.oc_ratio = self.oc_ratio, info["lam_last"] / self._scale
        self.oc_evals += info["evals"]
        self._te = self._tick()
        self.T["oc"] += self._te - self._td
        filt.physical(self.xnew, self.beta, self.phi, self.xphys, self.dxp)
        self._phys_key = (self.xnew.data_ptr(), self.beta)
        _lib.call("anneal_design_stats", _lib.ptr(self.x), _lib.ptr(self.xnew), _lib.ptr(self.xphys),
                  _lib.ptr(self.phi), None, None, _lib.ptr(flags), nel, _lib.ptr(self.stats), _lib.ptr(self.work),
                  stream)
        self.stats_host.copy_(self.stats, non_blocking=True)
        if self._fixed_coef:
            self.cs_host.copy_(self.cs_dev, non_blocking=True)

    def _record(self):
        """History entry and convergence test (optimize.py:162-175), from the statistics read after a wait."""
        fem, T = self.fem, self.T
        it, p, beta, rebuild, ndes = self.it, self._p, self.beta, self._rebuild, self._ndes
        st = self.stats_host.numpy()
        if self._fixed_coef:                                                   # optimize.py:132-134
            cs = self.cs_host.numpy().copy()
            if self.c0 is None:
                self.c0 = cs.copy()
            obj = float(self.w @ cs)
        else:
            cs, obj = self._cs, self._obj
        self.x, self.xnew = self.xnew, self.x
        change = float(st[2]) if ndes > 0 else 0.0
        grey = float(st[1] / ndes) if ndes > 0 else 0.0
        volfrac = float(st[0] / self.Vtot)
        T["stats"] += self._tick() - self._te
        ncase = self.ncase
        h = {"iter": it, "objective": obj, "compliance": cs.tolist(), "p": p, "beta": beta,
             "volumeFraction": volfrac, "change": change, "grey": grey,
             "cgIters": fem.cg_iters[-ncase:], "amgRebuilt": bool(rebuild),
             "seconds": time.time() - self._t0, "ocSteps": self._oc_steps, "ocEvals": self._oc_info["evals"]}
        self.hist.append(h)
        self.log(f"  it {it:3d}  p {p:.2f}  b {beta:.0f}  obj {obj:.4g}  vol {volfrac:.3f}  "
                 f"change {change:.3f}  grey {grey:.2f}  cg {fem.cg_iters[-1]}{' amg' if rebuild else ''}  "
                 f"{h['seconds']:.1f}s")
        if p >= self.p_final and beta >= self.beta_max and it > (5 if self.warm else 10) and change < self.tol:
            self.log(f"  converged at iteration {it}")
            self.converged = True
        return h

    def result(self):
        full = self.xphys.cpu().numpy().copy()
        full[~self.active] = 0.0
        return {"x": full, "history": self.hist, "fem_builds": self.fem.amg_builds, "rMM": self.r_vox * self.grid.h,
                "seconds": time.time() - self.t_start, "cases": self.cases}

    def run(self):
        while self.it < self.iters and not self.converged:
            self.step(ahead=True)
        return self.result()


def simp(grid, fill, active, frozen, fixed, loads, weights, E0, nu, vf,
         iters=80, r_mm=None, p_final=3.0, move=0.2, combine="sum", log=print, tol=0.01, x0=None,
         beta_max=BETA_MAX, *, emin=None, rtol=1e-6, maxiter=2000, fem=None, timers=None, learned=None,
         learned_mode=2, learned_level=1):
    """SIMP loop of optimize.py:50-180 on the GPU; returns the reference's result dict.

    Extra keyword-only arguments (defaults reproduce the reference): ``emin``
    (module ``EMIN`` when None, optimize.py:18), ``rtol`` of the inner solves
    (optimize.py:122 uses 1e-6), ``maxiter`` (fem.py:206), an existing ``fem``
    context to reuse, ``timers`` (a dict that receives per-phase seconds) and
    ``learned`` (an anneal_amd.learned.UNet used as a flexible-PCG stage with
    ``learned_mode`` / ``learned_level``, see FEM.set_learned).
    """
    return SimpRun(grid, fill, active, frozen, fixed, loads, weights, E0, nu, vf, iters=iters, r_mm=r_mm,
                   p_final=p_final, move=move, combine=combine, log=log, tol=tol, x0=x0, beta_max=beta_max,
                   emin=emin, rtol=rtol, maxiter=maxiter, fem=fem, timers=timers, learned=learned,
                   learned_mode=learned_mode, learned_level=learned_level).run()


def benchmark_cantilever(log=lambda *a: None):
    """V6 of the reference's validation (optimize.py:183-210).

    A 32 x 8 x 12 voxel cantilever, clamped on its x = 0 face and pulled down by
    a unit load spread over the tip line (x = 32, z = 6), is optimised at 30 %
    volume (E0 = 1000, 40 iterations, 1.6 mm filter).  Its compliance at
    penalty 3 is compared with that of a uniform design of the same mass.
    """
    g = Grid((32, 8, 12), 1.0, (0, 0, 0))
    xyz = g.node_coords(np.arange(g.nnode))
    wall = np.flatnonzero(xyz[:, 0] < 1e-9)
    tip_line = np.flatnonzero((np.abs(xyz[:, 0] - 32) < 1e-9) & (np.abs(xyz[:, 2] - 6) < 1e-9))
    fixed = np.repeat(3 * wall, 3) + np.tile(np.arange(3), wall.size)
    load = np.zeros(3 * g.nnode)
    load[3 * tip_line + 2] = -1.0 / tip_line.size
    solid = np.ones(g.nelem, bool)
    E0, vf = 1000.0, 0.3
    res = simp(g, np.ones(g.nelem), solid, ~solid, fixed, {"c": load}, {"c": 1.0}, E0, 0.3, vf, iters=40, r_mm=1.6,
               log=log)
    fem = FEM(g, solid, fixed, 0.3)
    f = fem.reduce_load(load)

    def compliance(density):
        K = fem.assemble(E0 * (EMIN + density ** 3 * (1 - EMIN)))
        return float(f @ fem.solve(K, [f], rebuild=True)[0])

    c_design, c_uniform = compliance(res["x"]), compliance(np.full(g.nelem, vf))
    return {"stiffer_pct": 100 * (1 - c_design / c_uniform), "vol_err": res["history"][-1]["volumeFraction"] - vf}
