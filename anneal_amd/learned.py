"""Learned preconditioner: a 3-D conv U-Net on the multigrid's coarse node grid.

No counterpart in the reference (anneal/fem.py solves with pyamg-preconditioned
CG).  The network (4 levels, 16/32/64/64 channels, 3x3x3 convolutions, BF16
weights and activations, FP32 accumulation) runs on the first coarse level of
the geometric hierarchy and serves as a stage of the flexible PCG
(``FEM.solve_device(..., learned=net)``); plain multigrid PCG remains the
fallback and the FP64 true residual judges every solve.  All convolutions,
their gradients, the loss and the Adam update are HIP kernels of
libanneal_hip.so (BF16 MFMA implicit GEMM); torch only holds device memory.
"""
import ctypes

import numpy as np

from . import _lib

WIDTHS = (16, 32, 64, 64)
CIN = 8     # density, residual x/y/z, fixed x/y/z, active-node flag
COUT = 3    # displacement x/y/z


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("anneal_amd needs a HIP device (MI355X); none is visible to torch")
    return torch


def layer_shapes(cin=CIN, cout=COUT, widths=WIDTHS):
    """[(taps, cin, cout, relu)] in the flat parameter order of the C ABI."""
    L = len(widths)
    out = []
    for l in range(L):
        out.append((27, cin if l == 0 else widths[l - 1], widths[l], True))
        out.append((27, widths[l], widths[l], True))
    for l in range(L - 1):
        out.append((27, widths[l + 1] + widths[l], widths[l], True))
        out.append((27, widths[l], widths[l], True))
    out.append((1, widths[0], cout, False))
    return out


def he_init(cin=CIN, cout=COUT, widths=WIDTHS, seed=0, out_scale=0.1):
    """He-normal weights (the linear output layer scaled by out_scale), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for taps, ci, co, relu in layer_shapes(cin, cout, widths):
        std = np.sqrt(2.0 / (taps * ci)) * (1.0 if relu else out_scale)
        parts.append(rng.normal(0.0, std, size=co * taps * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


class UNet:
    """Device U-Net context (anneal_unet_* entry points)."""

    def __init__(self, cin=CIN, cout=COUT, widths=WIDTHS, params=None, seed=0):
        torch = _torch()
        self.torch = torch
        self.cin, self.cout, self.widths = int(cin), int(cout), tuple(int(w) for w in widths)
        self._ctx = ctypes.c_void_p()
        w = (ctypes.c_int * len(self.widths))(*self.widths)
        _lib.call("anneal_unet_create", ctypes.byref(self._ctx), self.cin, self.cout, w, len(self.widths),
                  _lib.stream_ptr())
        n = ctypes.c_int64()
        _lib.call("anneal_unet_nparams", self._ctx, ctypes.byref(n))
        self.nparams = n.value
        self.set_params(he_init(self.cin, self.cout, self.widths, seed) if params is None else params)
        self.step_count = 0
        self._work = torch.zeros(130 * 64, dtype=torch.float64, device="cuda")
        self._loss = torch.zeros(1, dtype=torch.float64, device="cuda")

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_unet_destroy", self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.nparams, np.float32)
        _lib.call("anneal_unet_get_params", self._ctx, _lib.ptr(out))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(np.asarray(flat, np.float32).ravel())
        if flat.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {flat.size}")
        _lib.call("anneal_unet_set_params", self._ctx, _lib.ptr(flat))
        self.step_count = 0

    def grads(self):
        out = np.empty(self.nparams, np.float32)
        _lib.call("anneal_unet_get_grads", self._ctx, _lib.ptr(out))
        return out

    def forward(self, x, dims, out=None):
        """x: bf16 device tensor [B][V][cin] (V = nx ny nz nodes); returns FP32 [B][V][cout]."""
        torch = self.torch
        nx, ny, nz = (int(v) for v in dims)
        B = x.shape[0]
        if x.dtype != torch.bfloat16 or not x.is_contiguous() or x.shape[1:] != (nx * ny * nz, self.cin):
            raise ValueError("input must be a contiguous bf16 tensor [B][nx*ny*nz][cin]")
        if out is None:
            out = torch.empty(B, nx * ny * nz, self.cout, dtype=torch.float32, device=x.device)
        _lib.call("anneal_unet_forward", self._ctx, _lib.ptr(x), B, nx, ny, nz, _lib.ptr(out))
        return out

    def backward(self, dout):
        _lib.call("anneal_unet_backward", self._ctx, _lib.ptr(dout))

    def zero_grad(self):
        _lib.call("anneal_unet_zero_grad", self._ctx)

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step_count += 1
        _lib.call("anneal_unet_adam", self._ctx, lr, beta1, beta2, eps, self.step_count, weight_decay)

    def rel_mse(self, out, target, dims, dout=None):
        """Device loss mean_b |o - t|^2 / |t|^2 (FP64 scalar tensor) and, if given, its gradient into dout."""
        nx, ny, nz = (int(v) for v in dims)
        B = out.shape[0]
        _lib.call("anneal_unet_rel_mse", _lib.ptr(out), _lib.ptr(target), B, nx, ny, nz, self.cout,
                  _lib.ptr(dout) if dout is not None else None, _lib.ptr(self._loss), _lib.ptr(self._work),
                  _lib.stream_ptr())
        return self._loss

    def save(self, path):
        np.savez(path, params=self.params(), cin=self.cin, cout=self.cout, widths=np.array(self.widths))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        return cls(int(z["cin"]), int(z["cout"]), tuple(int(w) for w in z["widths"]), params=z["params"])


# ---- training data from SIMP runs ------------------------------------------------

TRAIN_SHAPE = (48, 24, 24)


def training_problem(seed, shape=TRAIN_SHAPE):
    """Seeded cantilever of the benchmark's family (the x = 0 face clamped).

    A unit point load sits on the x = nx face at a node drawn from the middle
    half of y and z, with a direction uniform on the sphere; the volume fraction
    is drawn from U(0.2, 0.4).  Returns (grid, fill, active, frozen, fixed, F, vf).
    """
    from .fem import Grid
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = shape
    g = Grid(shape, 1.0, (0, 0, 0))
    n = np.arange(g.nnode)
    clamp = n[n % (nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    j = int(rng.integers(ny // 4, 3 * ny // 4 + 1))
    k = int(rng.integers(nz // 4, 3 * nz // 4 + 1))
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    node = nx + (nx + 1) * (j + (ny + 1) * k)
    F = np.zeros(3 * g.nnode)
    F[3 * node:3 * node + 3] = d
    vf = float(rng.uniform(0.2, 0.4))
    return g, np.ones(g.nelem), np.ones(g.nelem, bool), np.zeros(g.nelem, bool), fixed, F, vf


def generate_pairs(n_runs, iters=50, shape=TRAIN_SHAPE, seed=0, level=1, r_mm=1.5, log=None, kind="residual"):
    """(X bf16 [N, V, 8], T FP32 [N, V, 3], node dims) from every solve of n_runs seeded SIMP runs.

    kind "residual" (what the solve hands the network): each solve contributes
    (density, residual r0 = F - K u0 of its starting point u0, correction
    u - u0), so the pairs follow the distribution the network sees inside the
    solver, where the start is extrapolated from the previous solutions and the
    residual is small and local; the first solve of a run starts from 0, so it
    contributes (density, load, displacement).  kind "load": every solve
    contributes (density, load, displacement).  Mapped onto multigrid level
    ``level`` by FEM.learned_pair.
    """
    torch = _torch()
    from .optimize import SimpRun
    xs, ts, dims = [], [], None

    def capture(run):
        nonlocal dims
        if kind == "residual":
            y = torch.empty_like(run.U_start)
            _lib.call("anneal_fem_apply_f64", run.fem._ctx, _lib.ptr(run.U_start), _lib.ptr(y), run.ncase)
            r0 = run.F - y          # the apply leaves padding nodes unwritten; learned_pair reads nodes only
            x, t, dims = run.fem.learned_pair(r0, run.U - run.U_start, run.ncase, level=level, e_ref=run.E0)
        else:
            x, t, dims = run.fem.learned_pair(run.F, run.U, run.ncase, level=level, e_ref=run.E0)
        xs.append(x)
        ts.append(t)

    if kind not in ("residual", "load"):
        raise ValueError(kind)
    for r in range(n_runs):
        g, fill, active, frozen, fixed, F, vf = training_problem(seed + r, shape)
        run = SimpRun(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, vf, iters=iters, r_mm=r_mm,
                      tol=0.0, emin=1e-9, rtol=1e-6, log=lambda m: None, on_solve=capture)
        run.keep_start = kind == "residual"
        for _ in range(iters):
            run.step()
        run.fem.close()
        if log and (r + 1) % max(1, n_runs // 10) == 0:
            log(f"pairs: {r + 1}/{n_runs} runs, {sum(x.shape[0] for x in xs)} samples")
    torch.cuda.synchronize()
    return torch.cat(xs), torch.cat(ts), dims


def train(net, X, T, dims, steps, batch=32, lr=2e-3, lr_final=1e-4, seed=0, val=None, log=None, log_every=100):
    """Adam on the per-sample relative squared error; cosine learning-rate decay.

    val: optional (Xv, Tv) evaluated every log_every steps.  Returns the list of
    (step, train loss, validation loss or None).
    """
    torch = _torch()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    N = X.shape[0]
    V = X.shape[1]
    dout = torch.empty(batch, V, net.cout, dtype=torch.float32, device=X.device)
    hist = []
    for s in range(1, steps + 1):
        idx = torch.randint(0, N, (batch,), generator=gen).to(X.device)
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        out = net.forward(xb, dims)
        loss = net.rel_mse(out, tb, dims, dout)
        net.zero_grad()
        net.backward(dout)
        c = 0.5 * (1.0 + np.cos(np.pi * (s - 1) / max(1, steps - 1)))
        net.adam(lr_final + (lr - lr_final) * c)
        if s % log_every == 0 or s == steps:
            tl = float(loss.item())
            vl = evaluate(net, *val, dims) if val is not None else None
            hist.append((s, tl, vl))
            if log:
                log(f"train step {s}: loss {tl:.4f}" + (f", validation {vl:.4f}" if vl is not None else ""))
    return hist


def evaluate(net, X, T, dims, batch=64):
    """Mean per-sample relative squared error over (X, T)."""
    torch = _torch()
    tot, n = 0.0, 0
    for i in range(0, X.shape[0], batch):
        xb, tb = X[i:i + batch].contiguous(), T[i:i + batch].contiguous()
        out = net.forward(xb, dims)
        tot += float(net.rel_mse(out, tb, dims).item()) * xb.shape[0]
        n += xb.shape[0]
    return tot / max(n, 1)


def main():
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser(description="Train the learned-preconditioner U-Net on SIMP solves")
    ap.add_argument("--runs", type=int, default=400)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--val-runs", type=int, default=20)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.time()
    X, T, dims = generate_pairs(a.runs, a.iters, seed=a.seed, log=log)
    Xv, Tv, _ = generate_pairs(a.val_runs, a.iters, seed=10_000 + a.seed)
    log(f"data: {X.shape[0]} training / {Xv.shape[0]} validation pairs on {dims} nodes in {time.time() - t0:.1f} s")
    net = UNet(seed=a.seed)
    t1 = time.time()
    hist = train(net, X, T, dims, a.steps, a.batch, a.lr, seed=a.seed, val=(Xv, Tv), log=log)
    log(f"trained {a.steps} steps in {time.time() - t1:.1f} s; final validation {hist[-1][2]:.4f}")
    net.save(a.out)


if __name__ == "__main__":
    main()
