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
