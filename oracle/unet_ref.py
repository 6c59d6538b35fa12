"""CPU oracle of the learned-preconditioner U-Net (test infrastructure, never product code).

The network has no counterpart in the reference (it is a build-only
component); its oracle is torch-CPU ``conv3d`` in FP32/FP64 on the same
architecture and parameter layout as the HIP implementation:

* node grids, channels-last [B][V][C] with voxel v = i + nx (j + ny k);
* per level l: conv a (3x3x3, ReLU), conv b (3x3x3, ReLU) -> s_l; the next
  level's input is restrict(s_l) / 8; decoder level l: concat[prolong(d_{l+1}),
  s_l] -> conv a (ReLU) -> conv b (ReLU) -> d_l; output = 1x1x1 conv of d_0;
* node-aligned transfer: coarse node I sits on fine node 2I, coarse size
  n // 2 + 1; restrict sums the 27 fine neighbours with weights
  prod(1 - |d|/2) (out-of-range nodes skipped), prolong is its transpose
  (trilinear interpolation);
* flat parameters: per layer W[co][tap][ci] (tap = dx+1 + 3(dy+1) + 9(dz+1))
  then b[co]; layers in the order enc (a, b) per level, dec (a, b) per level
  from the finest, output.

With ``bf16=True`` the oracle rounds to bfloat16 exactly where the device
stores bfloat16 (weights, every activation and every activation gradient) and
keeps FP32 elsewhere, so device and oracle differ only by summation order.

Pinning: restrict and prolong are checked to be transposes (<R x, y> =
<x, P y>), prolong to reproduce affine fields exactly and restrict/8 to keep
constants in the interior; the network gradient (autograd through these
operations) is checked against central finite differences in FP64.
"""
import numpy as np
import torch
import torch.nn.functional as F


def coarse_n(n):
    return n // 2 + 1


def layer_shapes(cin, cout, widths):
    """[(taps, cin, cout, relu)] in flat parameter order."""
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


def nparams(cin, cout, widths):
    return sum(co * t * ci + co for t, ci, co, _ in layer_shapes(cin, cout, widths))


def init_params(cin, cout, widths, seed=0, out_scale=0.1):
    """He-normal weights, zero biases (float32 flat vector)."""
    rng = np.random.default_rng(seed)
    parts = []
    for taps, ci, co, relu in layer_shapes(cin, cout, widths):
        std = np.sqrt(2.0 / (taps * ci)) * (1.0 if relu else out_scale)
        parts.append(rng.normal(0.0, std, size=co * taps * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def unflatten(flat, cin, cout, widths, dtype=torch.float32):
    """Flat vector -> list of (weight [co][ci][3][3][3] or [co][ci][1][1][1], bias)."""
    flat = torch.as_tensor(np.asarray(flat), dtype=dtype)
    res, pos = [], 0
    for taps, ci, co, _ in layer_shapes(cin, cout, widths):
        n = co * taps * ci
        W = flat[pos:pos + n].reshape(co, taps, ci)
        pos += n
        b = flat[pos:pos + co]
        pos += co
        k = 3 if taps == 27 else 1
        # tap = dx+1 + 3(dy+1) + 9(dz+1)  ->  torch kernel index [kz][ky][kx]
        Wt = W.permute(0, 2, 1).reshape(co, ci, k, k, k)
        res.append((Wt, b))
    return res


def flatten_grads(params, cin, cout, widths):
    parts = []
    for (Wt, b), (taps, ci, co, _) in zip(params, layer_shapes(cin, cout, widths)):
        parts.append(Wt.grad.reshape(co, ci, taps).permute(0, 2, 1).reshape(-1))
        parts.append(b.grad.reshape(-1))
    return torch.cat(parts).detach().numpy()


class _RoundBF16(torch.autograd.Function):
    """Round to bfloat16 (RNE) in the forward pass and the gradient in the backward pass."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


class _RoundGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def _tent(C, dtype):
    w1 = torch.tensor([0.5, 1.0, 0.5], dtype=dtype)
    k = w1[:, None, None] * w1[None, :, None] * w1[None, None, :]
    return k.expand(C, 1, 3, 3, 3).contiguous()


def restrict(x, s=1.0):
    """x [B][C][nz][ny][nx] -> [B][C][nz//2+1][ny//2+1][nx//2+1]: s * sum_d w(d) x[2I+d]."""
    C = x.shape[1]
    xp = F.pad(x, (0, 1, 0, 1, 0, 1))
    return s * F.conv3d(xp, _tent(C, x.dtype), stride=2, padding=1, groups=C)


def prolong(x, fine_shape, s=1.0):
    """Transpose of restrict: x [B][C][cz][cy][cx] -> [B][C][fz][fy][fx]."""
    C = x.shape[1]
    y = F.conv_transpose3d(x, _tent(C, x.dtype), stride=2, padding=1, groups=C)
    fz, fy, fx = fine_shape
    y = F.pad(y, (0, max(0, fx - y.shape[4]), 0, max(0, fy - y.shape[3]), 0, max(0, fz - y.shape[2])))
    return s * y[:, :, :fz, :fy, :fx]


def to_ncdhw(x_cl, B, dims):
    nx, ny, nz = dims
    C = x_cl.shape[-1]
    return x_cl.reshape(B, nz, ny, nx, C).permute(0, 4, 1, 2, 3)


def to_channels_last(x):
    B, C = x.shape[:2]
    return x.permute(0, 2, 3, 4, 1).reshape(B, -1, C)


def forward(params, x_cl, B, dims, widths, bf16=False):
    """x_cl [B][V][cin] -> out [B][V][cout] (same dtype as the parameters)."""
    rnd = _RoundBF16.apply if bf16 else (lambda t: t)
    L = len(widths)
    wt = [((_RoundBF16.apply(W) if bf16 else W), b) for W, b in params]
    x = to_ncdhw(x_cl, B, dims)
    if bf16:
        x = _RoundBF16.apply(x)
    conv = lambda t, i, relu: (F.relu if relu else (lambda u: u))(
        F.conv3d(t, wt[i][0], wt[i][1], padding=wt[i][0].shape[-1] // 2))
    skips = []
    h = x
    for l in range(L):
        a = rnd(conv(h, 2 * l, True))
        s = rnd(conv(a, 2 * l + 1, True))
        skips.append(s)
        if l + 1 < L:
            h = rnd(restrict(s, 0.125))
    d = skips[L - 1]
    base = 2 * L
    for l in range(L - 2, -1, -1):
        up = rnd(prolong(d, skips[l].shape[2:]))
        cat = torch.cat([up, skips[l]], dim=1)
        e = rnd(conv(cat, base + 2 * l, True))
        d = rnd(conv(e, base + 2 * l + 1, True))
    out = conv(d, len(params) - 1, False)
    if bf16:
        out = _RoundGrad.apply(out)
    return to_channels_last(out)


def rel_mse(o, t):
    """mean over samples of sum (o - t)^2 / sum t^2 (o, t: [B][V][c])."""
    e = ((o - t) ** 2).sum(dim=(1, 2))
    s = (t ** 2).sum(dim=(1, 2))
    return (e / s).mean()


def forward_backward(flat, x_cl, target, B, dims, cin, cout, widths, bf16=False, dtype=torch.float32):
    """Returns (out [B][V][cout] numpy, loss, flat gradient numpy)."""
    params = unflatten(flat, cin, cout, widths, dtype)
    for W, b in params:
        W.requires_grad_(True)
        b.requires_grad_(True)
    x = torch.as_tensor(np.asarray(x_cl), dtype=dtype)
    t = torch.as_tensor(np.asarray(target), dtype=dtype)
    out = forward(params, x, B, dims, widths, bf16)
    loss = rel_mse(out, t)
    loss.backward()
    return out.detach().numpy(), loss.item(), flatten_grads(params, cin, cout, widths)
