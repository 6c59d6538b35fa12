"""BF16 MFMA U-Net kernels against the torch-CPU oracle (oracle/unet_ref.py).

The oracle rounds to bfloat16 where the device stores bfloat16, so the two
differ by FP32 summation order only (and the occasional bf16 rounding flip it
causes): forward outputs within 1e-2 relative L2, parameter gradients within
3e-2 relative L2 per layer.
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu

CIN, COUT, WIDTHS = 8, 3, (16, 32, 64, 64)


def bf16_round(a):
    return torch.tensor(a, dtype=torch.float32).to(torch.bfloat16)


def layer_slices(cin, cout, widths):
    out, pos = [], 0
    for taps, ci, co, _ in U.layer_shapes(cin, cout, widths):
        n = co * taps * ci + co
        out.append(slice(pos, pos + n))
        pos += n
    return out


@pytest.mark.parametrize("fine", [(9, 7, 5), (12, 6, 5)])
def test_grid_transfer_matches_oracle(fine):
    from anneal_amd import _lib
    rng = np.random.default_rng(sum(fine))
    B, C = 2, 16
    coarse = tuple(U.coarse_n(n) for n in fine)
    Vf, Vc = int(np.prod(fine)), int(np.prod(coarse))
    xf = bf16_round(rng.normal(size=(B, Vf, C)))
    xc = bf16_round(rng.normal(size=(B, Vc, C)))
    dxf, dxc = xf.cuda(), xc.cuda()
    yc = torch.empty(B, Vc, C, dtype=torch.bfloat16, device="cuda")
    yf = torch.empty(B, Vf, C, dtype=torch.bfloat16, device="cuda")
    _lib.call("anneal_unet_restrict", _lib.ptr(dxf), B, *fine, C, _lib.ptr(yc), 0.125, _lib.stream_ptr())
    _lib.call("anneal_unet_prolong", _lib.ptr(dxc), B, *coarse, C, _lib.ptr(yf), *fine, 1.0, 0, _lib.stream_ptr())
    rc = U.to_channels_last(U.restrict(U.to_ncdhw(xf.float(), B, fine), 0.125))
    rf = U.to_channels_last(U.prolong(U.to_ncdhw(xc.float(), B, coarse), fine[::-1]))
    # device result is the bf16 rounding of the same FP32 sums
    assert torch.allclose(yc.cpu().float(), rc.to(torch.bfloat16).float(), rtol=1e-2, atol=1e-2)
    assert torch.allclose(yf.cpu().float(), rf.to(torch.bfloat16).float(), rtol=1e-2, atol=1e-2)
    assert (yc.cpu().float() - rc).abs().max() <= 2 ** -7 * rc.abs().max()


@pytest.mark.parametrize("dims,B", [((17, 9, 9), 2), ((12, 7, 6), 1), ((65, 33, 33), 1)])
def test_forward_and_gradients_match_oracle(dims, B):
    """17x9x9 and 12x7x6 run every convolution with the K dimension split over workgroups; 65x33x33 (level 1 of
    128x64x64, the learned-preconditioner size) runs the finest levels' convolutions unsplit."""
    from anneal_amd.learned import UNet
    rng = np.random.default_rng(dims[0])
    flat = U.init_params(CIN, COUT, WIDTHS, seed=5, out_scale=1.0)
    flat = flat + (0.02 * rng.normal(size=flat.size)).astype(np.float32) * (flat == 0)   # nonzero biases
    V = int(np.prod(dims))
    x = bf16_round(rng.normal(size=(B, V, CIN)))
    t = rng.normal(size=(B, V, COUT)).astype(np.float32)
    print(f"unet fwd/bwd dims {dims} B {B}", flush=True)
    net = UNet(CIN, COUT, WIDTHS, params=flat)
    xd = x.cuda()
    out = net.forward(xd, dims)
    td = torch.tensor(t, device="cuda")
    dout = torch.empty_like(out)
    loss = net.rel_mse(out, td, dims, dout)
    net.zero_grad()
    net.backward(dout)
    g = net.grads()
    o_ref, l_ref, g_ref = U.forward_backward(flat, x.float().numpy(), t, B, dims, CIN, COUT, list(WIDTHS), bf16=True)
    o = out.cpu().numpy()
    err = np.linalg.norm(o - o_ref) / np.linalg.norm(o_ref)
    print("forward rel err", err, "loss", float(loss.item()), l_ref, flush=True)
    assert err < 1e-2
    assert abs(float(loss.item()) - l_ref) <= 1e-2 * abs(l_ref)
    errs = []
    for sl in layer_slices(CIN, COUT, WIDTHS):
        e = np.linalg.norm(g[sl] - g_ref[sl]) / max(np.linalg.norm(g_ref[sl]), 1e-30)
        errs.append(e)
    print("per-layer gradient rel err", np.round(errs, 4), flush=True)
    assert max(errs) < 3e-2
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-2


def test_unet_is_deterministic_and_trains():
    from anneal_amd.learned import UNet
    rng = np.random.default_rng(9)
    dims, B = (13, 7, 7), 4
    V = int(np.prod(dims))
    # a smooth target that depends on the input: the network must learn a local linear map
    x = bf16_round(rng.normal(size=(B, V, CIN))).cuda()
    xf = x.float()
    t = torch.stack([xf[..., 1] + 0.5 * xf[..., 4], xf[..., 2] - xf[..., 0], 0.3 * xf[..., 3]], -1).contiguous()
    net = UNet(seed=1)
    dout = torch.empty(B, V, COUT, device="cuda")
    outs = []
    for rep in range(2):
        o = net.forward(x, dims).clone()
        net.rel_mse(o, t, dims, dout)
        net.zero_grad()
        net.backward(dout)
        outs.append((o.cpu(), net.grads()))
    assert torch.equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    first = None
    for it in range(300):
        o = net.forward(x, dims)
        loss = float(net.rel_mse(o, t, dims, dout).item())
        first = loss if first is None else first
        net.zero_grad()
        net.backward(dout)
        net.adam(2e-3)
    print("training loss", first, "->", loss, flush=True)
    assert loss < 0.1 * first
