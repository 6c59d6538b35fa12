"""CPU pinning of the U-Net oracle: grid transfer identities and FP64 finite-difference gradients."""
import numpy as np
import torch

from oracle import unet_ref as U


def test_restrict_prolong_are_transposes():
    rng = np.random.default_rng(0)
    for fine in [(5, 7, 9), (6, 4, 8), (3, 3, 3), (2, 5, 4)]:
        coarse = tuple(U.coarse_n(n) for n in fine)
        x = torch.tensor(rng.normal(size=(2, 3) + fine))
        y = torch.tensor(rng.normal(size=(2, 3) + coarse))
        Rx = U.restrict(x)
        assert Rx.shape == y.shape
        Py = U.prolong(y, fine)
        assert Py.shape == x.shape
        lhs, rhs = float((Rx * y).sum()), float((x * Py).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_prolong_reproduces_affine_and_restrict_keeps_constants():
    for fine in [(9, 5, 7), (8, 6, 4)]:
        coarse = tuple(U.coarse_n(n) for n in fine)
        a = np.array([0.3, -1.2, 2.0, 0.7])
        Z, Y, X = np.meshgrid(*[2.0 * np.arange(c) for c in coarse], indexing="ij")
        yc = torch.tensor((a[0] + a[1] * X + a[2] * Y + a[3] * Z)[None, None])
        z, yv, xv = np.meshgrid(*[np.arange(n, dtype=float) for n in fine], indexing="ij")
        exact = a[0] + a[1] * xv + a[2] * yv + a[3] * z
        assert np.abs(U.prolong(yc, fine)[0, 0].numpy() - exact).max() < 1e-12
        ones = torch.ones((1, 1) + fine, dtype=torch.float64)
        r = U.restrict(ones, 0.125)[0, 0].numpy()
        assert np.allclose(r[1:-1, 1:-1, 1:-1], 1.0)   # interior: the 27 weights sum to 8


def test_network_gradient_matches_finite_differences():
    cin, cout, widths = 8, 3, [16, 16, 32]
    dims, B = (5, 4, 3), 2
    rng = np.random.default_rng(1)
    flat = U.init_params(cin, cout, widths, seed=3, out_scale=1.0).astype(np.float64)
    flat += 0.05 * rng.normal(size=flat.size)            # nonzero biases as well
    V = dims[0] * dims[1] * dims[2]
    x = rng.normal(size=(B, V, cin))
    t = rng.normal(size=(B, V, cout))
    _, loss, g = U.forward_backward(flat, x, t, B, dims, cin, cout, widths, bf16=False, dtype=torch.float64)

    def L(p):
        params = U.unflatten(p, cin, cout, widths, torch.float64)
        o = U.forward(params, torch.tensor(x), B, dims, widths)
        return float(U.rel_mse(o, torch.tensor(t)))

    assert abs(L(flat) - loss) < 1e-12
    idx = rng.choice(flat.size, 24, replace=False)
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = 1e-6
        fd = (L(flat + e) - L(flat - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_bf16_emulation_stays_close_to_fp32():
    cin, cout, widths = 8, 3, [16, 32, 64, 64]
    dims, B = (9, 7, 5), 1
    rng = np.random.default_rng(2)
    flat = U.init_params(cin, cout, widths, seed=0)
    x = rng.normal(size=(B, dims[0] * dims[1] * dims[2], cin)).astype(np.float32)
    p = U.unflatten(flat, cin, cout, widths)
    o32 = U.forward(p, torch.tensor(x), B, dims, widths).numpy()
    o16 = U.forward(p, torch.tensor(x), B, dims, widths, bf16=True).numpy()
    assert np.linalg.norm(o16 - o32) <= 0.05 * np.linalg.norm(o32)
