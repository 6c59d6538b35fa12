"""anneal_dot_f64_noalloc (the entry point that allocates nothing) against anneal_dot_f64.

Both run the same two kernels; the no-alloc variant keeps its partial sums in
the library's static device array instead of a caller-supplied work buffer, so
the results must be bit-identical.  Two calls in a row, the second smaller, so
the second reads the array after a larger call has filled it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dot_pair(torch, nnode, nrhs, seed):
    from anneal_amd import _lib
    g = torch.Generator(device="cpu").manual_seed(seed)
    ldn = (nnode + 63) // 64 * 64
    a = torch.randn(nrhs, 3, ldn, dtype=torch.float64, generator=g).cuda()
    b = torch.randn(nrhs, 3, ldn, dtype=torch.float64, generator=g).cuda()
    stream = _lib.c_void_p(torch.cuda.current_stream().cuda_stream)
    want = torch.zeros(nrhs, dtype=torch.float64, device="cuda")
    work = torch.empty(512 * nrhs, dtype=torch.float64, device="cuda")
    _lib.call("anneal_dot_f64", _lib.ptr(a), _lib.ptr(b), nnode, ldn, nrhs, _lib.ptr(want), _lib.ptr(work), stream)
    got = torch.full((nrhs,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.call("anneal_dot_f64_noalloc", _lib.ptr(a), _lib.ptr(b), nnode, ldn, nrhs, _lib.ptr(got), stream)
    torch.cuda.synchronize()
    ref = (a[:, :, :nnode].cpu().numpy() * b[:, :, :nnode].cpu().numpy()).sum(axis=(1, 2))
    return got.cpu().numpy(), want.cpu().numpy(), ref


def test_noalloc_dot_matches_the_allocating_entry_point():
    import torch
    assert torch.cuda.is_available()
    for nnode, nrhs, seed in ((300_001, 64, 0), (1_000, 3, 1)):
        print(f"dot noalloc nnode={nnode} nrhs={nrhs}", flush=True)
        got, want, ref = _dot_pair(torch, nnode, nrhs, seed)
        assert np.array_equal(got, want)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_noalloc_dot_rejects_more_rhs_than_its_scratch():
    import torch
    from anneal_amd import _lib
    assert torch.cuda.is_available()
    a = torch.zeros(65, 3, 64, dtype=torch.float64, device="cuda")
    out = torch.zeros(65, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.AnnealLibError, match="nrhs must be in"):
        _lib.call("anneal_dot_f64_noalloc", _lib.ptr(a), _lib.ptr(a), 64, 64, 65, _lib.ptr(out), None)
