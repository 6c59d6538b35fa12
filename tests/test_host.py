"""CPU-side checks of the product: host math, the C ABI's exports, and loud failure without a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle.fem_ref import CORNERS, Grid as GridRef, hex8_matrices as ke_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "anneal_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(anneal_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from anneal_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert lib.anneal_abi_version() == 1


def test_library_ke_matches_oracle():
    from anneal_amd import _lib
    for nu, h in ((0.3, 1.0), (0.25, 0.5), (0.45, 2.0)):
        ke = np.zeros(576)
        coef = np.zeros(8)
        _lib.call("anneal_ke_host", nu, h, _lib.ptr(ke), _lib.ptr(coef))
        ke = ke.reshape(24, 24)
        perm = [c[0] + 2 * c[1] + 4 * c[2] for c in CORNERS]
        P = np.zeros((24, 24))
        for i, q in enumerate(perm):
            for c in range(3):
                P[3 * i + c, 3 * q + c] = 1
        ref = ke_ref(nu, h)[0]
        assert np.abs(P @ ke @ P.T - ref).max() <= 1e-15 * np.abs(ref).max() * 10


def test_library_rejects_bad_arguments():
    from anneal_amd import _lib
    ke, coef = np.zeros(576), np.zeros(8)
    with pytest.raises(_lib.AnnealLibError, match="nu"):
        _lib.call("anneal_ke_host", 0.5, 1.0, _lib.ptr(ke), _lib.ptr(coef))
    with pytest.raises(_lib.AnnealLibError, match="null"):
        _lib.call("anneal_fem_set_Ee", None, None)


def test_product_hex8_and_grid_match_oracle():
    from anneal_amd.fem import Grid, elasticity_D, hex8_matrices
    for nu, h in ((0.3, 1.0), (0.2, 1.7)):
        a, b = hex8_matrices(nu, h), ke_ref(nu, h)
        for x, y in zip(a, b):
            assert np.abs(x - y).max() <= 1e-14 * max(1.0, np.abs(y).max())
    g, r = Grid((5, 4, 3), 0.5, (1, 2, 3)), GridRef((5, 4, 3), 0.5, (1, 2, 3))
    e = np.arange(g.nelem)
    assert np.array_equal(g.element_nodes(e), r.element_nodes(e))
    assert np.allclose(g.node_coords(np.arange(g.nnode)), r.node_coords(np.arange(g.nnode)))
    v = np.arange(g.nelem)
    assert np.array_equal(g.to1d(g.to3d(v)), v) and g.to3d(v)[1, 0, 0] == 1


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from anneal_amd.fem import FEM, Grid
    g = Grid((2, 2, 2), 1.0, (0, 0, 0))
    with pytest.raises(RuntimeError, match="HIP device"):
        FEM(g, np.ones(g.nelem, bool), [0, 1, 2], 0.3)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_ctypes_signatures_cover_header_arity():
    """Each ctypes signature has as many arguments as the C declaration."""
    from anneal_amd import _lib
    text = " ".join(open(HEADER).read().split())
    for name, args in _lib.SIGNATURES.items():
        m = re.search(rf"{name}\s*\(([^)]*)\)", text)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), name
