"""Pins the FP64 CPU oracle against closed forms, brute force and library routines.

The reference holds no numeric fixtures for this path; its own known-answer
checks are the closed forms of anneal/validate.py (V1-V6), restated here
against the oracle.  Everything else is pinned by mathematics: dense solves,
brute-force neighbourhood sums, adjointness, finite differences and schedule
tables.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle.fem_ref import CORNERS, FEMRef, Grid, hex8_matrices
from oracle.optimize_ref import (FilterRef, benchmark_cantilever_ref, cone_kernel, heaviside, oc_update,
                                 penalty_schedule)

E0, NU = 1000.0, 0.3


def _nodes_where(g, fn):
    n = np.arange(g.nnode)
    return n[fn(g.node_coords(n))]


def _face_load(g, elems, weights, total, comp, face_nodes):
    F = np.zeros(3 * g.nnode)
    en = g.element_nodes(elems)
    share = total * weights / weights.sum()
    for k in face_nodes:
        np.add.at(F, 3 * en[:, k] + comp, share / 4.0)
    return F


def _tension(fill_layer=None, nz=4, L=10, ny=2, h=1.0, F=100.0):
    """Uniaxial bar of validate.py:48-64 (own restatement)."""
    g = Grid((L, ny, nz), h, (0, 0, 0))
    e = np.arange(g.nelem)
    i, j, k = g.elem_ijk(e)
    fill = np.ones(g.nelem) if fill_layer is None else np.where(k == nz - 1, fill_layer, 1.0)
    fixed = [3 * n for n in _nodes_where(g, lambda p: p[:, 0] < 1e-9)]
    n0 = _nodes_where(g, lambda p: (p[:, 0] < 1e-9) & (p[:, 1] < 1e-9) & (p[:, 2] < 1e-9))[0]
    n1 = _nodes_where(g, lambda p: (p[:, 0] < 1e-9) & (abs(p[:, 1] - ny * h) < 1e-9) & (p[:, 2] < 1e-9))[0]
    fixed += [3 * n0 + 1, 3 * n0 + 2, 3 * n1 + 2]
    fem = FEMRef(g, np.ones(g.nelem, bool), fixed, NU)
    end = e[i == L - 1]
    Fg = _face_load(g, end, fill[end], F, 0, [1, 2, 5, 6])
    u = fem.expand(fem.solve(fem.assemble(E0 * fill), [fem.reduce_load(Fg)], rtol=1e-12)[0])
    tip = _nodes_where(g, lambda p: abs(p[:, 0] - L * h) < 1e-9)
    return u[3 * tip].mean()


def test_v1_rigid_body_energy():
    ke, _, _ = hex8_matrices(NU, 1.0)
    xyz = CORNERS.astype(float)
    t = np.tile([1.0, 0, 0], 8)
    r = np.stack([-xyz[:, 1], xyz[:, 0], np.zeros(8)], axis=1).ravel()
    assert abs(t @ ke @ t) < 1e-10 and abs(r @ ke @ r) < 1e-10


def test_ke_structure():
    ke, B0, D = hex8_matrices(NU, 1.0)
    assert np.allclose(ke, ke.T, atol=1e-15)
    assert np.linalg.matrix_rank(ke) == 18
    assert np.allclose(np.diag(ke), ke[0, 0]) and abs(ke[0, 0] - 0.2350427350427350) < 1e-12
    assert np.unique(np.round(np.abs(ke[np.abs(ke) > 1e-12]), 10)).size == 10
    ke2, _, _ = hex8_matrices(NU, 2.5)
    assert np.allclose(ke2, 2.5 * ke)             # 3-D elasticity: ke scales with h


def test_v2_tension_closed_form():
    ux = _tension()
    exact = 100.0 * 10 / (2 * 4 * 1.0 * E0)
    assert abs(ux - exact) / exact < 1e-8


def test_v4_partial_volume():
    exact = 100.0 * 10 / (2 * 6.4 * E0)
    ep = abs(_tension(fill_layer=0.4, nz=7) - exact) / exact
    eb = abs(_tension(fill_layer=None, nz=6) - exact) / exact
    assert ep < 1e-8 and eb > 0.05


@pytest.mark.parametrize("nd,lo,hi", [(2, 0.84, 0.90), (4, 0.93, 0.98)])
def test_v3_cantilever_timoshenko(nd, lo, hi):
    d, b, L, P = 4.0, 4.0, 40.0, 10.0
    h = d / nd
    g = Grid((10 * nd, nd, nd), h, (0, 0, 0))
    e = np.arange(g.nelem)
    i = g.elem_ijk(e)[0]
    clamp = _nodes_where(g, lambda p: p[:, 0] < 1e-9)
    fem = FEMRef(g, np.ones(g.nelem, bool), (3 * clamp[:, None] + np.arange(3)).ravel(), NU)
    end = e[i == g.nx - 1]
    Fg = _face_load(g, end, np.ones(end.size), -P, 2, [1, 2, 5, 6])
    u = fem.expand(fem.solve(fem.assemble(E0 * np.ones(g.nelem)), [fem.reduce_load(Fg)], rtol=1e-12)[0])
    tip = _nodes_where(g, lambda p: abs(p[:, 0] - L) < 1e-9)
    I = b * d ** 3 / 12
    G = E0 / (2 * (1 + NU))
    timo = P * L ** 3 / (3 * E0 * I) + P * L / (5 / 6 * G * b * d)
    assert lo <= -u[3 * tip + 2].mean() / timo <= hi


def test_v5_determinism():
    out = []
    for _ in range(2):
        g = Grid((20, 4, 4), 1.0, (0, 0, 0))
        clamp = _nodes_where(g, lambda p: p[:, 0] < 1e-9)
        fem = FEMRef(g, np.ones(g.nelem, bool), (3 * clamp[:, None] + np.arange(3)).ravel(), NU)
        F = np.zeros(3 * g.nnode)
        F[3 * _nodes_where(g, lambda p: abs(p[:, 0] - 20) < 1e-9) + 2] = -0.04
        out.append(fem.solve(fem.assemble(E0 * np.ones(g.nelem)), [fem.reduce_load(F)])[0])
    assert out[0].tobytes() == out[1].tobytes()


def test_assembly_equals_dense_element_sum():
    rng = np.random.default_rng(0)
    g = Grid((3, 2, 2), 1.0, (0, 0, 0))
    active = rng.random(g.nelem) > 0.2
    active[0] = True
    fixed = np.array([0, 1, 2, 7])
    fem = FEMRef(g, active, fixed, NU)
    Ee = rng.uniform(0.1, 2, fem.act.size)
    Kd = np.zeros((3 * g.nnode, 3 * g.nnode))
    for k, e in enumerate(fem.act):
        dofs = fem.edof[k]
        Kd[np.ix_(dofs, dofs)] += Ee[k] * fem.ke
    assert np.allclose(fem.assemble(Ee).toarray(), Kd[np.ix_(fem.free, fem.free)], atol=1e-14)


def test_solve_matches_dense():
    rng = np.random.default_rng(1)
    g = Grid((4, 3, 2), 1.0, (0, 0, 0))
    active = np.ones(g.nelem, bool)
    clamp = _nodes_where(g, lambda p: p[:, 0] < 1e-9)
    fem = FEMRef(g, active, (3 * clamp[:, None] + np.arange(3)).ravel(), NU)
    K = fem.assemble(10 ** rng.uniform(-3, 0, g.nelem))
    f = rng.standard_normal(fem.nfree)
    u = fem.solve(K, [f], rtol=1e-12)[0]
    assert np.allclose(u, np.linalg.solve(K.toarray(), f), rtol=1e-8, atol=1e-10 * np.abs(u).max())


def test_filter_bruteforce_and_adjoint():
    rng = np.random.default_rng(2)
    g = Grid((7, 5, 4), 1.0, (0, 0, 0))
    active = rng.random(g.nelem) > 0.25
    for r in (1.5, 3.0):
        f = FilterRef(g, active, r)
        x = rng.random(f.act.size)
        # brute force: direct neighbourhood sum normalised by the active weight sum
        R = max(int(np.ceil(r)) - 1, 1)
        xs = np.zeros(g.nelem)
        xs[f.act] = x
        want = []
        for e in f.act:
            i, j, k = g.elem_ijk(e)
            num = den = 0.0
            for dk in range(-R, R + 1):
                for dj in range(-R, R + 1):
                    for di in range(-R, R + 1):
                        a, b, c = i + di, j + dj, k + dk
                        if 0 <= a < g.nx and 0 <= b < g.ny and 0 <= c < g.nz:
                            w = max(0.0, r - np.sqrt(di * di + dj * dj + dk * dk))
                            e2 = a + g.nx * (b + g.ny * c)
                            num += w * xs[e2] * active[e2]
                            den += w * active[e2]
            want.append(num / den)
        assert np.allclose(f.forward(x), want, rtol=1e-13, atol=1e-15)
        d = rng.standard_normal(f.act.size)
        assert np.isclose(f.forward(x) @ d, x @ f.backward(d), rtol=1e-12)
        assert np.allclose(f.forward(np.ones(f.act.size)), 1.0)
        assert (cone_kernel(r) > 0).sum() == (19 if r == 1.5 else 93)


def test_projection_closed_form_and_derivative():
    x = np.linspace(0, 1, 11)
    for beta in (1.0, 4.0):
        xp, dxp = heaviside(x, beta)
        assert abs(xp[0]) < 1e-15 and abs(xp[-1] - 1) < 1e-15 and abs(xp[5] - 0.5) < 1e-15
        eps = 1e-6
        fd = (heaviside(x + eps, beta)[0] - heaviside(x - eps, beta)[0]) / (2 * eps)
        assert np.allclose(dxp, fd, rtol=1e-7)
    xp, dxp = heaviside(x, 0.0)
    assert np.array_equal(xp, x) and np.all(dxp == 1)


def test_oc_invariants():
    rng = np.random.default_rng(3)
    g = Grid((10, 6, 5), 1.0, (0, 0, 0))
    active = np.ones(g.nelem, bool)
    f = FilterRef(g, active, 1.5)
    n = f.act.size
    frz = rng.random(n) < 0.1
    x = rng.uniform(0.1, 0.8, n)
    x[frz] = 1
    dcx, dvx, phi = -rng.random(n), rng.uniform(0.5, 1, n), np.ones(n)
    vt = 0.35 * n

    def physical(xd):
        xp = f.forward(xd)
        xp[frz] = 1
        return xp
    xn, steps = oc_update(x, dcx, dvx, phi, frz, ~frz, vt, 0.2, 0.0, physical)
    assert steps == 36
    assert np.all(xn[frz] == 1)
    assert np.all(xn[~frz] <= np.minimum(1, x[~frz] + 0.2) + 1e-15)
    assert np.all(xn[~frz] >= np.maximum(0, x[~frz] - 0.2) - 1e-15)
    assert abs((phi * physical(xn)).sum() - vt) / vt < 1e-6


def test_continuation_schedules():
    # 50 iterations: hold 2 -> p reaches 3 at it 17; 100 iterations: hold 3 -> it 25
    p50 = [penalty_schedule(it, 50, 3.0, False) for it in range(1, 51)]
    assert p50.index(3.0) + 1 == 17 and p50[:4] == [1.0, 1.0, 1.25, 1.25]
    p100 = [penalty_schedule(it, 100, 3.0, False) for it in range(1, 101)]
    assert p100.index(3.0) + 1 == 25
    assert all(v == 3.0 for v in [penalty_schedule(it, 60, 3.0, True) for it in range(1, 10)])


def test_v6_simp_beats_uniform():
    r = benchmark_cantilever_ref()
    assert r["stiffer_pct"] > 20 and abs(r["vol_err"]) < 1e-3
