"""Pins the oracle's element energy, sensitivity chain and stresses by mathematics.

The GPU kernels for element energy, sensitivities and stresses are compared
with the oracle elsewhere; these tests make sure the oracle itself is right,
without restating its formulas:

* the energy identity sum_e Ee_e ce_e = u^T K u (K from the CSR assembly,
  a separate code path), per element and summed, and = f.u at a solution;
* closed forms for affine displacement fields (constant strain eps): every
  element has ce = h^3 eps^T D eps and stress E D eps; uniaxial, pure-shear
  and hydrostatic states give von Mises |s|, sqrt(3)|tau| and 0; the normal
  stress across a direction is s cos^2;
* finite differences: d(f.u)/dEe_e = -ce_e (adjoint identity), and the whole
  chain x -> filter -> projection -> stiffness law -> solve -> objective
  ("sum" and the p-norm "worst") against the oracle's dcx, and the volume
  against dvx.  With frozen voxels the reference keeps dxp on frozen voxels
  (optimize.py:89-93 sets xp = 1 there but leaves dxp), so the oracle's dcx
  exceeds the true derivative by exactly the backward-filtered frozen term;
  that identity is tested too.
"""
import numpy as np
import pytest

from oracle.fem_ref import FEMRef, Grid, elasticity_D
from oracle.optimize_ref import FilterRef, combine_cases, element_sensitivity, heaviside

NU = 0.3


def _clamped(shape, seed, inactive=0.0):
    rng = np.random.default_rng(seed)
    g = Grid(shape, 1.0, (0, 0, 0))
    active = rng.random(g.nelem) >= inactive
    active[0] = True
    n = np.arange(g.nnode)
    clamp = n[n % (g.nx + 1) == 0]
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    return g, active, fixed, rng


def _affine(g, A):
    """Nodal displacement u(p) = A p over the whole grid (interleaved dofs)."""
    p = g.node_coords(np.arange(g.nnode))
    return (p @ A.T).ravel()


def _voigt_strain(A):
    """Engineering Voigt strain (xx yy zz xy yz zx) of the displacement gradient A."""
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]])


def test_energy_identity_against_assembly():
    g, active, fixed, rng = _clamped((4, 3, 2), 0, inactive=0.2)
    fem = FEMRef(g, active, fixed, NU)
    Ee = rng.uniform(0.2, 3.0, fem.act.size)
    K = fem.assemble(Ee)
    u = rng.standard_normal(fem.nfree)
    ce = fem.element_energy(fem.expand(u))
    assert np.isclose(Ee @ ce, u @ (K @ u), rtol=1e-12)
    # each element on its own: Ee = indicator of element k
    for k in rng.choice(fem.act.size, 5, replace=False):
        ind = np.zeros(fem.act.size)
        ind[k] = 1.0
        assert np.isclose(ce[k], u @ (fem.assemble(ind) @ u), rtol=1e-12, atol=1e-14)
    # at a solution the total energy equals the compliance f.u
    f = rng.standard_normal(fem.nfree)
    us = fem.solve(K, [f], rtol=1e-13)[0]
    assert np.isclose(Ee @ fem.element_energy(fem.expand(us)), f @ us, rtol=1e-10)


@pytest.mark.parametrize("h", [1.0, 0.7])
def test_energy_and_stress_of_constant_strain(h):
    rng = np.random.default_rng(1)
    g = Grid((3, 2, 2), h, (0.3, -1.0, 2.0))
    fem = FEMRef(g, np.ones(g.nelem, bool), [], NU)
    A = rng.standard_normal((3, 3))
    u = _affine(g, A)
    eps = _voigt_strain(A)
    D = elasticity_D(NU)
    ce = fem.element_energy(u)
    assert np.allclose(ce, h ** 3 * eps @ D @ eps, rtol=1e-12)
    E = rng.uniform(1.0, 5.0, fem.act.size)
    sig = fem.stress(u, E)
    assert np.allclose(sig, E[:, None] * (D @ eps)[None, :], rtol=1e-12, atol=1e-12)
    assert np.allclose(fem.stress(u, 2.0), 2.0 * (D @ eps)[None, :], rtol=1e-12, atol=1e-12)


def test_von_mises_and_normal_stress_closed_forms():
    g = Grid((2, 2, 2), 1.0, (0, 0, 0))
    fem = FEMRef(g, np.ones(g.nelem, bool), [], NU)
    Dinv = np.linalg.inv(elasticity_D(NU))

    def field_for(sig_voigt, E):
        eps = Dinv @ np.asarray(sig_voigt, float) / E
        A = np.diag(eps[:3])
        A[0, 1] = A[1, 0] = eps[3] / 2
        A[1, 2] = A[2, 1] = eps[4] / 2
        A[0, 2] = A[2, 0] = eps[5] / 2
        return _affine(g, A)
    E = 3.0
    s, tau, pr = 7.0, -2.5, 4.0
    vm = fem.von_mises(field_for([s, 0, 0, 0, 0, 0], E), E)
    assert np.allclose(vm, abs(s), rtol=1e-12)
    vm = fem.von_mises(field_for([0, 0, 0, tau, 0, 0], E), E)
    assert np.allclose(vm, np.sqrt(3) * abs(tau), rtol=1e-12)
    vm = fem.von_mises(field_for([0, 0, 0, 0, 0, tau], E), E)
    assert np.allclose(vm, np.sqrt(3) * abs(tau), rtol=1e-12)
    vm = fem.von_mises(field_for([pr, pr, pr, 0, 0, 0], E), E)
    assert np.all(vm < 1e-12 * pr)
    sig = fem.stress(field_for([s, 0, 0, 0, 0, 0], E), E)
    th = 0.6
    assert np.allclose(FEMRef.normal_stress(sig, [np.cos(th), np.sin(th), 0.0]), s * np.cos(th) ** 2, rtol=1e-12)
    # shear xy seen across the 45-degree direction: sigma_n = tau
    sig = fem.stress(field_for([0, 0, 0, tau, 0, 0], E), E)
    assert np.allclose(FEMRef.normal_stress(sig, [1.0, 1.0, 0.0]), tau, rtol=1e-12)


def test_compliance_derivative_wrt_stiffness_is_minus_energy():
    """d(f.u)/dEe_e = -u_e^T ke u_e with K(Ee) u = f (the adjoint identity behind optimize.py:126-129)."""
    g, active, fixed, rng = _clamped((4, 3, 2), 2)
    fem = FEMRef(g, active, fixed, NU)
    Ee = rng.uniform(0.5, 2.0, fem.act.size)
    f = rng.standard_normal(fem.nfree)

    def comp(E):
        return f @ fem.solve(fem.assemble(E), [f], rtol=1e-14)[0]
    u = fem.solve(fem.assemble(Ee), [f], rtol=1e-14)[0]
    ce = fem.element_energy(fem.expand(u))
    for k in rng.choice(fem.act.size, 6, replace=False):
        d = 1e-4 * Ee[k]
        Ep, Em = Ee.copy(), Ee.copy()
        Ep[k] += d
        Em[k] -= d
        fd = (comp(Ep) - comp(Em)) / (2 * d)
        assert abs(fd + ce[k]) <= 1e-6 * abs(ce[k]), (k, fd, -ce[k])


def _chain_problem(frozen_frac, seed):
    g, active, fixed, rng = _clamped((6, 4, 3), seed, inactive=0.1)
    fem = FEMRef(g, active, fixed, NU)
    n = fem.act.size
    phi = rng.uniform(0.5, 1.0, n)
    frz = rng.random(n) < frozen_frac
    filt = FilterRef(g, active, 1.5)
    Fs = [fem.reduce_load(rng.standard_normal(3 * g.nnode)) for _ in range(2)]
    x = rng.uniform(0.2, 0.8, n)
    return g, fem, filt, phi, frz, Fs, x, rng


def _objective_and_chain(fem, filt, phi, frz, Fs, x, beta, p, E0, emin, w, c0, combine):
    xp, dxp = heaviside(filt.forward(x), beta)
    xp = xp.copy()
    xp[frz] = 1.0
    Ee = E0 * phi * (emin + xp ** p * (1 - emin))
    U = fem.solve(fem.assemble(Ee), Fs, rtol=1e-14)
    cs = np.array([f @ u for f, u in zip(Fs, U)])
    dce = [element_sensitivity(fem, u, xp, p, E0, phi, emin) for u in U]
    obj, dc = combine_cases(cs, dce, w, c0, combine)
    vol = (phi * xp).sum()
    return obj, vol, dc, dxp


@pytest.mark.parametrize("combine", ["sum", "worst"])
@pytest.mark.parametrize("frozen_frac", [0.0, 0.15])
def test_design_sensitivity_chain_finite_differences(combine, frozen_frac):
    """dcx = backward(dc * dxp) and dvx = backward(phi * dxp) (optimize.py:141-142) vs central differences
    of the objective and volume over the design, through filter, projection (beta 2), penalty 3 and solve."""
    g, fem, filt, phi, frz, Fs, x, rng = _chain_problem(frozen_frac, 3 if combine == "sum" else 4)
    beta, p, E0, emin = 2.0, 3.0, 2.0, 1e-6
    w = np.array([1.0, 0.6])
    c0 = np.array([0.8, 1.7]) * 1e2
    args = (fem, filt, phi, frz, Fs)
    obj, vol, dc, dxp = _objective_and_chain(*args, x, beta, p, E0, emin, w, c0, combine)
    assert np.all(dc <= 0)      # compliance sensitivities are non-positive: the clamp is inactive here
    dcx = filt.backward(dc * dxp)
    dvx = filt.backward(phi * dxp)
    # what the reference's chain carries through frozen voxels beyond the true derivative
    dcx_true = dcx - filt.backward(dc * dxp * frz)
    dvx_true = dvx - filt.backward(phi * dxp * frz)
    if frozen_frac == 0.0:
        assert np.array_equal(dcx_true, dcx)
    else:
        assert frz.any() and np.abs(dcx - dcx_true).max() > 1e-3 * np.abs(dcx).max()
    des = np.flatnonzero(~frz)
    for k in rng.choice(des, 6, replace=False):
        d = 1e-6
        xp_, xm_ = x.copy(), x.copy()
        xp_[k] += d
        xm_[k] -= d
        op, vp, _, _ = _objective_and_chain(*args, xp_, beta, p, E0, emin, w, c0, combine)
        om, vm, _, _ = _objective_and_chain(*args, xm_, beta, p, E0, emin, w, c0, combine)
        fd_o, fd_v = (op - om) / (2 * d), (vp - vm) / (2 * d)
        assert abs(fd_o - dcx_true[k]) <= 1e-6 * np.abs(dcx_true).max(), (k, fd_o, dcx_true[k])
        assert abs(fd_v - dvx_true[k]) <= 1e-7 * np.abs(dvx_true).max(), (k, fd_v, dvx_true[k])


def test_worst_case_combination_is_the_pnorm_gradient():
    """combine_cases('worst'): obj = ||w cs / c0||_8 and its chain factor, against finite differences in cs."""
    rng = np.random.default_rng(5)
    cs = rng.uniform(1, 3, 4)
    w = rng.uniform(0.5, 1.5, 4)
    c0 = rng.uniform(1, 2, 4)
    # dce_i = unit vectors: dc then equals the gradient d obj / d cs (all entries are negated to stay <= 0)
    dce = [-np.eye(4)[i] for i in range(4)]
    obj, dc = combine_cases(cs, dce, w, c0, "worst")
    assert np.isclose(obj, np.linalg.norm(w * cs / c0, 8), rtol=1e-14)
    for i in range(4):
        e = np.eye(4)[i] * 1e-6
        fd = (combine_cases(cs + e, dce, w, c0, "worst")[0] - combine_cases(cs - e, dce, w, c0, "worst")[0]) / 2e-6
        assert np.isclose(-dc[i], fd, rtol=1e-7)
    obj, dc = combine_cases(cs, dce, w, c0, "sum")
    assert np.isclose(obj, w @ cs) and np.allclose(-dc, w)
