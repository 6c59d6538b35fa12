"""Shared helpers for the GPU parity tests (random problems, layout conversion)."""
import numpy as np

from oracle.fem_ref import FEMRef, Grid as GridRef


def random_problem(shape, seed, inactive_frac=0.0, fix="clamp"):
    """A voxel grid with an optional random inactive set, fixed dofs and stiffness scales."""
    rng = np.random.default_rng(seed)
    g = GridRef(shape, 1.0, (0, 0, 0))
    active = np.ones(g.nelem, bool)
    if inactive_frac > 0:
        active = rng.random(g.nelem) >= inactive_frac
        active[0] = True
    n = np.arange(g.nnode)
    i = n % (g.nx + 1)
    if fix == "clamp":
        clamp = n[i == 0]
        fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    else:
        fixed = rng.choice(3 * g.nnode, size=max(6, g.nnode // 10), replace=False)
    Ee = 10.0 ** rng.uniform(-3, 0, size=g.nelem)
    return g, active, np.sort(fixed), Ee


def aos_to_soa(u_full, nnode, ldn):
    out = np.zeros((3, ldn))
    out[:, :nnode] = np.asarray(u_full).reshape(nnode, 3).T
    return out


def soa_to_aos(U, nnode):
    return np.asarray(U)[:, :nnode].T.ravel()


def oracle_fem(g, active, fixed, nu=0.3):
    return FEMRef(g, active, fixed, nu)
