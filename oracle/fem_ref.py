"""FP64 CPU oracle for the hex8 voxel solve (test infrastructure, never product code).

This module restates, in plain numpy/scipy, the arithmetic of the reference's
``anneal/fem.py`` so that the HIP path can be checked against it.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
it.  Each function names the reference lines it follows.

Pinning (what checks this restatement independently of itself):
  * ``hex8_matrices``: rigid-body energy < 1e-10 (validate.py:38-45), symmetry,
    rank 18, constant diagonal, closed-form uniaxial tension FL/(AE) to 1e-8
    (validate.py:48-72) and the partial-volume bar (validate.py:101-108).
  * ``Grid``: index round trips and brute-force corner enumeration.
  * ``FEMRef.assemble``: equals a dense element-by-element sum on tiny grids.
  * ``FEMRef.solve``: dense ``numpy.linalg.solve`` on tiny grids, tension and
    Timoshenko windows (validate.py:75-98), determinism (validate.py:111-121).
  * ``FEMRef.element_energy``: sum_e Ee_e ce_e = u^T K u against the CSR
    assembly (per element and summed), = f.u at a solution, h^3 eps^T D eps for
    constant-strain fields, and -ce = d(f.u)/dEe by central differences
    (tests/test_oracle_sensitivity.py).
  * ``FEMRef.stress`` / ``von_mises`` / ``normal_stress``: E D eps for
    constant-strain fields; uniaxial, pure-shear and hydrostatic closed forms.
The reference preconditions CG with pyamg (fem.py:212-215); pyamg is not
installed, so this oracle preconditions scipy's CG with its own Galerkin
geometric V-cycle.  The converged answer does not depend on that choice; the
iteration counts do ("parity unpinned" for ``cg_iters``).
"""
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# Corner order of a hex8 brick (fem.py:20-21): bottom face counter-clockwise, then top.
CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


class SolverError(RuntimeError):
    """Raised on non-convergence / empty design (fem.py:24-25, 138-139, 225-229)."""


def elasticity_D(nu):
    """Isotropic D for E=1, Voigt xx yy zz xy yz zx with engineering shear (fem.py:36-43)."""
    lam_over = nu
    scale = 1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu))
    D = np.zeros((6, 6))
    D[0:3, 0:3] = lam_over
    for a in range(3):
        D[a, a] = 1.0 - nu
        D[3 + a, 3 + a] = 0.5 * (1.0 - 2.0 * nu)
    return scale * D


def shape_gradients(xi, eta, zeta):
    """d N_a / d(xi, eta, zeta) of the trilinear shape functions (fem.py:46-52)."""
    sgn = 2 * CORNERS - 1
    pt = np.array([xi, eta, zeta])
    g = np.empty((8, 3))
    for d in range(3):
        o1, o2 = [t for t in range(3) if t != d]
        g[:, d] = sgn[:, d] * (1 + sgn[:, o1] * pt[o1]) * (1 + sgn[:, o2] * pt[o2]) / 8.0
    return g


def strain_matrix(grad):
    """6x24 strain-displacement matrix from physical gradients (fem.py:55-68)."""
    B = np.zeros((6, 24))
    for a in range(8):
        gx, gy, gz = grad[a]
        c = 3 * a
        B[0, c] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c], B[3, c + 1] = gy, gx
        B[4, c + 1], B[4, c + 2] = gz, gy
        B[5, c], B[5, c + 2] = gz, gx
    return B


def hex8_matrices(nu, h=1.0):
    """(ke, B0, D): 2x2x2 Gauss element stiffness for E=1, centre B, D (fem.py:71-84)."""
    D = elasticity_D(nu)
    half = h / 2.0
    gp = 1.0 / np.sqrt(3.0)
    ke = np.zeros((24, 24))
    for xi in (-gp, gp):
        for eta in (-gp, gp):
            for zeta in (-gp, gp):
                B = strain_matrix(shape_gradients(xi, eta, zeta) / half)
                ke += (B.T @ D @ B) * half ** 3
    B0 = strain_matrix(shape_gradients(0.0, 0.0, 0.0) / half)
    return ke, B0, D


class Grid:
    """Regular voxel grid, elements x-fastest (fem.py:87-127)."""

    def __init__(self, shape, h, origin):
        self.shape = tuple(int(s) for s in shape)
        self.nx, self.ny, self.nz = self.shape
        self.h = float(h)
        self.origin = np.asarray(origin, dtype=float)
        self.nelem = self.nx * self.ny * self.nz
        self.nnode = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self._corner_off = (CORNERS[:, 0]
                            + (self.nx + 1) * (CORNERS[:, 1] + (self.ny + 1) * CORNERS[:, 2]))

    def elem_ijk(self, e):
        e = np.asarray(e)
        return e % self.nx, (e // self.nx) % self.ny, e // (self.nx * self.ny)

    def element_nodes(self, e):
        i, j, k = self.elem_ijk(e)
        first = i + (self.nx + 1) * (j + (self.ny + 1) * k)
        return first[:, None] + self._corner_off[None, :]

    def node_coords(self, n):
        n = np.asarray(n)
        sx, sy = self.nx + 1, self.ny + 1
        ijk = np.stack([n % sx, (n // sx) % sy, n // (sx * sy)], axis=-1)
        return self.origin + self.h * ijk

    def elem_centers(self, e=None):
        if e is None:
            e = np.arange(self.nelem)
        i, j, k = self.elem_ijk(e)
        return self.origin + self.h * (np.stack([i, j, k], axis=-1) + 0.5)

    def to3d(self, flat):
        return np.asarray(flat).reshape(self.nz, self.ny, self.nx).transpose(2, 1, 0)

    def to1d(self, arr3):
        return np.asarray(arr3).transpose(2, 1, 0).ravel()


def _prolongation_1d(n_fine):
    """Linear interpolation from ceil(n/2)+1 coarse nodes to n+1 fine nodes (1-D)."""
    nc = (n_fine + 1) // 2
    rows, cols, vals = [], [], []
    for i in range(n_fine + 1):
        if i % 2 == 0:
            rows.append(i), cols.append(i // 2), vals.append(1.0)
        else:
            rows += [i, i]
            cols += [i // 2, i // 2 + 1]
            vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(n_fine + 1, nc + 1)), nc


class _GalerkinVCycle:
    """Symmetric V-cycle (Chebyshev smoothing, Galerkin coarse operators, sparse LU at the bottom).

    Oracle-only preconditioner standing in for pyamg (fem.py:212-215), which is
    absent: it changes the CG iteration count, not the converged answer.  The
    smoother is a degree-2 Chebyshev polynomial in D^-1 A on [lmax/15, lmax]
    with lmax 1.1x a Lanczos estimate, which keeps the cycle positive definite
    for strongly heterogeneous stiffness (damped Jacobi with a fixed weight
    does not: lambda_max(D^-1 A) ranges from about 3.1 to 5.3 for hex8).
    """

    def __init__(self, A, grid_shape, free, sweeps=2, ratio=15.0, coarse_limit=4000):
        self.levels = []
        shape = tuple(grid_shape)
        keep = free          # fine-level DOF ids (into 3*nnode_full of `shape`)
        while A.shape[0] > coarse_limit and min(shape) >= 2:
            P1 = []
            cshape = []
            for n in shape:
                p, nc = _prolongation_1d(n)
                P1.append(p)
                cshape.append(nc)
            Pn = sp.kron(P1[2], sp.kron(P1[1], P1[0])).tocsr()     # node-level, x fastest
            P = sp.kron(Pn, sp.identity(3)).tocsr()[keep]
            used = np.flatnonzero(np.asarray(abs(P).sum(axis=0)).ravel() > 0)
            P = P[:, used]
            Ac = (P.T @ A @ P).tocsr()
            dinv = 1.0 / A.diagonal()
            dh = sp.diags(np.sqrt(dinv))
            lmax = 1.1 * spla.eigsh(dh @ A @ dh, k=1, which="LA", return_eigenvectors=False, tol=1e-3)[0]
            self.levels.append((A, P, dinv, lmax))
            A, shape, keep = Ac, tuple(cshape), used
        self.coarse = spla.splu(A.tocsc())
        self.sweeps, self.ratio = sweeps, ratio

    def _smooth(self, A, dinv, lmax, b, x):
        lmin = lmax / self.ratio
        th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sg = th / de
        rho = 1.0 / sg
        d = dinv * (b if x is None else b - A @ x) / th
        x = d if x is None else x + d
        for _ in range(self.sweeps - 1):
            rn = 1.0 / (2.0 * sg - rho)
            d = rn * rho * d + (2.0 * rn / de) * dinv * (b - A @ x)
            rho = rn
            x = x + d
        return x

    def _cycle(self, lvl, b):
        if lvl == len(self.levels):
            return self.coarse.solve(b)
        A, P, dinv, lmax = self.levels[lvl]
        x = self._smooth(A, dinv, lmax, b, None)
        x = x + P @ self._cycle(lvl + 1, P.T @ (b - A @ x))
        return self._smooth(A, dinv, lmax, b, x)

    def operator(self, n):
        return spla.LinearOperator((n, n), matvec=lambda b: self._cycle(0, np.asarray(b).ravel()))


class FEMRef:
    """Assembled-CSR hex8 solve on the active voxels (fem.py:130-263)."""

    def __init__(self, grid, active, fixed_dofs, nu, log=None):
        t0 = time.time()
        self.grid = grid
        self.log = log or (lambda *a: None)
        self.act = np.flatnonzero(active)
        if self.act.size == 0:
            raise SolverError("No active voxels: the design space is empty.")
        self.ke, self.B0, self.D = hex8_matrices(nu, grid.h)
        en = grid.element_nodes(self.act)
        self.edof = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
        # free dofs: dofs of nodes touched by an active element, minus fixed (fem.py:145-155)
        mask = np.zeros(3 * grid.nnode, dtype=bool)
        mask[self.edof.ravel()] = True
        mask[np.asarray(fixed_dofs, dtype=np.int64)] = False
        self.free = np.flatnonzero(mask)
        self.nfree = self.free.size
        self.red = np.full(3 * grid.nnode, -1, dtype=np.int64)
        self.red[self.free] = np.arange(self.nfree)
        # fixed CSR pattern + refill map (fem.py:156-172)
        redof = self.red[self.edof]
        r = np.repeat(redof, 24, axis=1).ravel()
        c = np.tile(redof, (1, 24)).ravel()
        ok = np.flatnonzero((r >= 0) & (c >= 0))
        key = r[ok] * self.nfree + c[ok]
        ukey, inv = np.unique(key, return_inverse=True)
        self.indices = (ukey % self.nfree).astype(np.int32)
        self.indptr = np.searchsorted(ukey // self.nfree, np.arange(self.nfree + 1)).astype(np.int64)
        self.map = inv
        self.ei = ok // 576
        self.kev = self.ke.ravel()[ok % 576]
        self.nnz = ukey.size
        self.cg_iters = []
        self.amg_builds = 0
        self._M = None
        self.log(f"  system: {self.act.size:,} elements, {self.nfree:,} free dofs, "
                 f"{self.nnz:,} nonzeros, built in {time.time() - t0:.1f} s")

    def assemble(self, Ee):
        """CSR refill by bincount (fem.py:198-201)."""
        data = np.bincount(self.map, weights=self.kev * np.asarray(Ee)[self.ei], minlength=self.nnz)
        return sp.csr_matrix((data, self.indices, self.indptr), shape=(self.nfree, self.nfree))

    def reduce_load(self, F_global):
        return np.asarray(F_global)[self.free]

    def solve(self, K, F_list, x0_list=None, rebuild=True, rtol=1e-8, maxiter=2000):
        """scipy CG per right-hand side, stop at ||f-Ku|| <= rtol ||f|| (fem.py:206-232)."""
        if rebuild or self._M is None:
            self._M = _GalerkinVCycle(K, self.grid.shape, self.free).operator(self.nfree)
            self.amg_builds += 1
        out = []
        for i, f in enumerate(F_list):
            x0 = None if x0_list is None else x0_list[i]
            n = [0]

            def count(_):
                n[0] += 1
            u, info = spla.cg(K, f, x0=x0, rtol=rtol, maxiter=maxiter, M=self._M, callback=count)
            if info != 0 or not np.all(np.isfinite(u)):
                raise SolverError(
                    f"The solver did not converge in {maxiter} iterations. Usual causes: the part can move "
                    f"or turn freely (add a fixed anchor, or fix more directions on one), or the design has "
                    f"pieces held only by near-empty voxels (raise volumeFraction or iterations).")
            self.cg_iters.append(n[0])
            out.append(u)
        return out

    def expand(self, u_red):
        u = np.zeros(3 * self.grid.nnode)
        u[self.free] = u_red
        return u

    def element_energy(self, u_full):
        """u_e^T ke u_e per active element (fem.py:239-242)."""
        ue = np.asarray(u_full)[self.edof]
        return np.sum((ue @ self.ke) * ue, axis=1)

    def stress(self, u_full, E_material):
        """Centre stress, Voigt order (fem.py:244-248)."""
        ue = np.asarray(u_full)[self.edof]
        sig = (ue @ self.B0.T) @ self.D.T
        E = np.asarray(E_material)
        return sig * (E[:, None] if E.ndim else E)

    @staticmethod
    def normal_stress(sig, d):
        d = np.asarray(d, float)
        d = d / np.linalg.norm(d)
        x, y, z = d
        return (sig[:, 0] * x * x + sig[:, 1] * y * y + sig[:, 2] * z * z
                + 2.0 * (sig[:, 3] * x * y + sig[:, 4] * y * z + sig[:, 5] * z * x))

    def von_mises(self, u_full, E_material):
        s = self.stress(u_full, E_material)
        d1, d2, d3 = s[:, 0] - s[:, 1], s[:, 1] - s[:, 2], s[:, 2] - s[:, 0]
        return np.sqrt(0.5 * (d1 ** 2 + d2 ** 2 + d3 ** 2) + 3.0 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2))
