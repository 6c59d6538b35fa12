"""Voxel hex8 finite elements on the GPU (drop-in for the reference's anneal/fem.py).

Same public names and signatures as the reference (fem.py:20-263).  What
changes underneath:

* ``FEM.assemble(Ee)`` forms no matrix.  It hands the stiffness scales to a
  device context (libanneal_hip.so) and returns a :class:`DeviceOperator`;
  the operator is matrix-free (ke, Ee, free-dof mask).
* ``FEM.solve`` runs geometric-multigrid-preconditioned CG on the GPU for all
  right-hand sides at once: FP32 operator and V-cycle, FP64 dot products and
  FP64 true-residual refinement, stopping at ``||f - K u|| <= rtol ||f||``
  exactly like the reference's scipy CG (fem.py:28-33, 206-232).
* Returned displacements are host float64 vectors over the free dofs, as in
  the reference, so callers (``f @ u``, ``expand``) work unchanged.

There is no CPU fallback: without the built library or a HIP device the
constructor raises.
"""
import time

import numpy as np

from . import _lib

# Corner order of the hex8 brick (fem.py:20-21): the reference's local node order.
LOCAL = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


class SolverError(RuntimeError):
    """Solver failure (fem.py:24-25); message text kept from fem.py:138-139, 225-229."""


NOT_CONVERGED = ("The solver did not converge in {maxiter} iterations. Usual causes: the part can move "
                 "or turn freely (add a fixed anchor, or fix more directions on one), or the design has "
                 "pieces held only by near-empty voxels (raise volumeFraction or iterations).")


def elasticity_D(nu):
    """Isotropic constitutive matrix for E = 1, Voigt xx yy zz xy yz zx (fem.py:36-43)."""
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 0.5 / (1 + nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def _strain_operator(xi, eta, zeta, h):
    """6x24 B at a natural coordinate for a cube of side h (fem.py:46-68)."""
    s = 2.0 * LOCAL - 1.0
    p = np.array([xi, eta, zeta])
    lin = 1.0 + s * p                                   # (8, 3): 1 + s_d * p_d
    grad = np.empty((8, 3))
    grad[:, 0] = s[:, 0] * lin[:, 1] * lin[:, 2]
    grad[:, 1] = s[:, 1] * lin[:, 0] * lin[:, 2]
    grad[:, 2] = s[:, 2] * lin[:, 0] * lin[:, 1]
    grad *= 2.0 / (8.0 * h)
    B = np.zeros((6, 8, 3))
    for d in range(3):
        B[d, :, d] = grad[:, d]
    for row, (a, b) in zip((3, 4, 5), ((0, 1), (1, 2), (2, 0))):
        B[row, :, a] = grad[:, b]
        B[row, :, b] = grad[:, a]
    return B.reshape(6, 24)


def hex8_matrices(nu, h=1.0):
    """(ke, B0, D): 2x2x2 Gauss element stiffness for E = 1, centre B, D (fem.py:71-84)."""
    D = elasticity_D(nu)
    g = 1.0 / np.sqrt(3.0)
    Bs = np.stack([_strain_operator(a, b, c, h) for c in (-g, g) for b in (-g, g) for a in (-g, g)])
    ke = np.einsum("qki,kl,qlj->ij", Bs, D, Bs) * (h / 2.0) ** 3
    return 0.5 * (ke + ke.T), _strain_operator(0.0, 0.0, 0.0, h), D


class Grid:
    """Box of nx x ny x nz cubic voxels of side h whose lowest corner sits at ``origin``.

    Voxels and lattice nodes are numbered with x varying fastest, then y, then z;
    the public methods and that numbering are those of fem.py:87-127.
    """

    def __init__(self, shape, h, origin):
        self.shape = tuple(int(s) for s in shape)
        self.nx, self.ny, self.nz = self.shape
        self.h = float(h)
        self.origin = np.asarray(origin, dtype=float)
        self.nelem = self.nx * self.ny * self.nz
        self.nnode = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        # node-index step of one lattice unit along x, y, z; the 8 corners' offsets from corner (0,0,0)
        self._nstep = np.array([1, self.nx + 1, (self.nx + 1) * (self.ny + 1)])
        self._noff = LOCAL @ self._nstep

    def elem_ijk(self, e):
        rest, i = np.divmod(np.asarray(e), self.nx)
        k, j = np.divmod(rest, self.ny)
        return i, j, k

    def element_nodes(self, e):
        """Node ids of each voxel's 8 corners, in the LOCAL corner order: shape (len(e), 8)."""
        ijk = np.stack(self.elem_ijk(e), axis=-1)
        return (ijk @ self._nstep)[:, None] + self._noff[None, :]

    def node_coords(self, n):
        rest, i = np.divmod(np.asarray(n), self.nx + 1)
        k, j = np.divmod(rest, self.ny + 1)
        return self.origin + self.h * np.stack([i, j, k], axis=-1)

    def elem_centers(self, e=None):
        ijk = np.stack(self.elem_ijk(np.arange(self.nelem) if e is None else e), axis=-1)
        return self.origin + self.h * (ijk + 0.5)

    def to3d(self, flat):
        """Per-voxel values in voxel order -> a 3-D array indexed [i, j, k] (a view where possible)."""
        return np.asarray(flat).reshape(self.nz, self.ny, self.nx).T

    def to1d(self, arr3):
        """Inverse of to3d: a 3-D array indexed [i, j, k] -> values in voxel order."""
        return np.asarray(arr3).T.ravel()


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("anneal_amd needs a HIP device (MI355X); none is visible to torch")
    return torch


class DeviceOperator:
    """What ``FEM.assemble`` returns: the matrix-free stiffness K(Ee) on the device.

    ``K @ u_red`` (a free-dof float64 vector) applies K in FP64 on the GPU,
    mirroring ``csr_matrix @ u`` of the reference's assembled K.
    """

    def __init__(self, fem, Ee_full_dev):
        self.fem = fem
        self.Ee = Ee_full_dev
        self.shape = (fem.nfree, fem.nfree)

    def __matmul__(self, u_red):
        fem = self.fem
        fem._activate(self)
        torch = _torch()
        U = fem._to_device([np.asarray(u_red, float)])
        Y = torch.empty_like(U)
        _lib.call("anneal_fem_apply_f64", fem._ctx, _lib.ptr(U), _lib.ptr(Y), 1)
        return fem._from_device(Y, 1)[0]

    dot = __matmul__


class FEM:
    """Linear elastic solve on the active voxels of a grid (fem.py:130-263), on the GPU."""

    def __init__(self, grid, active, fixed_dofs, nu, log=None, nrhs_max=4, max_coarse_nodes=None):
        t0 = time.time()
        torch = _torch()
        self.grid = grid
        self.log = log or (lambda *a: None)
        active = np.asarray(active).astype(bool).ravel()
        self.act = np.flatnonzero(active)
        if self.act.size == 0:
            raise SolverError("No active voxels: the design space is empty.")
        self.ke, self.B0, self.D = hex8_matrices(nu, grid.h)
        self.nu = float(nu)
        fixed = np.unique(np.asarray(fixed_dofs, dtype=np.int64).ravel())
        if fixed.size and (fixed.min() < 0 or fixed.max() >= 3 * grid.nnode):
            raise IndexError("fixed dof index out of range")
        fixed_bits = np.zeros(grid.nnode, np.uint8)
        np.bitwise_or.at(fixed_bits, fixed // 3, (1 << (fixed % 3)).astype(np.uint8))
        self.nrhs_max = int(nrhs_max)
        self.stream = torch.cuda.current_stream()
        ctx = _lib.c_void_p()
        act_u8 = np.ascontiguousarray(active, dtype=np.uint8)
        _lib.call("anneal_fem_create", _lib.ctypes.byref(ctx), grid.nx, grid.ny, grid.nz, grid.h, self.nu,
                  _lib.ptr(act_u8), _lib.ptr(fixed_bits), self.nrhs_max, int(max_coarse_nodes or 0),
                  _lib.c_void_p(self.stream.cuda_stream))
        self._ctx = ctx
        ldn, lev = _lib.c_int64(), _lib.c_int()
        mask = np.zeros(grid.nnode, np.uint8)
        _lib.call("anneal_fem_info", ctx, _lib.ctypes.byref(ldn), _lib.ctypes.byref(lev), _lib.ptr(mask))
        self.ldn = int(ldn.value)
        self.levels = int(lev.value)
        self.dof_mask = mask
        dofmask = ((mask[:, None] >> np.arange(3)[None, :]) & 1).astype(bool).ravel()
        self.free = np.flatnonzero(dofmask)
        self.nfree = self.free.size
        red = np.full(3 * grid.nnode, -1, dtype=np.int64)
        red[self.free] = np.arange(self.nfree)
        self.red = red
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.set_option("rebuild_each_solve", 0)   # the coarse inverse follows solve(rebuild=...)
        self._active_dev = torch.from_numpy(act_u8).to(self.device)
        self._current = None
        self.amg_builds = 0
        self.cg_iters = []
        self.log(f"  system: {self.act.size:,} elements, {self.nfree:,} free dofs, matrix-free hex8 operator "
                 f"on the GPU ({self.levels} multigrid levels), built in {time.time() - t0:.1f} s")

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_fem_destroy", self._ctx)
            self._ctx = _lib.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, name, value):
        """Solver knobs of the device context (anneal_fem_set_option): graphs, smoother_degree (fine level),
        smoother_degree_coarse (levels >= 1), smoother_degree_deep (levels >= 2, 0 = as coarse),
        deep_escalate_iters / deep_escalate_degree (slow solves raise the deep-level degree once), cheb_ratio,
        coarsen_power, fuse_first_step, fuse_first_level, inner_tol_factor, inner_tol_floor, inner_max,
        rebuild_each_solve."""
        _lib.call("anneal_fem_set_option", self._ctx, name.encode(), float(value))

    def autotune(self, variant=-1):
        """Pick the FP32 apply tile shape (variant < 0: time all on this grid); returns (chosen, us per shape).

        A context picks its shapes when it is created (the process's earlier choice for the same grid
        hierarchy, else this sweep); call this to re-time or to pin one shape on every level."""
        chosen, n = _lib.ctypes.c_int(), _lib.ctypes.c_int()
        us = (_lib.ctypes.c_float * 16)()
        _lib.call("anneal_fem_autotune", self._ctx, int(variant), _lib.ctypes.byref(chosen), us,
                  _lib.ctypes.byref(n))
        return chosen.value, ([us[i] for i in range(n.value)] if variant < 0 else [])

    def level_info(self):
        """[(nx, ny, nz, tile shape id)] per multigrid level, finest first."""
        out = []
        for l in range(self.levels):
            v = [_lib.ctypes.c_int() for _ in range(4)]
            _lib.call("anneal_fem_level_info", self._ctx, l, *[_lib.ctypes.byref(x) for x in v])
            out.append(tuple(x.value for x in v))
        return out

    def set_level_tiles(self, tiles):
        """Pin the FP32 apply tile shape id of each multigrid level, finest first (ids as in level_info)."""
        for l, t in enumerate(tiles):
            _lib.call("anneal_fem_set_level_tile", self._ctx, int(l), int(t))

    def set_learned(self, net, level=1, e_ref=1.0, mode=2):
        """Attach a U-Net (anneal_amd.learned.UNet) as a preconditioner stage of every solve.

        mode 1: warm start x = z1 (z1 = gamma P N(R r), gamma energy-optimal) before plain PCG;
        mode 2: z = z1 + V(r - K z1) in every flexible-CG iteration;
        mode 0 or net None: plain multigrid PCG.  The network runs on the node
        grid of multigrid level ``level``; e_ref is the stiffness scale (E0).
        """
        if net is None or mode == 0:
            _lib.call("anneal_fem_set_learned", self._ctx, None, 1, 1.0, 0)
            self._net = None
            return
        _lib.call("anneal_fem_set_learned", self._ctx, net._ctx, int(level), float(e_ref), int(mode))
        self._net = net          # the context must outlive its use here

    def learned_stage(self, r, z, nrhs):
        """z = P^l N(R^l r) on FP32 SoA device blocks [nrhs, 3, ldn]."""
        _lib.call("anneal_fem_learned_stage", self._ctx, _lib.ptr(r), _lib.ptr(z), int(nrhs))

    def learned_pair(self, F, U, nrhs, level=1, e_ref=1.0):
        """(input bf16 [nrhs, V, 8], target FP32 [nrhs, V, 3]) for loads F and solutions U (FP64 SoA device)."""
        torch = _torch()
        nx, ny, nz, _ = self.level_info()[level]
        V = (nx + 1) * (ny + 1) * (nz + 1)
        x = torch.empty(nrhs, V, 8, dtype=torch.bfloat16, device=self.device)
        t = torch.empty(nrhs, V, 3, dtype=torch.float32, device=self.device)
        _lib.call("anneal_fem_learned_pair", self._ctx, _lib.ptr(F), _lib.ptr(U), int(nrhs), int(level),
                  float(e_ref), _lib.ptr(x), _lib.ptr(t))
        return x, t, (nx + 1, ny + 1, nz + 1)

    def tile(self):
        """Tile shape id in use by the FP32 apply."""
        chosen = _lib.ctypes.c_int()
        _lib.call("anneal_fem_autotune", self._ctx, -2, _lib.ctypes.byref(chosen), None, None)
        return chosen.value

    # -- layout helpers (AoS dof 3n+c on the host <-> SoA [c][n] on the device)
    def _to_device(self, vecs, reduced=True):
        torch = _torch()
        nn = self.grid.nnode
        out = torch.zeros((len(vecs), 3, self.ldn), dtype=torch.float64, device=self.device)
        for r, v in enumerate(vecs):
            full = np.zeros(3 * nn)
            if reduced:
                full[self.free] = v
            else:
                full[:] = np.asarray(v, float)
            out[r, :, :nn] = torch.from_numpy(full.reshape(nn, 3).T.copy()).to(self.device)
        return out

    def _from_device(self, U, nrhs, reduced=True):
        nn = self.grid.nnode
        host = U[:nrhs, :, :nn].transpose(1, 2).contiguous().cpu().numpy().reshape(nrhs, 3 * nn)
        return [h[self.free].copy() if reduced else h.copy() for h in host]

    def _activate(self, K):
        if self._current is not K:
            _lib.call("anneal_fem_set_Ee", self._ctx, _lib.ptr(K.Ee))
            self._current = K

    # -- reference API -----------------------------------------------------
    def assemble(self, Ee):
        """Ee: stiffness scale per active element (fem.py:198-201) -> DeviceOperator."""
        torch = _torch()
        Ee = np.asarray(Ee, dtype=float).ravel()
        if Ee.size != self.act.size:
            raise ValueError(f"Ee has {Ee.size} entries, expected one per active element ({self.act.size})")
        full = np.zeros(self.grid.nelem)
        full[self.act] = Ee
        K = DeviceOperator(self, torch.from_numpy(full).to(self.device))
        self._activate(K)
        return K

    def assemble_device(self, Ee_full_dev):
        """Device variant: Ee over the full grid as a float64 CUDA tensor (0 on inactive voxels)."""
        K = DeviceOperator(self, Ee_full_dev)
        self._activate(K)
        return K

    def reduce_load(self, F_global):
        return np.asarray(F_global)[self.free]

    def solve(self, K, F_list, x0_list=None, rebuild=True, rtol=1e-8, maxiter=2000):
        """K u = f for every f in F_list (fem.py:206-232); returns free-dof float64 vectors."""
        self._activate(K)
        if rebuild or self.amg_builds == 0:
            self.rebuild()
        out = []
        for s in range(0, len(F_list), self.nrhs_max):
            Fs = [np.asarray(f, float) for f in F_list[s:s + self.nrhs_max]]
            x0s = None if x0_list is None else x0_list[s:s + self.nrhs_max]
            F = self._to_device(Fs)
            U = self._to_device(x0s) if x0s is not None else _torch().zeros_like(F)
            iters = self.solve_device(F, U, len(Fs), rtol, maxiter, use_x0=x0s is not None)
            self.cg_iters.extend(iters)
            out.extend(self._from_device(U, len(Fs)))
        return out

    def rebuild(self):
        """Refresh the multigrid hierarchy's coarse inverse for the current operator (fem.py:207-216)."""
        _lib.call("anneal_fem_rebuild", self._ctx)
        self.amg_builds += 1

    def solve_device(self, F, U, nrhs, rtol, maxiter, use_x0):
        """Device-resident solve on SoA FP64 blocks [nrhs, 3, ldn]; returns PCG iterations per rhs."""
        its = (_lib.ctypes.c_int * nrhs)()
        try:
            _lib.call("anneal_fem_solve", self._ctx, _lib.ptr(F), _lib.ptr(U), nrhs, float(rtol), int(maxiter),
                      int(bool(use_x0)), its)
        except _lib.AnnealLibError as e:
            if e.status in (3, 4):
                raise SolverError(NOT_CONVERGED.format(maxiter=maxiter)) from None
            raise
        return [int(v) for v in its]

    def expand(self, u_red):
        u = np.zeros(3 * self.grid.nnode)
        u[self.free] = u_red
        return u

    def element_energy(self, u_full):
        """u_e^T ke u_e per active element for E = 1 (fem.py:239-242), computed on the GPU."""
        torch = _torch()
        U = self._to_device([np.asarray(u_full, float)], reduced=False)
        ce = torch.empty(self.grid.nelem, dtype=torch.float64, device=self.device)
        _lib.call("anneal_fem_element_energy", self._ctx, _lib.ptr(U), 1, _lib.ptr(ce))
        return ce.cpu().numpy()[self.act]

    @property
    def edof(self):
        en = self.grid.element_nodes(self.act)
        return (3 * en[:, :, None] + np.arange(3)).reshape(-1, 24)

    def stress_device(self, U, nrhs, E_material, von_mises=True):
        """Element-centre stress on the device for SoA FP64 solutions U [nrhs, 3, ldn].

        Returns (sig [nrhs, nelem, 6], vm [nrhs, nelem] or None) over the full grid
        (zeros on inactive voxels); E_material is a scalar or one value per active
        element (fem.py:244-263).
        """
        torch = _torch()
        nel = self.grid.nelem
        sig = torch.empty(nrhs, nel, 6, dtype=torch.float64, device=self.device)
        vm = torch.empty(nrhs, nel, dtype=torch.float64, device=self.device) if von_mises else None
        E_dev, Es = None, 0.0
        if np.ndim(E_material):
            Ef = np.zeros(nel)
            Ef[self.act] = np.asarray(E_material, float)
            E_dev = torch.from_numpy(Ef).to(self.device)
        else:
            Es = float(E_material)
        B0 = np.ascontiguousarray(self.B0, dtype=np.float64)
        D = np.ascontiguousarray(self.D, dtype=np.float64)
        loc = np.ascontiguousarray(LOCAL, dtype=np.int32)
        _lib.call("anneal_fem_stress", self._ctx, _lib.ptr(U), int(nrhs), _lib.ptr(B0), _lib.ptr(D), _lib.ptr(loc),
                  _lib.ptr(E_dev) if E_dev is not None else None, Es, _lib.ptr(sig),
                  _lib.ptr(vm) if vm is not None else None)
        return sig, vm

    def stress(self, u_full, E_material):
        """Element-centre stress (MPa), Voigt order xx yy zz xy yz zx, per active element (fem.py:244-248)."""
        U = self._to_device([np.asarray(u_full, float)], reduced=False)
        sig, _ = self.stress_device(U, 1, E_material, von_mises=False)
        return sig[0].cpu().numpy()[self.act]

    @staticmethod
    def normal_stress(sig, d):
        """Normal stress n.S.n on the planes whose normal is the direction d (fem.py:250-256), from Voigt
        stresses (xx yy zz xy yz zx) per element."""
        n = np.asarray(d, float)
        n = n / np.linalg.norm(n)
        tensor = np.asarray(sig)[:, [[0, 3, 5], [3, 1, 4], [5, 4, 2]]]      # (m, 3, 3) symmetric stress tensors
        return np.einsum("i,mij,j->m", n, tensor, n)

    def von_mises(self, u_full, E_material):
        """Element-centre von Mises stress per active element (fem.py:258-263), computed on the GPU."""
        U = self._to_device([np.asarray(u_full, float)], reduced=False)
        _, vm = self.stress_device(U, 1, E_material)
        return vm[0].cpu().numpy()[self.act]
