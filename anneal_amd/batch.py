"""Several independent SIMP problems (any grid sizes, several load cases each) advanced together.

No counterpart in the reference, which runs one part at a time
(anneal/optimize.py:50-180).  The solves of all problems go through one
``BatchSolve``: problems of the same grid shape form a ``FemGroup`` (one block
of right-hand sides, every kernel of the solve covering all of them in one
launch), and the groups and remaining single problems advance together
(anneal_batch_solve: each context's PCG iteration on its own stream, the next
iteration launched before the host reads the current one's norms, one host
round per iteration for the whole batch).  The rest of each SIMP
iteration (filter, projection, sensitivities, OC) is the per-problem device
code of ``SimpRun``.
"""
import ctypes

import numpy as np

from . import _lib
from .fem import NOT_CONVERGED, SolverError
from .optimize import SimpRun


class FemGroup:
    """Problems of one grid shape solved as one block of right-hand sides (anneal_fem_group_create).

    Problem p owns rhs [p * k, (p + 1) * k) of a solve with nrhs = k * len(fems).
    The member FEMs keep their stiffness and coarse rebuilds and must outlive the group."""

    def __init__(self, fems, nrhs_max):
        import torch
        self.fems = list(fems)
        self.ldn = self.fems[0].ldn
        self.nrhs_max = int(nrhs_max)
        self._ctx = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(self.fems))(*[f._ctx.value for f in self.fems])
        _lib.call("anneal_fem_group_create", ctypes.byref(self._ctx), arr, len(self.fems), self.nrhs_max,
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        self._F = self._U = None

    def buffers(self, k):
        """Device blocks [k * len(fems), 3, ldn] FP64 for k right-hand sides per problem."""
        import torch
        n = k * len(self.fems)
        if self._F is None or self._F.shape[0] != n:
            dev = torch.device("cuda", torch.cuda.current_device())
            self._F = torch.zeros((n, 3, self.ldn), dtype=torch.float64, device=dev)
            self._U = torch.zeros_like(self._F)
        return self._F, self._U

    def solve_device(self, F, U, nrhs, rtol, maxiter, use_x0):
        its = (ctypes.c_int * nrhs)()
        try:
            _lib.call("anneal_fem_solve", self._ctx, _lib.ptr(F), _lib.ptr(U), nrhs, float(rtol), int(maxiter),
                      int(bool(use_x0)), its)
        except _lib.AnnealLibError as e:
            if e.status in (3, 4):
                raise SolverError(NOT_CONVERGED.format(maxiter=maxiter)) from None
            raise
        return [int(v) for v in its]

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_fem_destroy", self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _shape_key(f):
    g = f.grid
    return (tuple(g.shape), float(g.h), float(f.nu), int(f.levels))


class BatchSolve:
    """anneal_batch_* over FEM objects that share the current stream.

    With ``group=True`` (default) problems of the same grid shape, voxel size,
    Poisson ratio and hierarchy depth are solved as a FemGroup (at most 64
    right-hand sides per group, so ``group_rhs`` cases per problem fit); the
    batch then runs over the groups and the remaining single problems.  A
    grouped problem's F and U are copied into and out of the group's blocks
    around each solve (device to device)."""

    def __init__(self, fems, group=True, group_rhs=4):
        self.fems = list(fems)
        self.groups = []          # (FemGroup, [problem indices])
        self.units = []           # per batch member: ("group", gi) or ("single", p)
        if group:
            by = {}
            for p, f in enumerate(self.fems):
                by.setdefault(_shape_key(f), []).append(p)
            cap = max(1, 64 // max(1, int(group_rhs)))
            for key, idx in by.items():
                for s0 in range(0, len(idx), cap):
                    part = idx[s0:s0 + cap]
                    if len(part) >= 2:
                        self.groups.append((FemGroup([self.fems[p] for p in part], len(part) * int(group_rhs)), part))
                        self.units.append(("group", len(self.groups) - 1))
                    else:
                        self.units.append(("single", part[0]))
        else:
            self.units = [("single", p) for p in range(len(self.fems))]
        ctxs = [self.groups[i][0]._ctx.value if kind == "group" else self.fems[i]._ctx.value for kind, i in self.units]
        self._ctx = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(ctxs))(*ctxs)
        _lib.call("anneal_batch_create", ctypes.byref(self._ctx), arr, len(ctxs))

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_batch_destroy", self._ctx)
            self._ctx = ctypes.c_void_p()
        for g, _ in self.groups:
            g.close()
        self.groups = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, F, U, nrhs, rtol, maxiter=2000, use_x0=False):
        """F[p], U[p]: FP64 SoA device blocks; returns per-problem lists of PCG iterations per rhs."""
        nrhs = [int(v) for v in nrhs]
        m = len(self.units)
        Fs, Us, nr = [], [], []
        for kind, i in self.units:
            if kind == "single":
                Fs.append(F[i])
                Us.append(U[i])
                nr.append(nrhs[i])
                continue
            g, idx = self.groups[i]
            k = nrhs[idx[0]]
            if any(nrhs[p] != k for p in idx):
                raise ValueError("problems of one group must be solved with the same number of load cases")
            GF, GU = g.buffers(k)
            for j, p in enumerate(idx):
                GF[j * k:(j + 1) * k].copy_(F[p][:k])
                if use_x0:
                    GU[j * k:(j + 1) * k].copy_(U[p][:k])
            Fs.append(GF)
            Us.append(GU)
            nr.append(k * len(idx))
        Fp = (ctypes.c_void_p * m)(*[f.data_ptr() for f in Fs])
        Up = (ctypes.c_void_p * m)(*[u.data_ptr() for u in Us])
        nra = (ctypes.c_int * m)(*nr)
        status = (ctypes.c_int * m)()
        iters = (ctypes.c_int * (64 * m))()
        lib = _lib.load()
        st = lib.anneal_batch_solve(self._ctx, Fp, Up, nra, float(rtol), int(maxiter), int(bool(use_x0)), status,
                                    iters)
        if st not in (0, 3, 4):
            raise _lib.AnnealLibError("anneal_batch_solve", st, lib.anneal_last_error().decode())
        if st != 0:
            raise SolverError(NOT_CONVERGED.format(maxiter=maxiter))
        out = [None] * len(self.fems)
        for u, (kind, i) in enumerate(self.units):
            if kind == "single":
                out[i] = [iters[64 * u + r] for r in range(nrhs[i])]
                continue
            g, idx = self.groups[i]
            k = nrhs[idx[0]]
            GF, GU = g.buffers(k)
            for j, p in enumerate(idx):
                U[p][:k].copy_(GU[j * k:(j + 1) * k])
                out[p] = [iters[64 * u + j * k + r] for r in range(k)]
        return out


class BatchSimp:
    """Lock-step SIMP over several problems; ``problems`` are dicts of SimpRun arguments.

    Each dict holds grid, fill, active, frozen, fixed, loads, weights, E0, nu,
    vf and optional keyword arguments of SimpRun (iters, r_mm, tol, emin,
    rtol, ...).  All runs must use the same rtol.  A run leaves the batch when
    it has done its own ``iters`` or met its stopping test (optimize.py:173-175),
    so each run's history and design equal those of its own ``simp()``.
    """

    def __init__(self, problems, log=lambda m: None):
        self.runs = []
        for pr in problems:
            pr = dict(pr)
            args = [pr.pop(k) for k in ("grid", "fill", "active", "frozen", "fixed", "loads", "weights", "E0", "nu",
                                        "vf")]
            pr.setdefault("log", log)
            self.runs.append(SimpRun(*args, **pr))
        rt = {r.rtol for r in self.runs}
        if len(rt) != 1:
            raise ValueError("all problems of a batch must use the same rtol")
        self.rtol = rt.pop()
        self.maxiter = max(r.maxiter for r in self.runs)
        self.solver = None
        self._members = None
        self._solver_for(self.runs)
        self.it = 0

    def active(self):
        """The runs that still iterate: fewer than their iters done and not converged."""
        return [r for r in self.runs if not r.converged and r.it < r.iters]

    def _solver_for(self, runs):
        """The batch solver over the FEMs of ``runs``, rebuilt when a run has left the batch."""
        key = tuple(id(r) for r in runs)
        if key != self._members:
            if self.solver is not None:
                self.solver.close()
            self.solver = BatchSolve([r.fem for r in runs], group_rhs=max(r.ncase for r in runs))
            self._members = key
        return self.solver

    def step(self):
        """One SIMP iteration of every active run; returns one history entry per run (None for runs that
        have finished)."""
        runs = self.active()
        if not runs:
            return [None] * len(self.runs)
        for r in runs:
            r._pre_solve()
        use_x0 = all(r.have_u for r in runs)   # every run starts cold on the first iteration
        its = self._solver_for(runs).solve([r.F for r in runs], [r.U for r in runs], [r.ncase for r in runs],
                                           self.rtol, self.maxiter, use_x0=use_x0)
        self.it += 1
        # the density side of every run in phases, one host wait per phase for the whole batch;
        # the launches are those of SimpRun.step
        for r, i in zip(runs, its):
            r._sensitivities(i)
        for r in runs:
            r._oc_begin()          # bracketed on the device: no wait for the statistics
        pending = runs
        while pending:
            pending = [r for r in pending if not r.filt.oc_poll()]
        for r in runs:
            r._oc_end()
        self._sync(runs)
        entries = {id(r): r._record() for r in runs}
        return [entries.get(id(r)) for r in self.runs]

    @staticmethod
    def _sync(runs):
        for st in {r.stream.value: r._torch_stream for r in runs}.values():
            st.synchronize()

    def run(self, iters=None):
        """Step until every run has finished (or for at most ``iters`` batch iterations); returns each
        run's ``simp()`` result."""
        n = 0
        while self.active() and (iters is None or n < iters):
            self.step()
            n += 1
        return [r.result() for r in self.runs]

    def close(self):
        if self.solver is not None:
            self.solver.close()
        for r in self.runs:
            r.fem.close()
