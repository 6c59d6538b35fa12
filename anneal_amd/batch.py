"""Several independent SIMP problems (any grid sizes, several load cases each) advanced together.

No counterpart in the reference, which runs one part at a time
(anneal/optimize.py:50-180).  The solves of all problems go through one
``BatchSolve`` (anneal_batch_solve): every PCG iteration of the whole batch
is a single hipGraph launch whose per-problem branches run concurrently, with
one host wait per iteration for all problems.  The rest of each SIMP iteration
(filter, projection, sensitivities, OC) is the per-problem device code of
``SimpRun``.
"""
import ctypes

import numpy as np

from . import _lib
from .fem import NOT_CONVERGED, SolverError
from .optimize import SimpRun


class BatchSolve:
    """anneal_batch_* over FEM objects that share the current stream."""

    def __init__(self, fems):
        self.fems = list(fems)
        self._ctx = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(self.fems))(*[f._ctx.value for f in self.fems])
        _lib.call("anneal_batch_create", ctypes.byref(self._ctx), arr, len(self.fems))

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.call("anneal_batch_destroy", self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, F, U, nrhs, rtol, maxiter=2000, use_x0=False):
        """F[p], U[p]: FP64 SoA device blocks; returns per-problem lists of PCG iterations per rhs."""
        n = len(self.fems)
        Fp = (ctypes.c_void_p * n)(*[f.data_ptr() for f in F])
        Up = (ctypes.c_void_p * n)(*[u.data_ptr() for u in U])
        nr = (ctypes.c_int * n)(*[int(v) for v in nrhs])
        status = (ctypes.c_int * n)()
        iters = (ctypes.c_int * (64 * n))()
        lib = _lib.load()
        st = lib.anneal_batch_solve(self._ctx, Fp, Up, nr, float(rtol), int(maxiter), int(bool(use_x0)), status,
                                    iters)
        if st not in (0, 3, 4):
            raise _lib.AnnealLibError("anneal_batch_solve", st, lib.anneal_last_error().decode())
        if st != 0:
            raise SolverError(NOT_CONVERGED.format(maxiter=maxiter))
        return [[iters[64 * p + r] for r in range(int(nrhs[p]))] for p in range(n)]


class BatchSimp:
    """Lock-step SIMP over several problems; ``problems`` are dicts of SimpRun arguments.

    Each dict holds grid, fill, active, frozen, fixed, loads, weights, E0, nu,
    vf and optional keyword arguments of SimpRun (iters, r_mm, tol, emin,
    rtol, ...).  All runs must use the same rtol.
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
        self.solver = BatchSolve([r.fem for r in self.runs])
        self.it = 0

    def step(self):
        """One SIMP iteration of every problem; returns their history entries."""
        for r in self.runs:
            r._pre_solve()
        use_x0 = all(r.have_u for r in self.runs)   # every run starts cold on the first iteration
        its = self.solver.solve([r.F for r in self.runs], [r.U for r in self.runs], [r.ncase for r in self.runs],
                                self.rtol, self.maxiter, use_x0=use_x0)
        self.it += 1
        return [r._post_solve(i) for r, i in zip(self.runs, its)]

    def run(self, iters):
        for _ in range(iters):
            self.step()
        return [r.result() for r in self.runs]

    def close(self):
        self.solver.close()
        for r in self.runs:
            r.fem.close()
