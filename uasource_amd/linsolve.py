"""Newton-step linear solve on the GPU (``uvh_lin_*`` in include/uvh.h).

The device form of ``CtrlVar.AsymmSolver = 'EliminateBCsSolveSystemIterativly'`` in
``UaMain/solveKApe.m:117-151``: constrained degrees of freedom eliminated, a no-fill
incomplete LU (here the block ILU(0) of the 3 x 3 node blocks), and a preconditioned Krylov
solve (here BiCGStab) to ``CtrlVar.LinSolveTol`` (1e-8 by default,
``UaMain/Ua2D_DefaultParameters.m:659``). K and the right-hand side never leave the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import uvh

EXPORTED_LIN = ("uvh_lin_create", "uvh_lin_destroy", "uvh_lin_last_error", "uvh_lin_set_krylov", "uvh_lin_factor",
                "uvh_lin_solve", "uvhb_lin_create", "uvhb_lin_destroy", "uvhb_lin_last_error", "uvhb_lin_solve",
                "uvhb_lin_begin", "uvhb_lin_iterate", "uvhb_lin_stopped", "uvhb_lin_end")

UVH_E_NOCONV = 8
METHODS = {"bicgstab": 0, "gmres": 1, "auto": 2}

_bound = False


def _bind(lib):
    global _bound
    if _bound:
        return lib
    vp = C.c_void_p
    lib.uvh_lin_create.argtypes = [vp, vp, C.POINTER(vp)]
    lib.uvh_lin_create.restype = C.c_int
    lib.uvh_lin_destroy.argtypes = [vp]
    lib.uvh_lin_destroy.restype = None
    lib.uvh_lin_last_error.argtypes = [vp]
    lib.uvh_lin_last_error.restype = C.c_char_p
    lib.uvh_lin_set_krylov.argtypes = [vp, C.c_int32, C.c_int32]
    lib.uvh_lin_set_krylov.restype = C.c_int
    lib.uvh_lin_factor.argtypes = [vp, vp, vp]
    lib.uvh_lin_factor.restype = C.c_int
    lib.uvh_lin_solve.argtypes = [vp, vp, vp, C.c_double, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double), vp]
    lib.uvh_lin_solve.restype = C.c_int
    lib.uvhb_lin_create.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    lib.uvhb_lin_create.restype = C.c_int
    lib.uvhb_lin_destroy.argtypes = [vp]
    lib.uvhb_lin_destroy.restype = None
    lib.uvhb_lin_last_error.argtypes = [vp]
    lib.uvhb_lin_last_error.restype = C.c_char_p
    lib.uvhb_lin_solve.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int32, vp, vp, vp, vp]
    lib.uvhb_lin_solve.restype = C.c_int
    lib.uvhb_lin_begin.argtypes = [vp, vp, vp, C.c_double, C.c_int32, vp]
    lib.uvhb_lin_begin.restype = C.c_int
    lib.uvhb_lin_iterate.argtypes = [vp, C.c_int32, vp]
    lib.uvhb_lin_iterate.restype = C.c_int
    lib.uvhb_lin_stopped.argtypes = [vp]
    lib.uvhb_lin_stopped.restype = C.c_int
    lib.uvhb_lin_end.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.uvhb_lin_end.restype = C.c_int
    _bound = True
    return lib


class LinearSolveError(RuntimeError):
    """The solve did not reach its tolerance (solveKApe's Krylov flag != 0) or hit a zero pivot."""


class GpuLinearSolver:
    """Block-ILU(0) preconditioned Krylov solve (BiCGStab, GMRES or BiCGStab falling back to
    GMRES) for ``K d = rhs`` of one context's mesh, with the degrees of freedom in ``fixed``
    (indices into [ub; vb; h]) held at zero.

    ``factor(K)`` takes the device K values of ``UvhContext.assemble_device``; ``solve(rhs, x)``
    takes and fills device vectors of 3*Nn values. ``on_device`` tells ``newton.solve`` to hand
    over the device K and residual instead of host copies."""

    on_device = True

    def __init__(self, ctx: uvh.UvhContext, fixed=None, rtol: float = 1e-8, maxit: int = 5000,
                 raise_on_noconv: bool = False, method: str = "auto", restart: int = 30):
        self.lib = _bind(uvh.load_library())
        self.ctx = ctx
        self.nn = ctx.nn
        self.fixed = None if fixed is None or len(fixed) == 0 else np.unique(np.asarray(fixed, dtype=np.int64))
        mask = None
        if self.fixed is not None:
            mask = np.zeros(3 * self.nn, dtype=np.uint8)
            mask[self.fixed] = 1
        h = C.c_void_p()
        st = self.lib.uvh_lin_create(ctx.h, None if mask is None else mask.ctypes.data, C.byref(h))
        if st != uvh.UVH_OK:
            raise uvh.UvhLibraryError(f"uvh_lin_create failed with status {st}")
        self.h = h
        self.set_krylov(method, restart)
        self.rtol, self.maxit = float(rtol), int(maxit)
        self.raise_on_noconv = raise_on_noconv
        self.last_iterations = 0
        self.last_relres = float("nan")
        self.last_message = ""          # empty when the last solve reached its tolerance
        self.total_iterations = 0

    def close(self):
        if self.h:
            self.lib.uvh_lin_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _err(self, st, what):
        msg = self.lib.uvh_lin_last_error(self.h).decode()
        if st == UVH_E_NOCONV:
            return LinearSolveError(f"{what}: {msg}")
        return uvh.UvhLibraryError(f"{what}: status {st}: {msg}")

    def set_krylov(self, method: str = "auto", restart: int = 30):
        """'bicgstab', 'gmres' (GMRES(restart), the reference branch's method) or 'auto'
        (BiCGStab, then GMRES from its best iterate if it fails)."""
        st = self.lib.uvh_lin_set_krylov(self.h, METHODS[method], int(restart))
        if st != uvh.UVH_OK:
            raise self._err(st, "uvh_lin_set_krylov")
        self.method, self.restart = method, int(restart)

    def factor(self, K, stream=None):
        st = self.lib.uvh_lin_factor(self.h, uvh._ptr(K), stream)
        if st != uvh.UVH_OK:
            raise self._err(st, "uvh_lin_factor")

    def solve(self, rhs, x, stream=None, rtol: float | None = None, maxit: int | None = None):
        """x = K^-1 rhs; returns (iterations, true relative residual). Raises LinearSolveError
        when the tolerance is missed only if ``raise_on_noconv`` (x then holds the last iterate,
        as the reference's gmres returns its last iterate with flag != 0)."""
        its, rel = C.c_int32(), C.c_double()
        st = self.lib.uvh_lin_solve(self.h, uvh._ptr(rhs), uvh._ptr(x), self.rtol if rtol is None else rtol,
                                    self.maxit if maxit is None else maxit, C.byref(its), C.byref(rel), stream)
        self.last_iterations, self.last_relres = its.value, rel.value
        self.last_message = "" if st == uvh.UVH_OK else self.lib.uvh_lin_last_error(self.h).decode()
        self.total_iterations += its.value
        if st != uvh.UVH_OK and (st != UVH_E_NOCONV or self.raise_on_noconv):
            raise self._err(st, "uvh_lin_solve")
        return its.value, rel.value


class GpuBatchLinearSolver:
    """BiCGStab on many members' systems at once (``uvhb_lin_*``): one launch per step of the
    iteration for all members, each member a ``GpuLinearSolver`` of its own mesh (factored with
    its ``factor`` before ``solve``). The solvers must outlive this object."""

    def __init__(self, solvers, rtol: float = 1e-8, maxit: int = 5000):
        self.lib = _bind(uvh.load_library())
        self.solvers = list(solvers)
        self.n = len(self.solvers)
        arr = (C.c_void_p * self.n)(*[sv.h.value for sv in self.solvers])
        h = C.c_void_p()
        st = self.lib.uvhb_lin_create(arr, self.n, C.byref(h))
        if st != uvh.UVH_OK:
            raise uvh.UvhLibraryError(f"uvhb_lin_create failed with status {st}")
        self.h = h
        self.rtol, self.maxit = float(rtol), int(maxit)
        self.last_iterations = np.zeros(self.n, dtype=np.int32)
        self.last_relres = np.zeros(self.n)
        self.last_converged = np.zeros(self.n, dtype=np.int32)
        self.last_message = ""

    def close(self):
        if self.h:
            self.lib.uvhb_lin_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def solve(self, rhs, x, active=None, stream=None, rtol: float | None = None, maxit: int | None = None):
        """x[m] = K_m^-1 rhs[m] for the active members (device vectors); returns (iterations,
        true relative residuals, converged) per member. A member that misses the tolerance keeps
        its best iterate, as the single solver does (no exception)."""
        vp = C.c_void_p
        r = (vp * self.n)(*[uvh._ptr(v) for v in rhs])
        xx = (vp * self.n)(*[uvh._ptr(v) for v in x])
        act = None if active is None else np.ascontiguousarray(np.asarray(active, dtype=np.uint8))
        st = self.lib.uvhb_lin_solve(self.h, r, xx, None if act is None else act.ctypes.data,
                                     self.rtol if rtol is None else rtol, self.maxit if maxit is None else maxit,
                                     self.last_iterations.ctypes.data, self.last_relres.ctypes.data,
                                     self.last_converged.ctypes.data, stream)
        self.last_message = "" if st == uvh.UVH_OK else self.lib.uvhb_lin_last_error(self.h).decode()
        if st != uvh.UVH_OK and st != UVH_E_NOCONV:
            raise uvh.UvhLibraryError(f"uvhb_lin_solve: status {st}: {self.last_message}")
        return self.last_iterations.copy(), self.last_relres.copy(), self.last_converged.copy()

    # the same solve in steps (uvhb_lin_begin / _iterate / _stopped / _end), so that a caller can
    # capture ``iterate`` in a HIP graph and replay it until ``stopped``
    def begin(self, rhs, active=None, stream=None, rtol: float | None = None, maxit: int | None = None):
        vp = C.c_void_p
        self._rhs = (vp * self.n)(*[uvh._ptr(v) for v in rhs])
        act = None if active is None else np.ascontiguousarray(np.asarray(active, dtype=np.uint8))
        st = self.lib.uvhb_lin_begin(self.h, self._rhs, None if act is None else act.ctypes.data,
                                     self.rtol if rtol is None else rtol, self.maxit if maxit is None else maxit, stream)
        if st != uvh.UVH_OK:
            raise uvh.UvhLibraryError(f"uvhb_lin_begin: status {st}: {self.lib.uvhb_lin_last_error(self.h).decode()}")

    def iterate(self, iterations: int, stream=None):
        st = self.lib.uvhb_lin_iterate(self.h, int(iterations), stream)
        if st != uvh.UVH_OK:
            raise uvh.UvhLibraryError(f"uvhb_lin_iterate: status {st}: {self.lib.uvhb_lin_last_error(self.h).decode()}")

    def stopped(self) -> bool:
        return self.lib.uvhb_lin_stopped(self.h) == 1

    def end(self, x, stream=None):
        xx = (C.c_void_p * self.n)(*[uvh._ptr(v) for v in x])
        st = self.lib.uvhb_lin_end(self.h, xx, self.last_iterations.ctypes.data, self.last_relres.ctypes.data,
                                   self.last_converged.ctypes.data, stream)
        self.last_message = "" if st == uvh.UVH_OK else self.lib.uvhb_lin_last_error(self.h).decode()
        if st != uvh.UVH_OK and st != UVH_E_NOCONV:
            raise uvh.UvhLibraryError(f"uvhb_lin_end: status {st}: {self.last_message}")
        return self.last_iterations.copy(), self.last_relres.copy(), self.last_converged.copy()
