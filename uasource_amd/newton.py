"""Newton-Raphson solve of the implicit uvh system around the GPU assembly.

Restates the uvh iteration of ``UaMain/SSTREAM_TransientImplicit.m:224-418`` for problems
without boundary-condition constraints (``L`` empty):

* normalisation ``Fext0 = R`` of a zero-fields, R-only assembly (``:224-227``);
* the cost of ``UaMain/CalcCostFunctionNRuvh.m:22-98`` at ``x + gamma d``:
  ``frhs = -R``, ``rForce = frhs'frhs / (Fext0'Fext0 + 1000 eps)``, ``D2 = frhs'd``,
  ``rWork = D2^2``;
* the exit tests of ``:269-350`` (desired/acceptable work-and-force tolerances depending
  on the last step length, ``NRitmax``, backtracking failure, and stagnation of ``r/r0``
  over three iterations);
* the Newton step ``K d = -R`` (``:356-372``) and the line search in the Newton direction of
  ``UaMain/rLineminUa.m:60-135`` with the backtracking of ``UaMain/BackTracking.m``
  (full step accepted below ``NewtonAcceptRatio * r0``, else guarded quadratic fits down to
  a step of 1e-3). ``rLineminUa`` always runs its "-Auto-" mode (``rLineminUa.m:38``): where
  the Newton line search ends below gamma = 0.75 with r/r0 above 0.7 (``:94-103``), it also
  tries a Cauchy step with the mass matrix in place of the Hessian (``:137-259``), the path
  from that point towards the Newton point (``fminbnd``) and steepest descent, and keeps the
  best point. With ``CtrlVar.rLineMinUa = "-Auto-"`` this driver adds these three stages (host
  linear solves only; BackTracking's extrapolation, which the steepest-descent stage may use
  there, is not restated); the default here, "-Newton Step-",
  keeps the Newton point and stops as "backtracking stagnated" where the Newton direction
  gives no reduction.

The residual cost is device-resident (the R-only assemblies, the axpy to the trial state
and the dot products stay in HBM; three scalars come back per evaluation) and the residual
at ``gamma = 0`` is taken from the R+K assembly instead of being assembled again. The linear
solve is either on the host (SciPy's SuperLU, the default, as Úa's direct ``solveKApe``
branches) or, with a ``linsolve.GpuLinearSolver``, on the device (``solveKApe``'s iterative
branch: block ILU(0) and BiCGStab), in which case K and R never leave HBM.

Precision form: with ``fp32=True`` the Jacobian and the residual are assembled with FP32
element arithmetic while ``rForce`` is above ``fp64_switch``, the FP32 line search still
reduces it and no two steps in a row are shorter than 0.25; after that both are FP64 (the step's right-hand side, the line search and the
convergence test with the residual), so the iteration ends with FP64 Newton steps, converges to
the FP64 solution and makes its final convergence check in FP64. (An FP32 Jacobian kept to the
end converges, linearly, on small meshes, but on the 199,712-element config-2 mesh its steps
stall at gamma ~0.2 with rForce ~1e-7: the Jacobian's FP32 error is too large a part of the
ill-conditioned system's steps.)
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import uvh

EPS = np.finfo(float).eps


@dataclass
class NewtonResult:
    converged: bool
    iterations: int
    reason: str
    rForce: float
    rWork: float
    history: list = field(default_factory=list)     # (iteration, gamma, rForce, rWork, fp64 residual)
    t_assembly: float = 0.0
    t_solve: float = 0.0
    t_linesearch: float = 0.0
    n_rk: int = 0
    n_r: int = 0
    linear_iterations: list = field(default_factory=list)   # Krylov iterations per Newton step (device solver)
    checkpoint: dict | None = None   # set when the run stopped at its time limit: pass as ``resume``
    cauchy_steps: int = 0            # iterations that took a Cauchy or Cauchy-to-Newton point (rLineMinUa "-Auto-")
    luvh: np.ndarray | None = None   # Lagrange multipliers of the linear constraints L x = cuvh
    ub: np.ndarray | None = None
    vb: np.ndarray | None = None
    h: np.ndarray | None = None


def superlu_solve(K: sp.csr_matrix, rhs: np.ndarray, rtol: float = 1e-10) -> np.ndarray:
    """Host direct solve (SciPy SuperLU). K is structurally symmetric, so the fill-reducing
    order is minimum degree on A'A + A with diagonal pivots kept unless one falls below 1e-3
    of its column (at 12,800 P2 elements: 1.0 s and 15 M factor entries, against 4-5.5 s and
    53 M with COLAMD and partial pivoting). If that factorization fails or leaves a relative
    residual above ``rtol``, the COLAMD factorization with partial pivoting solves instead, and
    so for every later system of the same pattern (the Newton systems of one run share it: at
    the full 200k-element size the diagonal-pivot factorization fails from the second step on,
    after ~40-90 s of factoring)."""
    A = K.tocsc()
    key = (A.shape[0], A.nnz, hash(A.indices[:: max(1, A.nnz // 4096)].tobytes()))
    if key not in _DIAGONAL_PIVOTS_FAILED:
        try:
            x = spla.splu(A, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=1e-3,
                          options=dict(SymmetricMode=True)).solve(rhs)
            bn = np.linalg.norm(rhs)
            if bn == 0 or np.linalg.norm(rhs - A @ x) <= rtol * bn:
                return x
        except RuntimeError:   # a singular pivot in the diagonal-pivot factorization
            pass
        _DIAGONAL_PIVOTS_FAILED.add(key)
    return spla.splu(A, permc_spec="COLAMD").solve(rhs)


_DIAGONAL_PIVOTS_FAILED: set = set()   # patterns whose diagonal-pivot factorization has failed


class LaggedLU:
    """Host solve of the Newton steps with one sparse LU kept across steps: the LU of an
    earlier K preconditions GMRES on the current one, and K is factored again only when GMRES
    misses ``rtol`` within ``max_krylov`` iterations. The K of successive Newton steps share
    one pattern and change less and less as the iteration converges, so most steps cost a few
    triangular solves instead of a factorization (at the full 200k-element P2 size: ~65 s per
    factorization, ~0.8 s per solve). SuperLU orders by minimum degree on A'A + A (K is
    structurally symmetric) and keeps diagonal pivots unless one is below 1e-3 of its column:
    a threshold of 0.1 already lets the second Newton system pivot off the diagonal, which
    destroys the ordering (factorization 1 s -> minutes at 12,800 elements).

    ``rtol`` bounds the true relative residual of every step; Úa's default branch solves
    directly (``solveKApe.m:91-116``), which this matches to that residual."""

    def __init__(self, rtol: float = 1e-10, max_krylov: int = 30, restart: int = 30,
                 permc_spec: str = "MMD_AT_PLUS_A", log=None):
        self.rtol, self.max_krylov, self.restart = rtol, max_krylov, restart
        self.permc_spec = permc_spec
        self.lu = None
        self.factorizations = 0
        self.krylov = []          # GMRES iterations per call (0 where the fresh LU solved directly)
        self.t_factor = 0.0
        self.log = log

    def _factor(self, K):
        t = time.perf_counter()
        self.lu = spla.splu(K.tocsc(), permc_spec=self.permc_spec, diag_pivot_thresh=1e-3,
                            options=dict(SymmetricMode=True))
        self.factorizations += 1
        self.t_factor += time.perf_counter() - t
        if self.log:
            self.log(f"    LU factorization {self.factorizations}: {time.perf_counter() - t:.1f} s")

    def _gmres(self, K, rhs):
        n = K.shape[0]
        M = spla.LinearOperator((n, n), matvec=self.lu.solve, dtype=np.float64)
        its = [0]

        def count(_):
            its[0] += 1
        x, info = spla.gmres(K, rhs, M=M, rtol=self.rtol, atol=0.0, restart=self.restart,
                             maxiter=max(1, self.max_krylov // self.restart + 1), callback=count,
                             callback_type="pr_norm")
        bn = np.linalg.norm(rhs)
        rel = np.linalg.norm(rhs - K @ x) / bn if bn > 0 else 0.0
        return x, its[0], rel <= self.rtol and its[0] <= self.max_krylov

    def __call__(self, K, rhs):
        if self.lu is not None:
            x, its, ok = self._gmres(K, rhs)
            if ok:
                self.krylov.append(its)
                return x
        self._factor(K)
        x = self.lu.solve(rhs)
        bn = np.linalg.norm(rhs)
        if bn > 0 and np.linalg.norm(rhs - K @ x) / bn > self.rtol:   # refine with the fresh LU
            x, its, ok = self._gmres(K, rhs)
            self.krylov.append(its)
            if not ok:
                # the diagonal-pivot factorization is unstable on this K: solve it with partial
                # pivoting (COLAMD), and keep that LU for the following steps
                t = time.perf_counter()
                self.lu = spla.splu(K.tocsc(), permc_spec="COLAMD")
                self.factorizations += 1
                self.t_factor += time.perf_counter() - t
                x = self.lu.solve(rhs)
        else:
            self.krylov.append(0)
        return x


class GpuAssembler:
    """R and K of one mesh on the GPU, with F0, F1 and the Newton state resident in HBM.

    The interface the driver uses: ``normalisation()``, ``assemble_rk(fp32)``,
    ``set_step(d)``, ``cost(gamma, fp32)``, ``accept(gamma)``, ``state()``, ``set_state(x)``,
    ``set_fixed(fixed)`` and ``set_shift(a0, a1)`` (constraint terms added to R)."""

    def __init__(self, CtrlVar, MUA, F0, F1, device: int = 0, ctx: uvh.UvhContext | None = None):
        import torch
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        uvh.check_supported(CtrlVar)
        self.ctx = ctx or uvh.UvhContext(device)
        if ctx is None or ctx.mesh_key != uvh.mesh_key(MUA):
            self.ctx.set_mesh(MUA)
        self.mua = MUA
        self.nn = MUA.Nnodes
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)
        self.f1 = {k: t(getattr(F1, k)) for k in uvh.F1_KEYS}
        self.f0 = {k: t(getattr(F0, k)) for k in uvh.F0_KEYS}
        # Newton state x = [ub; vb; h] and the trial state the assembly reads
        self.x = torch.cat([self.f1["ub"], self.f1["vb"], self.f1["h"]])
        self.xt = self.x.clone()
        n = self.nn
        trial = {"ub": self.xt[:n], "vb": self.xt[n:2 * n], "h": self.xt[2 * n:]}

        class _F:
            pass
        fd1, fd0 = _F(), _F()
        for k in uvh.F1_KEYS:
            setattr(fd1, k, trial.get(k, self.f1[k]))
        for k in uvh.F0_KEYS:
            setattr(fd0, k, self.f0[k])
        for k in ("q", "muk", "V0"):   # read by the sliding laws other than Weertman
            a = getattr(F1, k, None)
            if a is not None and np.size(a) > 0:
                self.f1[k] = t(a)
                setattr(fd1, k, self.f1[k])
        self.fields = uvh.make_fields(fd0, fd1)
        self.ctrl = uvh.make_ctrl(CtrlVar, F1)
        self.R = torch.zeros(3 * n, dtype=torch.float64, device=self.dev)
        self.R0 = torch.zeros_like(self.R)
        self.K = torch.zeros(self.ctx.nnz, dtype=torch.float64, device=self.dev)
        self.d = torch.zeros_like(self.R)
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        self.n_rk = self.n_r = 0
        self.free = None
        self.shift = None

    def next_time_step(self, F1=None, dt: float | None = None):
        """Moves to the next implicit time step with the mesh, pattern and fields resident: F0
        takes the current state (u, v, h) and F1's a_s, a_b, as a converged step hands over to
        the next (``SSTREAM_TransientImplicit.m``: F0 = F1); with ``F1``, its nodal fields
        (those the assembler holds, q, muk and V0 included), its scalars g, rhow, alpha, and its
        u, v, h as the first guess replace the resident ones in place, else the last solution is
        the first guess; ``dt`` replaces CtrlVar.dt."""
        n = self.nn
        self.f0["ub"].copy_(self.x[:n])
        self.f0["vb"].copy_(self.x[n:2 * n])
        self.f0["h"].copy_(self.x[2 * n:])
        self.f0["as_"].copy_(self.f1["as_"])
        self.f0["ab"].copy_(self.f1["ab"])
        if F1 is not None:
            t = self.torch
            for k, v in self.f1.items():
                a = getattr(F1, k, None)
                if a is not None and k not in ("ub", "vb", "h"):
                    v.copy_(t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)))
            self.set_state(np.concatenate([F1.ub, F1.vb, F1.h]))
            self.ctrl.g, self.ctrl.rhow, self.ctrl.alpha = float(F1.g), float(F1.rhow), float(F1.alpha)
        if dt is not None:
            self.ctrl.dt = float(dt)

    def set_fixed(self, fixed):
        """Residual rows of fixed degrees of freedom are left out of the cost."""
        if fixed is None or len(fixed) == 0:
            self.free = None
        else:
            m = np.ones(3 * self.nn)
            m[fixed] = 0.0
            self.free = self.torch.from_numpy(m).to(self.dev)

    def set_shift(self, a0=None, a1=None):
        """Terms added to R in the cost, R + a0 + gamma a1: L'luvh and L'dl with linear
        constraints L x = cuvh (CalcCostFunctionNRuvh.m:47, frhs = -R - L'luvh)."""
        if a0 is None:
            self.shift = None
        else:
            t = self.torch
            self.shift = tuple(t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev) for a in (a0, a1))

    def _assemble(self, flags, K=None):
        self.ctx.assemble_device(self.ctrl, self.fields, flags, self.R, K, None, None, stream=self.stream)
        self.ctx.sync_status()

    def normalisation(self) -> float:
        """Fext0'Fext0 + 1000 eps: R of the zero-fields, R-only assembly at the current state."""
        self.xt.copy_(self.x)
        self._assemble(uvh.UVH_RONLY | uvh.UVH_ZEROFIELDS)
        self.n_r += 1
        return float(self.torch.dot(self.R, self.R).item()) + 1000 * EPS

    def assemble_rk_device(self, fp32: bool, fp32_residual: bool):
        """K (device, ``self.K``) and R(gamma = 0) (device, ``self.R0``) at the current state.
        With an FP32 Jacobian and an FP64 residual, the residual comes from an extra R-only
        assembly."""
        self.xt.copy_(self.x)
        self._assemble(uvh.UVH_FP32 if fp32 else 0, self.K)
        self.n_rk += 1
        if fp32 and not fp32_residual:
            self._assemble(uvh.UVH_RONLY)
            self.n_r += 1
        self.R0.copy_(self.R)

    def assemble_rk(self, fp32: bool, fp32_residual: bool):
        """As ``assemble_rk_device``; returns K (host CSR) and -R(gamma = 0) (host)."""
        self.assemble_rk_device(fp32, fp32_residual)
        return self.ctx.K_matrix(self.K.cpu().numpy()), (-self.R0).cpu().numpy()

    def step_on_device(self, solver):
        """d = K^-1 (-R0) with a device solver (``linsolve.GpuLinearSolver``), into ``self.d``."""
        rhs = self.torch.neg(self.R0)
        solver.factor(self.K, self.stream)
        solver.solve(rhs, self.d, self.stream)

    def set_step(self, d: np.ndarray):
        self.d.copy_(self.torch.from_numpy(d))

    def cost(self, gamma: float, fp32: bool, N: float):
        """(rForce, rWork, D2) at x + gamma d; the residual at gamma = 0 is the cached R0."""
        if gamma == 0.0:
            R = self.R0
        else:
            self.torch.add(self.x, self.d, alpha=gamma, out=self.xt)
            self._assemble(uvh.UVH_RONLY | (uvh.UVH_FP32 if fp32 else 0))
            self.n_r += 1
            R = self.R
        if self.shift is not None:
            R = (R + self.shift[0]).add_(self.shift[1], alpha=gamma)
        if self.free is not None:
            R = R * self.free
        frhs_frhs = self.torch.dot(R, R)
        D2 = -self.torch.dot(R, self.d)
        v = self.torch.stack([frhs_frhs, D2]).cpu().numpy()
        return v[0] / N, v[1] * v[1], v[1]

    def accept(self, gamma: float):
        self.x.add_(self.d, alpha=gamma)

    def state(self):
        x = self.x.cpu().numpy()
        n = self.nn
        return x[:n].copy(), x[n:2 * n].copy(), x[2 * n:].copy()

    def set_state(self, x: np.ndarray):
        """Newton state [ub; vb; h] (resuming a run)."""
        self.x.copy_(self.torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)))


class GpuUvAssembler:
    """R and K of the uv (diagnostic SSTREAM) system on the GPU (uvMatrixAssemblySSTREAM.m),
    with the geometry resident and the state x = [ub; vb]: the interface of ``GpuAssembler``
    for ``solve(..., system="uv")``. ``F`` carries h and the upper surface s besides the
    uvh fields. The normalisation is the zero-fields residual, with the velocities zero except
    at the fixed degrees of freedom, which keep their values (SSTREAM2dNR2.m:116-118,
    uvMatrixAssemblySSTREAM.m:24-28)."""

    def __init__(self, CtrlVar, MUA, F, device: int = 0, ctx: uvh.UvhContext | None = None):
        import torch
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        self.ctx = ctx or uvh.UvhContext(device)
        if ctx is None or ctx.mesh_key != uvh.mesh_key(MUA):
            self.ctx.set_mesh(MUA)
        self.mua = MUA
        self.nn = n = MUA.Nnodes
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)  # noqa: E731
        self.f = {k: t(getattr(F, k)) for k in uvh.UV_KEYS}
        self.x = torch.cat([self.f["ub"], self.f["vb"]])
        self.xt = self.x.clone()

        class _F:
            pass
        fd = _F()
        for k in uvh.UV_KEYS:
            setattr(fd, k, self.f[k])
        fd.ub, fd.vb = self.xt[:n], self.xt[n:]
        for k in ("q", "muk", "V0"):
            a = getattr(F, k, None)
            if a is not None and np.size(a) > 0:
                self.f[k] = t(a)
                setattr(fd, k, self.f[k])
        self.fields = uvh.make_uv_fields(fd)
        self.ctrl = uvh.make_ctrl(CtrlVar, F)
        self.uvc = uvh.make_uv_ctrl(CtrlVar)
        self.nnz = 4 * (self.ctx.nnz // 9)
        self.R = torch.zeros(2 * n, dtype=torch.float64, device=self.dev)
        self.R0 = torch.zeros_like(self.R)
        self.K = torch.zeros(self.nnz, dtype=torch.float64, device=self.dev)
        self.d = torch.zeros_like(self.R)
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        self.n_rk = self.n_r = 0
        self.free = None
        self.fixed_mask = torch.zeros_like(self.R)

    def set_fixed(self, fixed):
        m = np.ones(2 * self.nn)
        if fixed is not None and len(fixed):
            m[fixed] = 0.0
        self.free = None if fixed is None or len(fixed) == 0 else self.torch.from_numpy(m).to(self.dev)
        self.fixed_mask = self.torch.from_numpy(1.0 - m).to(self.dev)

    def set_shift(self, a0=None, a1=None):
        if a0 is not None:
            raise ValueError("the uv system takes no constraint terms")

    def _assemble(self, flags, K=None):
        self.ctx.assemble_uv_device(self.ctrl, self.uvc, self.fields, flags, self.R, K, stream=self.stream)
        self.ctx.sync_status()

    def normalisation(self) -> float:
        self.torch.mul(self.x, self.fixed_mask, out=self.xt)
        self._assemble(uvh.UVH_RONLY)
        self.n_r += 1
        return float(self.torch.dot(self.R, self.R).item()) + 1000 * EPS

    def assemble_rk_device(self, fp32: bool, fp32_residual: bool):
        """K (device, ``self.K``) and R(gamma = 0) (device, ``self.R0``) at the current state."""
        self.xt.copy_(self.x)
        self._assemble(uvh.UVH_FP32 if fp32 else 0, self.K)
        self.n_rk += 1
        if fp32 and not fp32_residual:
            self._assemble(uvh.UVH_RONLY)
            self.n_r += 1
        self.R0.copy_(self.R)

    def assemble_rk(self, fp32: bool, fp32_residual: bool):
        self.assemble_rk_device(fp32, fp32_residual)
        return self.ctx.uv_K_matrix(self.K.cpu().numpy()), (-self.R0).cpu().numpy()

    def uvh_layout_index(self):
        """Device copy of ``uv_to_uvh_index`` for this mesh, and its mask of uv values."""
        if getattr(self, "_uvh_index", None) is None:
            rowptr_n = self.ctx.pattern()[0][:self.nn + 1] // 3   # block row 0 starts rows at 3 rowptr_n
            idx = uv_to_uvh_index(rowptr_n)
            self._uvh_index = self.torch.from_numpy(idx).to(self.dev)
            self._uvh_mask = (self._uvh_index >= 0)
        return self._uvh_index, self._uvh_mask

    def device_solver(self, fixed_uv, rtol: float):
        """A ``linsolve.GpuLinearSolver`` for the uv steps: the uvh solver of the same mesh with
        the uv fixed dofs and every thickness dof held (their rows and columns are eliminated)."""
        from .linsolve import GpuLinearSolver
        n = self.nn
        held = np.concatenate([np.asarray(fixed_uv if fixed_uv is not None else [], dtype=np.int64),
                               np.arange(2 * n, 3 * n)])
        solver = GpuLinearSolver(self.ctx, held, rtol=rtol)
        self._Kuvh = self.torch.zeros(self.ctx.nnz, dtype=self.torch.float64, device=self.dev)
        self._rhs3 = self.torch.zeros(3 * n, dtype=self.torch.float64, device=self.dev)
        self._d3 = self.torch.zeros(3 * n, dtype=self.torch.float64, device=self.dev)
        self.uvh_layout_index()
        return solver

    def step_on_device(self, solver):
        """d = K^-1 (-R0) with the solver of ``device_solver``, into ``self.d``."""
        idx, mask = self._uvh_index, self._uvh_mask
        torch = self.torch
        torch.index_select(self.K, 0, idx.clamp(min=0), out=self._Kuvh)
        self._Kuvh.mul_(mask)
        n = self.nn
        self._rhs3[:2 * n].copy_(self.R0).neg_()
        solver.factor(self._Kuvh, self.stream)
        solver.solve(self._rhs3, self._d3, self.stream)
        self.d.copy_(self._d3[:2 * n])

    def set_step(self, d: np.ndarray):
        self.d.copy_(self.torch.from_numpy(d))

    def cost(self, gamma: float, fp32: bool, N: float):
        if gamma == 0.0:
            R = self.R0
        else:
            self.torch.add(self.x, self.d, alpha=gamma, out=self.xt)
            self._assemble(uvh.UVH_RONLY | (uvh.UVH_FP32 if fp32 else 0))
            self.n_r += 1
            R = self.R
        if self.free is not None:
            R = R * self.free
        frhs_frhs = self.torch.dot(R, R)
        D2 = -self.torch.dot(R, self.d)
        v = self.torch.stack([frhs_frhs, D2]).cpu().numpy()
        return v[0] / N, v[1] * v[1], v[1]

    def accept(self, gamma: float):
        self.x.add_(self.d, alpha=gamma)

    def state(self):
        x = self.x.cpu().numpy()
        n = self.nn
        return x[:n].copy(), x[n:].copy(), None

    def set_state(self, x: np.ndarray):
        self.x.copy_(self.torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)))


def uv_to_uvh_index(rowptr_n: np.ndarray) -> np.ndarray:
    """For each value of K in the uvh layout (9 blocks: row (bi, n) at bi 3 nnzN + 3 rowptr_n[n],
    blocks bj = 0..2 of deg(n) columns each), the index of the same entry in the uv layout
    (4 blocks: row (bi, n) at bi 2 nnzN + 2 rowptr_n[n]), or -1 in the thickness blocks."""
    rp = np.asarray(rowptr_n, dtype=np.int64)
    deg = np.diff(rp)
    nnzN = int(rp[-1])
    node = np.repeat(np.arange(len(deg)), deg)
    c = np.arange(nnzN) - rp[node]
    idx = np.full(9 * nnzN, -1, dtype=np.int64)
    for bi in range(2):
        for bj in range(2):
            idx[bi * 3 * nnzN + 3 * rp[node] + bj * deg[node] + c] = bi * 2 * nnzN + 2 * rp[node] + bj * deg[node] + c
    return idx


def uv_fixed(fixed_uvh: np.ndarray | None, nn: int):
    """The velocity degrees of freedom among uvh fixed dofs (indices into [ub; vb])."""
    if fixed_uvh is None:
        return None
    f = np.asarray(fixed_uvh)
    return f[f < 2 * nn]


def transient_start(CtrlVar, MUA, F0, F1, fixed, uv_assembler, linear_solver=superlu_solve, verbose=False,
                    log=print, device_solver: bool = False):
    """The start of a transient run (Ua2D.m:598-611, "always needed at a start of a transient
    run"): a uv (diagnostic) solve on F1's geometry, whose velocities then become both F0's and
    F1's, so that the first implicit uvh step starts from the balanced state. ``uv_assembler``
    builds the uv assembler from (CtrlVar, MUA, F) with F carrying the upper surface s (from the
    floatation of h). With ``device_solver`` the uv steps take the device solver (the uv
    assembler's ``device_solver``) instead of ``linear_solver``. Returns (F0, F1, uv NewtonResult)."""
    from .fields import surface_from_floatation
    F = F1.copy()
    F.s = surface_from_floatation(F1, CtrlVar)
    asm = uv_assembler(CtrlVar, MUA, F)
    fuv = uv_fixed(fixed, MUA.Nnodes)
    solver = asm.device_solver(fuv, CtrlVar.LinSolveTol) if device_solver else linear_solver
    try:
        res = solve(asm, CtrlVar, linear_solver=solver, fixed=fuv, verbose=verbose, log=log, system="uv")
    finally:
        if device_solver:
            solver.close()
    ub, vb, _ = res.ub, res.vb, None
    G0, G1 = F0.copy(), F1.copy()
    G0.ub, G0.vb = ub.copy(), vb.copy()
    G1.ub, G1.vb = ub.copy(), vb.copy()
    return G0, G1, res


# Lagrange shape functions in Úa's local node order (UaMain/shape_fun.m:63-82), as barycentric
# multi-indices (c1, c2, c3) with c1 = xi, c2 = eta, c3 = 1 - xi - eta
_NODE_INDEX = {
    3: [(1, 0, 0), (0, 0, 1), (0, 1, 0)],
    6: [(2, 0, 0), (1, 0, 1), (0, 0, 2), (0, 1, 1), (0, 2, 0), (1, 1, 0)],
    10: [(3, 0, 0), (2, 0, 1), (1, 0, 2), (0, 0, 3), (0, 1, 2), (0, 2, 1), (0, 3, 0), (1, 2, 0), (2, 1, 0), (1, 1, 1)],
}
_CORNERS = {3: (0, 1, 2), 6: (0, 2, 4), 10: (0, 3, 6)}   # nodes at c1 = 1, c3 = 1, c2 = 1


def shape_functions(nod: int, points: np.ndarray) -> np.ndarray:
    """N (npoints x nod) at (xi, eta) points: the degree-p Lagrange basis on the triangle,
    prod_i prod_{j < a_i} (p c_i - j) / (j + 1) for node multi-index a (|a| = p)."""
    p = {3: 1, 6: 2, 10: 3}[nod]
    c = np.stack([points[:, 0], points[:, 1], 1.0 - points[:, 0] - points[:, 1]], axis=1)
    N = np.ones((points.shape[0], nod))
    for k, a in enumerate(_NODE_INDEX[nod]):
        for i in range(3):
            for j in range(a[i]):
                N[:, k] *= (p * c[:, i] - j) / (j + 1)
    return N


def mass_matrix(MUA) -> sp.csr_matrix:
    """Nodal mass matrix M_ij = integral of N_i N_j (UaMain/MassMatrix2D1dof.m), with MUA's
    quadrature; the elements are straight-sided, so each element's matrix is its detJ times
    that of the reference triangle. Symmetrised as (M + M')/2, as there."""
    nod = MUA.nod
    N = shape_functions(nod, np.asarray(MUA.points, dtype=float))
    Mref = (N * np.asarray(MUA.weights, dtype=float)[:, None]).T @ N
    conn = np.asarray(MUA.connectivity, dtype=np.int64)
    x, y = MUA.coordinates[:, 0], MUA.coordinates[:, 1]
    a, b, c = _CORNERS[nod]
    detJ = (x[conn[:, a]] - x[conn[:, b]]) * (y[conn[:, c]] - y[conn[:, b]]) - \
           (y[conn[:, a]] - y[conn[:, b]]) * (x[conn[:, c]] - x[conn[:, b]])
    rows = np.repeat(conn, nod, axis=1).ravel()
    cols = np.tile(conn, (1, nod)).ravel()
    vals = (detJ[:, None] * Mref.ravel()[None, :]).ravel()
    n = MUA.Nnodes
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return ((M + M.T) * 0.5).tocsr()


class _CauchyM:
    """The Cauchy step with the mass matrix in place of the Hessian (rLineminUa.m:137-259):
    direction s = Mblock^-1 frhs on the free degrees of freedom (Mblock = diag(M, M, M)), the
    minimum of the quadratic model along it as the first trial, and backtracking with
    NewtonAcceptRatio 0.9 down to a step of 1e-10."""

    def __init__(self, MUA, fixed):
        self.nn = MUA.Nnodes
        M = mass_matrix(MUA)
        free = np.ones(3 * self.nn, dtype=bool)
        if fixed is not None and len(fixed):
            free[np.asarray(fixed, dtype=np.int64)] = False
        self.free = free
        self.lu = []
        for b in range(3):
            f = free[b * self.nn:(b + 1) * self.nn]
            self.lu.append((f, spla.splu(M[f][:, f].tocsc())))

    def direction(self, frhs):
        s = np.zeros(3 * self.nn)
        for b, (f, lu) in enumerate(self.lu):
            blk = slice(b * self.nn, (b + 1) * self.nn)
            sb = np.zeros(self.nn)
            sb[f] = lu.solve(frhs[blk][f])
            s[blk] = sb
        return s

    def step(self, CtrlVar, K, frhs, cost, r0, N):
        """(trial step along s that the backtracking returned, its cost triple, s, the quadratic
        model's step) or None when the model gives no usable slope."""
        free = self.free
        s = self.direction(frhs)
        Kf = K[free][:, free]
        Hs = Kf @ s[free]
        Rf = frhs[free]
        RHs = float(Rf @ Hs)
        HsHs = float(Hs @ Hs)
        if HsHs == 0.0:
            return None
        slope0 = -2.0 * RHs / N
        gC = RHs / HsHs
        if slope0 > 0:   # the other way along s (rLineminUa.m:180-199)
            s, RHs = -s, -RHs
            slope0, gC = -2.0 * RHs / N, RHs / HsHs
        if slope0 > 0:
            slope0 = -slope0
            gC = -0.1 * r0 / slope0
        if gC < 0:
            gC = -0.1 * r0 / slope0
        c1 = cost(gC, s)
        g, best, _ = _line_search(CtrlVar, lambda g: cost(g, s), r0, c1, slope0=slope0, b0=gC, gamma_min=1e-10,
                                  accept_ratio=0.9)
        return g, best, s, gC

    def steepest(self, CtrlVar, K, frhs, cost, r0, N):
        """The steepest-descent step (rLineminUa.m:333-400), tried where the Cauchy step gives no
        reduction: direction frhs on the free degrees of freedom (the identity in place of the
        mass matrix), the quadratic model's step as the first trial (a positive slope is negated,
        the direction kept, as there), backtracking with NewtonAcceptRatio 0.999 down to 1/1000
        of that step. BackTracking's extrapolation, allowed there, is not restated."""
        free = self.free
        s = np.where(free, frhs, 0.0)
        Hs = K[free][:, free] @ s[free]
        RHs = float(frhs[free] @ Hs)
        HsHs = float(Hs @ Hs)
        if HsHs == 0.0:
            return None
        slope0 = -2.0 * RHs / N
        gD = RHs / HsHs
        if slope0 > 0:
            slope0 = -slope0
        if gD < 0:
            gD = -0.1 * r0 / slope0
        c1 = cost(gD, s)
        g, best, _ = _line_search(CtrlVar, lambda g: cost(g, s), r0, c1, slope0=slope0, b0=gD, gamma_min=gD / 1000,
                                  accept_ratio=0.999)
        return g, best, s, gD


def _criteria(CtrlVar, gamma, rWork, rForce, iteration, system="uvh"):
    """Exit test of SSTREAM_TransientImplicit.m:269-287 (``system="uvh"``), or of
    uvResidualsCriteria.m for the uv system (the same form with the uv tolerances)."""
    tol = lambda name: CtrlVar.get(system + name)  # noqa: E731
    if gamma > max(tol("ExitBackTrackingStepLength"), CtrlVar.BacktrackingGammaMin):
        aw, af = tol("DesiredWorkAndForceTolerances")
        ow, of = tol("DesiredWorkOrForceTolerances")
    else:
        aw, af = tol("AcceptableWorkAndForceTolerances")
        ow, of = tol("AcceptableWorkOrForceTolerances")
    return (rWork < aw and rForce < af) and (rWork < ow or rForce < of) and iteration >= CtrlVar.NRitmin


def _line_search_steps(CtrlVar, r0, r1, slope0=None, b0=1.0, gamma_min=1e-3, accept_ratio=None):
    """Backtracking (BackTracking.m) from the trial step ``b0`` with value ``r1``, as a generator:
    it yields each further trial step g and is sent the cost (rForce, rWork, D2) at g; its
    return value is (gamma, (rForce, rWork, D2) at gamma, number of trial evaluations). The
    defaults are the Newton direction's (rLineminUa.m:60-135: slope -2 r0, first trial 1,
    floor 1e-3)."""
    if slope0 is None:
        slope0 = -2.0 * r0
    ratio = CtrlVar.NewtonAcceptRatio if accept_ratio is None else accept_ratio
    target = max(ratio * r0, min(r0 + CtrlVar.BackTrackBeta * slope0 * b0, ratio * r0))
    best_g, best = b0, r1
    if r1[0] < target:
        return b0, r1, 0
    lower, upper = CtrlVar.BackTrackGuardLower, CtrlVar.BackTrackGuardUpper
    b, fb = b0, r1[0]
    evals = 0
    for _ in range(CtrlVar.BackTrackMaxIterations):
        # minimum of the quadratic through f(0) = r0, f'(0) = slope0 and f(b) = fb (BackTracking.m:268)
        den = 2.0 * ((fb - r0) / b - slope0)
        g = -b * slope0 / den if den > 0 else lower * b
        g = min(max(g, lower * b), upper * b)
        if g < gamma_min:
            break
        rg = yield g
        evals += 1
        if rg[0] < best[0]:
            best_g, best = g, rg
        if rg[0] < target:
            break
        b, fb = g, rg[0]
    if best[0] >= r0:
        return 0.0, (r0, float("nan"), float("nan")), evals
    return best_g, best, evals


def _line_search(CtrlVar, cost, r0, r1, slope0=None, b0=1.0, gamma_min=1e-3, accept_ratio=None):
    """Backtracking with the cost function ``cost(g)``; see ``_line_search_steps``."""
    gen = _line_search_steps(CtrlVar, r0, r1, slope0, b0, gamma_min, accept_ratio)
    try:
        g = next(gen)
        while True:
            g = gen.send(cost(g))
    except StopIteration as stop:
        return stop.value


def eliminate(linear_solver, fixed: np.ndarray | None):
    """Wraps a linear solver for Dirichlet conditions on the degrees of freedom in ``fixed``
    (indices into [ub; vb; h]): the step is zero there and the rows and columns are removed.
    This is what Úa's Lagrange-multiplier form (``solveKApe`` with ``L``) gives at a feasible
    point, whose multipliers absorb the residual of the fixed rows."""
    if fixed is None or len(fixed) == 0:
        return linear_solver

    def solve_free(K, rhs):
        free = np.ones(K.shape[0], dtype=bool)
        free[fixed] = False
        d = np.zeros(K.shape[0])
        d[free] = linear_solver(K[free][:, free], rhs[free])
        return d
    return solve_free


def solve(assembler, CtrlVar, linear_solver=superlu_solve, fp32: bool = False, fp64_switch: float = 1e-8,
          fixed: np.ndarray | None = None, verbose: bool = False, log=print, time_limit: float | None = None,
          resume: dict | None = None, L=None, cuvh: np.ndarray | None = None,
          luvh: np.ndarray | None = None, system: str = "uvh") -> NewtonResult:
    """Newton-Raphson uvh iteration (SSTREAM_TransientImplicit.m:224-418). ``fixed``: degrees of
    freedom held at their current values (their residual rows are left out of the cost).
    ``time_limit`` (seconds): once at least one iteration has run in this call, stop before
    starting an iteration past it, with the loop state in ``NewtonResult.checkpoint``;
    ``resume`` (such a checkpoint) continues that run where it stopped, with the same result as
    one uninterrupted run.

    ``L`` (m x 3Nn sparse), ``cuvh`` (m): linear constraints L [ub; vb; h] = cuvh held with
    Lagrange multipliers ``luvh`` (start value, zero by default), as Úa does
    (SSTREAM_TransientImplicit.m:195, :373-383, :481): each step solves
    [K L'; L 0] [d; dl] = [-R - L'luvh; cuvh - L x] (solveKApe.m) with the host linear solver,
    and the cost carries both residual blocks (CalcCostFunctionNRuvh.m:45-85). The start need
    not be feasible.

    ``system="uv"``: the assembler is a ``GpuUvAssembler`` (state [ub; vb]) and the exit
    tests take the uv tolerances (SSTREAM2dNR2.m with uvResidualsCriteria.m, whose loop has the
    same normalisation, step, line search and exit order)."""
    on_device = getattr(linear_solver, "on_device", False)
    if system not in ("uvh", "uv"):
        raise ValueError(f"unknown system {system!r}")
    if system == "uv" and (L is not None or "-Auto-" in str(CtrlVar.get("rLineMinUa", ""))):
        raise ValueError("the uv system takes Dirichlet conditions and the Newton line search only")
    if L is not None:
        if on_device:
            raise ValueError("linear constraints L need the host linear solver")
        if fixed is not None and len(fixed):
            raise ValueError("give the boundary conditions either as fixed degrees of freedom or as L, not both")
        if "-Auto-" in str(CtrlVar.get("rLineMinUa", "")):
            raise ValueError('rLineMinUa "-Auto-" is not available with linear constraints L')
        L = sp.csr_matrix(L, dtype=np.float64)
        cuvh = np.asarray(cuvh, dtype=np.float64).ravel()
        if L.shape[1] != 3 * assembler.nn or cuvh.shape != (L.shape[0],):
            raise ValueError(f"L must be m x {3 * assembler.nn} and cuvh of length m; got {L.shape} and {cuvh.shape}")
        lam = np.zeros(L.shape[0]) if luvh is None else np.array(luvh, dtype=np.float64).ravel()
        LT = L.T.tocsr()
    if on_device:
        # the device solver holds its own fixed set (identity rows and columns); for the uv
        # system, which it solves in the uvh layout, that set also holds every thickness dof
        want = np.unique(np.asarray(fixed if fixed is not None else [], dtype=np.int64))
        if system == "uv":
            want = np.unique(np.concatenate([want, np.arange(2 * assembler.nn, 3 * assembler.nn)]))
        have = linear_solver.fixed
        same = (have is None and want.size == 0) or (have is not None and np.array_equal(have, want))
        if not same:
            raise ValueError("the device linear solver was built for a different set of fixed degrees of freedom")
    else:
        linear_solver = eliminate(linear_solver, fixed)
    cauchy = None
    if "-Auto-" in str(CtrlVar.get("rLineMinUa", "")):
        if on_device:
            raise ValueError('rLineMinUa "-Auto-" needs K on the host (the host linear solver)')
        cauchy = _CauchyM(assembler.mua, fixed)
    assembler.set_fixed(fixed)
    assembler.set_shift(None)   # no constraint terms left from a solve that ended in an error
    N = assembler.normalisation()
    res = NewtonResult(False, 0, "", float("inf"), float("nan"))
    iteration, gamma = 0, 1.0
    r, r0 = float("inf"), float("inf")
    rForce, rWork = float("inf"), float("nan")
    residual64 = not fp32
    ratios = np.zeros(3)
    short32 = 0
    backtrack_ok = True
    if resume is not None:
        iteration, gamma, r, r0 = int(resume["iteration"]), float(resume["gamma"]), float(resume["r"]), float(resume["r0"])
        rForce, rWork = float(resume["rForce"]), float(resume["rWork"])
        residual64, backtrack_ok = bool(resume["residual64"]), bool(resume["backtrack_ok"])
        ratios = np.array(resume["ratios"], dtype=float)
        short32 = int(resume.get("short32", 0))
        res.history = [tuple(h) for h in resume["history"]]
        assembler.set_state(resume["x"])
        if L is not None:
            lam = np.array(resume["luvh"], dtype=np.float64)
    t_start = time.perf_counter()
    done_here = 0
    while True:
        # rWork is NaN after a Cauchy or steepest-descent point (no work measure there); it is
        # then taken as 0, so that the exit and stagnation tests judge such a point on rForce
        # alone (SSTREAM_TransientImplicit.m:275-281, 345-350)
        rw = 0.0 if np.isnan(rWork) else rWork
        if residual64 and _criteria(CtrlVar, gamma, rw, rForce, iteration, system):
            res.converged, res.reason = True, "converged"
            break
        if iteration > CtrlVar.NRitmax:
            res.reason = "NRitmax"
            break
        if not backtrack_ok or gamma == 0:
            if fp32 and not residual64:
                # the FP32 residual has reached its rounding floor: continue with the FP64 one
                residual64, backtrack_ok, gamma = True, True, 1.0
                r, r0, ratios[:] = float("inf"), float("inf"), 0.0
                rForce, rWork = float("inf"), float("nan")   # no exit test on an FP32 residual
                continue
            res.reason = "backtracking stagnated"
            break
        ratios[iteration % 3] = r / r0 if np.isfinite(r0) and r0 > 0 else 0.0
        if np.all(ratios > 0.98):
            if fp32 and not residual64:
                residual64, gamma = True, 1.0
                r, r0, ratios[:] = float("inf"), float("inf"), 0.0
                rForce, rWork = float("inf"), float("nan")
                continue
            res.converged = rw < 1e-14
            res.reason = "stagnated"
            break
        if time_limit is not None and done_here > 0 and time.perf_counter() - t_start > time_limit:
            res.reason = "time limit"
            res.checkpoint = dict(iteration=iteration, gamma=gamma, r=r, r0=r0, rForce=rForce, rWork=rWork,
                                  residual64=residual64, backtrack_ok=backtrack_ok, ratios=ratios.copy(), short32=short32,
                                  history=list(res.history), x=np.concatenate(assembler.state()))
            if L is not None:
                res.checkpoint["luvh"] = lam.copy()
            break
        iteration += 1
        done_here += 1
        if fp32 and not residual64 and (rForce < fp64_switch or short32 >= 2):
            # from here on the residual, the Jacobian, the step and the exit test are FP64
            residual64 = True
        fp32_res = fp32 and not residual64   # the Jacobian too: FP32 only while the residual is
        t0 = time.perf_counter()
        if on_device:
            assembler.assemble_rk_device(fp32_res, fp32_res)
            t1 = time.perf_counter()
            assembler.step_on_device(linear_solver)
            t2 = time.perf_counter()
            res.linear_iterations.append(linear_solver.last_iterations)
        elif L is not None:
            K, rhs = assembler.assemble_rk(fp32_res, fp32_res)
            t1 = time.perf_counter()
            n3 = K.shape[0]
            G0 = cuvh - L @ np.concatenate(assembler.state())
            sol = linear_solver(sp.bmat([[K, LT], [L, None]], format="csr"),
                                np.concatenate([rhs - LT @ lam, G0]))
            d, dl = sol[:n3], sol[n3:]
            t2 = time.perf_counter()
            assembler.set_step(d)
            assembler.set_shift(LT @ lam, LT @ dl)
            G1 = L @ d
        else:
            K, rhs = assembler.assemble_rk(fp32_res, fp32_res)
            t1 = time.perf_counter()
            d = linear_solver(K, rhs)
            t2 = time.perf_counter()
            assembler.set_step(d)
        if L is None:
            cost = lambda g: assembler.cost(g, fp32_res, N)
        else:
            def cost(g):
                # the assembler's terms hold frhs = -(R + L'(luvh + g dl)); the constraint block
                # grhs = cuvh - L (x + g d) is added here
                c = assembler.cost(g, fp32_res, N)
                gr = G0 - g * G1
                D2 = c[2] + float(gr @ dl)
                return c[0] + float(gr @ gr) / N, D2 * D2, D2
        c0 = cost(0.0)
        c1 = cost(1.0)
        gamma, best, _ = _line_search(CtrlVar, cost, c0[0], c1)
        step = gamma   # the length taken along the direction the assembler holds
        rmin = best[0] if gamma > 0 else c0[0]
        if cauchy is not None and gamma < 0.75 and rmin / c0[0] > 0.7:
            # rLineminUa.m:94-103: the Newton step fell short; try the Cauchy step with the mass matrix
            def cost_along(g, s):
                assembler.set_step(s)
                return assembler.cost(g, fp32_res, N)
            out = cauchy.step(CtrlVar, K, rhs, cost_along, c0[0], N)
            chosen = d
            if out is None or out[1][0] >= c0[0]:
                # no reduction along the mass-matrix direction: steepest descent (rLineminUa.m:320),
                # whose point then takes the Cauchy point's place if it is the better one
                sd = cauchy.steepest(CtrlVar, K, rhs, cost_along, c0[0], N)
                if sd is not None and (out is None or sd[1][0] < out[1][0]):
                    out = sd
            if out is not None:
                gC, bestC, s, gModel = out
                if bestC[0] < rmin:
                    chosen, rmin = s, bestC[0]
                    step, best = gC, (bestC[0], float("nan"), float("nan"))
                    gamma = gC / gModel   # returned normalised by the quadratic model's step, as there
                    res.cauchy_steps += 1
                if bestC[0] < c0[0] and rmin / c0[0] >= 0.7:
                    # rLineminUa.m:270-300, :420-470: from the Cauchy point towards the full Newton
                    # step, x + xC + t (xN - xC) for t in [0, 1], minimised by Brent's bounded method
                    # (MATLAB's fminbnd, TolX 0.02)
                    from scipy.optimize import minimize_scalar
                    xC = gC * s

                    def along(t):
                        assembler.set_step(xC + t * (d - xC))
                        return assembler.cost(1.0, fp32_res, N)[0]
                    o = minimize_scalar(along, bounds=(0.0, 1.0), method="bounded", options={"xatol": 0.02})
                    if o.fun < rmin:
                        chosen = xC + float(o.x) * (d - xC)
                        assembler.set_step(chosen)
                        step, best, gamma = 1.0, assembler.cost(1.0, fp32_res, N), float(o.x)
                        rmin = best[0]
                        res.cauchy_steps += 1
            assembler.set_step(chosen)
        backtrack_ok = gamma > 0
        # two FP32 steps in a row shorter than 0.25: the FP32 Jacobian no longer gives Newton-
        # quality steps, and the next iteration switches to FP64
        short32 = short32 + 1 if (fp32_res and gamma < 0.25) else 0
        if gamma > 0:
            assembler.accept(step)
            if L is not None:
                lam = lam + step * dl
            rForce, rWork = best[0], best[1]
            r = rForce if CtrlVar.uvhMinimisationQuantity == "Force Residuals" else rWork
        r0 = c0[0] if CtrlVar.uvhMinimisationQuantity == "Force Residuals" else c0[1]
        t3 = time.perf_counter()
        res.t_assembly += t1 - t0
        res.t_solve += t2 - t1
        res.t_linesearch += t3 - t2
        res.history.append((iteration, gamma, rForce, rWork, not fp32_res))
        if verbose:
            lin = (f" Krylov {res.linear_iterations[-1]} relres {linear_solver.last_relres:.1e}"
                   + (f" [{linear_solver.last_message}]" if linear_solver.last_message else "")) if on_device else ""
            log(f"  it {iteration}: gamma {gamma:.3g} rForce {rForce:.3e} rWork {rWork:.3e} "
                f"(assembly {t1 - t0:.3f} s, solve {t2 - t1:.3f} s{lin})"
                f"{'' if not fp32 else (' (fp64 R)' if not fp32_res else ' (fp32 R)')}")
    res.iterations, res.rForce, res.rWork = iteration, rForce, rWork
    res.n_rk, res.n_r = assembler.n_rk, assembler.n_r
    res.ub, res.vb, res.h = assembler.state()
    if L is not None:
        assembler.set_shift(None)
        res.luvh = lam
    return res
