"""Newton-Raphson uvh solves of many independent meshes together, with the iterations'
device work captured in HIP graphs.

Each member is its own problem (mesh, F0, F1, fixed degrees of freedom), solved as
``newton.solve`` solves one (``UaMain/SSTREAM_TransientImplicit.m:224-418``: the zero-fields
normalisation, the cost of ``CalcCostFunctionNRuvh.m:22-98``, Úa's exit tests, the Newton step
and the backtracking of ``BackTracking.m``), but the members step together:

* graph A: the trial state of every member set to its state, then the R and K of every member
  in one batched assembly (``uvhb_assemble_device``);
* the step ``K d = -R`` of every member on the device: each member's block ILU(0)
  (``uvh_lin_factor``), then one batched BiCGStab for all members (``uvhb_lin_*``: one launch
  per step of the iteration for every member, convergence, restarts and the iteration limit
  per member on the device); graph S holds ``KRYLOV_CHUNK`` iterations of every member and is
  replayed until every member has stopped (an iteration after a member's stop does nothing
  for it); a member that misses the tolerance is solved again alone with BiCGStab and the
  GMRES fall-back, as a single mesh is;
* graph C: the cost at gamma = 0 of every member (|w R|^2 and -w R . d, ``uvhb_cost_device``);
* graph B: the trial state x + gamma_m d_m of every member, with each member's gamma in a
  device vector, its R (one batched R-only assembly) and its cost.

The line search of every member runs as ``newton._line_search_steps``; one replay of graph B
evaluates the trial steps of all members that are still searching. A member that has met its
exit test keeps its state and drops out of the solves and line searches. With
``graphs=False`` the same calls run eagerly: the results are the same bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import newton, uvh
from .batch import UvhBatch
from .linsolve import GpuBatchLinearSolver, GpuLinearSolver

EPS = np.finfo(float).eps
KRYLOV_CHUNK = 8   # BiCGStab iterations per replay of graph S (per eager call in the eager run)


class BatchNewton:
    def __init__(self, CtrlVar, members, device: int = 0, graphs: bool = True, batched_solve: bool = True):
        """``members``: list of (MUA, F0, F1, fixed) with ``fixed`` indices into [ub; vb; h]
        (or None). The members share CtrlVar and F1's scalars g, rhow and alpha. With
        ``batched_solve=False`` the members' Newton steps are solved one after another."""
        import torch
        self.torch = torch
        self.cv = CtrlVar
        self.dev = torch.device(f"cuda:{device}")
        uvh.check_supported(CtrlVar)
        self.graphs = graphs
        self.n = len(members)
        if self.n == 0:
            raise ValueError("no members")
        F1a = members[0][2]
        for (_, _, F1, _) in members[1:]:
            if (float(F1.g), float(F1.rhow), float(F1.alpha)) != (float(F1a.g), float(F1a.rhow), float(F1a.alpha)):
                raise ValueError("the members of one batch take the same g, rhow and alpha")
        self.batch = UvhBatch(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)  # noqa: E731
        self.x, self.xt, self.d, self.R, self.K, self.w, self.rhs = [], [], [], [], [], [], []
        self.fields, self.keep, self.fixed, self.nn = [], [], [], []
        for (M, F0, F1, fixed) in members:
            self.batch.add_mesh(M)
        for (M, F0, F1, fixed), mv in zip(members, self.batch.members):
            nn = M.Nnodes
            x = torch.cat([t(F1.ub), t(F1.vb), t(F1.h)])
            xt = x.clone()
            f1, f0 = type("F", (), {})(), type("F", (), {})()
            trial = {"ub": xt[:nn], "vb": xt[nn:2 * nn], "h": xt[2 * nn:]}
            for k in uvh.F1_KEYS:
                setattr(f1, k, trial[k] if k in trial else t(getattr(F1, k)))
                self.keep.append(getattr(f1, k))
            for k in uvh.F0_KEYS:
                setattr(f0, k, t(getattr(F0, k)))
                self.keep.append(getattr(f0, k))
            for k in ("q", "muk", "V0"):   # read by the sliding laws other than Weertman
                a = getattr(F1, k, None)
                if a is not None and np.size(a) > 0:
                    setattr(f1, k, t(a))
                    self.keep.append(getattr(f1, k))
            self.fields.append(uvh.make_fields(f0, f1))
            w = torch.ones(3 * nn, dtype=torch.float64, device=self.dev)
            if fixed is not None and len(fixed):
                w[torch.from_numpy(np.asarray(fixed, dtype=np.int64))] = 0.0
            self.x.append(x)
            self.xt.append(xt)
            self.d.append(torch.zeros_like(x))
            self.rhs.append(torch.zeros_like(x))
            self.R.append(torch.zeros_like(x))
            self.K.append(torch.zeros(mv.nnz, dtype=torch.float64, device=self.dev))
            self.w.append(w)
            self.fixed.append(fixed)
            self.nn.append(nn)
        self.ctrl = uvh.make_ctrl(CtrlVar, members[0][2])
        self.gamma = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.rr = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.d2 = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.solvers = [GpuLinearSolver(mv, fx, rtol=CtrlVar.LinSolveTol) for mv, fx in zip(self.batch.members, self.fixed)]
        self.bsolver = GpuBatchLinearSolver(self.solvers, rtol=CtrlVar.LinSolveTol) if batched_solve else None
        self.stream = torch.cuda.Stream(self.dev)   # default priority
        self.sptr = self.stream.cuda_stream
        self.n_graph_replays = 0
        self._graphs = None

    def close(self):
        if getattr(self, "bsolver", None) is not None:
            self.bsolver.close()
            self.bsolver = None
        for s in self.solvers:
            s.close()
        self.solvers = []
        self.batch.close()

    # ---------------------------------------------------------------- the captured stages
    def _stage_assemble_rk(self):
        for x, xt in zip(self.x, self.xt):
            xt.copy_(x)
        self.batch.assemble_device(self.ctrl, self.fields, 0, self.R, self.K, stream=self.sptr)

    def _stage_cost(self):
        self.batch.cost_device(self.R, self.d, self.w, self.rr, self.d2, stream=self.sptr)

    def _stage_trial(self):
        for m in range(self.n):
            self.torch.addcmul(self.x[m], self.d[m], self.gamma[m:m + 1], out=self.xt[m])
        self.batch.assemble_device(self.ctrl, self.fields, uvh.UVH_RONLY, self.R, self.K, stream=self.sptr)
        self._stage_cost()

    def _run(self, name):
        torch = self.torch
        stage = {"A": self._stage_assemble_rk, "B": self._stage_trial, "C": self._stage_cost}[name]
        with torch.cuda.stream(self.stream):
            if self.graphs and self._graphs is not None:
                self._graphs[name].replay()
                self.n_graph_replays += 1
            else:
                stage()

    def _capture(self):
        """Eager runs of every stage first (the batch's member and cost tables are uploaded
        then, so that nothing is uploaded inside a capture), then one graph per stage."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            self._stage_assemble_rk()
            self._stage_cost()
            self._stage_trial()
        self.stream.synchronize()
        self._graphs = {}
        for name, stage in (("A", self._stage_assemble_rk), ("B", self._stage_trial), ("C", self._stage_cost)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self.stream):
                stage()
            self._graphs[name] = g
        self.stream.synchronize()

    def _krylov(self, active):
        """The batched BiCGStab of the active members' steps: graph S (captured at the first
        call, inside a begun solve) or the same iterations eagerly, until every member stops."""
        torch, bs = self.torch, self.bsolver
        bs.begin(self.rhs, active=active, stream=self.sptr)
        with torch.cuda.stream(self.stream):
            if self.graphs and "S" not in self._graphs:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    bs.iterate(KRYLOV_CHUNK, self.sptr)
                self._graphs["S"] = g
            while True:
                if self.graphs:
                    self._graphs["S"].replay()
                    self.n_graph_replays += 1
                else:
                    bs.iterate(KRYLOV_CHUNK, self.sptr)
                self.stream.synchronize()
                if bs.stopped():
                    break
        return bs.end(self.d, stream=self.sptr)

    def _status(self):
        self.stream.synchronize()
        st = self.batch.sync_status()
        if any(st):
            raise uvh.UaError("uvhMatrixAssembly:NaNinR" if any(v == uvh.UVH_E_NAN_R for v in st) else "",
                              f"NaN in the assembly of members {[i for i, v in enumerate(st) if v]}")

    def normalisation(self):
        """Fext0'Fext0 + 1000 eps per member: R of the zero-fields R-only assembly (eager)."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for x, xt in zip(self.x, self.xt):
                xt.copy_(x)
            self.batch.assemble_device(self.ctrl, self.fields, uvh.UVH_RONLY | uvh.UVH_ZEROFIELDS, self.R, self.K,
                                       stream=self.sptr)
            self._status()
            return np.array([float(torch.dot(R, R).item()) + 1000 * EPS for R in self.R])

    # ---------------------------------------------------------------- the iteration
    def solve(self):
        """Runs every member to its exit; returns one ``newton.NewtonResult`` per member."""
        with self.torch.cuda.stream(self.stream):   # every torch op in the order of the library calls
            return self._solve()

    def _solve(self):
        torch, cv = self.torch, self.cv
        force = cv.uvhMinimisationQuantity == "Force Residuals"
        N = self.normalisation()
        if self.graphs and self._graphs is None:
            self._capture()
        res = [newton.NewtonResult(False, 0, "", float("inf"), float("nan")) for _ in range(self.n)]
        it = np.zeros(self.n, dtype=int)
        gamma = np.ones(self.n)
        r, r0 = np.full(self.n, np.inf), np.full(self.n, np.inf)
        rForce, rWork = np.full(self.n, np.inf), np.full(self.n, np.nan)
        ratios = np.zeros((self.n, 3))
        backtrack_ok = np.ones(self.n, dtype=bool)
        active = np.ones(self.n, dtype=bool)
        while True:
            # exit tests of the members still iterating (SSTREAM_TransientImplicit.m:269-350)
            for m in np.nonzero(active)[0]:
                rw = 0.0 if np.isnan(rWork[m]) else rWork[m]
                if newton._criteria(cv, gamma[m], rw, rForce[m], it[m]):
                    res[m].converged, res[m].reason, active[m] = True, "converged", False
                elif it[m] > cv.NRitmax:
                    res[m].reason, active[m] = "NRitmax", False
                elif not backtrack_ok[m] or gamma[m] == 0:
                    res[m].reason, active[m] = "backtracking stagnated", False
                else:
                    ratios[m, it[m] % 3] = r[m] / r0[m] if np.isfinite(r0[m]) and r0[m] > 0 else 0.0
                    if np.all(ratios[m] > 0.98):
                        res[m].converged, res[m].reason, active[m] = rw < 1e-14, "stagnated", False
            if not active.any():
                break
            act = np.nonzero(active)[0]
            it[act] += 1
            self._run("A")
            self._status()
            for m in act:   # the Newton steps: every member's factorization, then the solves
                torch.neg(self.R[m], out=self.rhs[m])
                self.solvers[m].factor(self.K[m], self.sptr)
            if self.bsolver is not None:
                its, _, conv = self._krylov(active)
                for m in act:
                    n_it = int(its[m])
                    if not conv[m]:   # BiCGStab again alone, then GMRES from its best iterate
                        self.solvers[m].solve(self.rhs[m], self.d[m], self.sptr)
                        n_it += self.solvers[m].last_iterations
                    res[m].linear_iterations.append(n_it)
            else:
                for m in act:
                    self.solvers[m].solve(self.rhs[m], self.d[m], self.sptr)
                    res[m].linear_iterations.append(self.solvers[m].last_iterations)
            for m in np.nonzero(~active)[0]:
                self.d[m].zero_()   # members that are done stay where they are
            self._run("C")
            c0 = self._costs(N)
            self.gamma.fill_(1.0)
            self._run("B")
            self._status()
            c1 = self._costs(N)
            searches = {m: newton._line_search_steps(cv, c0[m][0], c1[m]) for m in act}
            trial, done = {}, {}
            for m, gen in searches.items():
                try:
                    trial[m] = next(gen)
                except StopIteration as stop:
                    done[m] = stop.value
            while trial:   # one replay of graph B per round evaluates every member's trial step
                g = np.ones(self.n)
                for m, gm in trial.items():
                    g[m] = gm
                self.gamma.copy_(torch.from_numpy(g))
                self._run("B")
                self._status()
                cg = self._costs(N)
                nxt = {}
                for m in trial:
                    try:
                        nxt[m] = searches[m].send(cg[m])
                    except StopIteration as stop:
                        done[m] = stop.value
                trial = nxt
            for m in act:
                gm, best, _ = done[m]
                gamma[m] = gm
                backtrack_ok[m] = gm > 0
                if gm > 0:
                    self.x[m].add_(self.d[m], alpha=gm)
                    rForce[m], rWork[m] = best[0], best[1]
                    r[m] = rForce[m] if force else rWork[m]
                r0[m] = c0[m][0] if force else c0[m][1]
                res[m].history.append((int(it[m]), float(gm), float(rForce[m]), float(rWork[m]), True))
        self.stream.synchronize()
        for m in range(self.n):
            res[m].iterations, res[m].rForce, res[m].rWork = int(it[m]), float(rForce[m]), float(rWork[m])
            x = self.x[m].cpu().numpy()
            nn = self.nn[m]
            res[m].ub, res[m].vb, res[m].h = x[:nn].copy(), x[nn:2 * nn].copy(), x[2 * nn:].copy()
        return res

    def _costs(self, N):
        """(rForce, rWork, D2) of every member from the last cost stage."""
        self.stream.synchronize()
        rr = self.rr.cpu().numpy()
        d2 = self.d2.cpu().numpy()
        return [(rr[m] / N[m], d2[m] * d2[m], d2[m]) for m in range(self.n)]
