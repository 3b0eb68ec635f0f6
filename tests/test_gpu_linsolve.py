"""Newton-step linear solve on the GPU (uvh_lin_*, uasource_amd.linsolve): the device form of
solveKApe's 'EliminateBCsSolveSystemIterativly' branch (UaMain/solveKApe.m:117-151).

Checked against SciPy's SuperLU on the eliminated system (rows and columns of the fixed
degrees of freedom removed, as solveKApe.m:126-129 does): at rtol 1e-12 the solution agrees
to 1e-8 (the solve's residual bound times the conditioning of these systems) and the true
residual, recomputed on the host, is below 2e-12; fixed degrees of freedom are exactly zero;
a solve is bit for bit reproducible; a zero right-hand side gives zero in no iterations; the
iteration limit and the call order are enforced. Each Krylov method (BiCGStab, GMRES(m) as in
the reference branch, and the default BiCGStab with a GMRES fall-back) meets the same bars, and
so do the triangular solves on the FP32-rounded factors (the default) and on the FP64 factors
(UVH_LIN_FACTOR=fp64). The Newton iteration with the device solver
(LinSolveTol = 1e-8, Úa's default) converges like the oracle's Newton with a direct solve:
iteration counts within one, rForce < 1e-15, the same u, v, h."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from uasource_amd import newton, uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, newton_problem
from uasource_amd.linsolve import GpuLinearSolver, LinearSolveError

pytestmark = pytest.mark.gpu

_sys = {}


def system(name):
    """(M, fixed, ctx, host K, host -R, device K, device -R) of the Newton problem's first step."""
    if name not in _sys:
        import torch
        w = WORKLOADS[name]
        M = build_mesh(w)
        F0, F1, _, fixed = newton_problem(M, w)
        ctx = uvh.UvhContext(0)
        ctx.set_mesh(M)
        out = ctx.assemble(default_ctrlvar(), F0, F1)
        K = ctx.K_matrix(out["Kval"])
        rhs = -out["R"]
        dev = torch.device("cuda:0")
        _sys[name] = (M, fixed, ctx, K, rhs, torch.from_numpy(out["Kval"]).to(dev), torch.from_numpy(rhs).to(dev))
    return _sys[name]


def direct(K, rhs, fixed):
    free = np.ones(K.shape[0], dtype=bool)
    free[fixed] = False
    x = np.zeros(K.shape[0])
    x[free] = spla.splu(K[free][:, free].tocsc(), permc_spec="COLAMD").solve(rhs[free])
    return x, free


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2", "newton_p2"])
def test_solution_matches_direct_solve(name):
    import torch
    print(f"[gpu] linear solve {name}", flush=True)
    M, fixed, ctx, K, rhs, dK, drhs = system(name)
    solver = GpuLinearSolver(ctx, fixed, rtol=1e-12, maxit=20000)
    try:
        x = torch.empty_like(drhs)
        solver.factor(dK)
        its, relres = solver.solve(drhs, x)
        torch.cuda.synchronize()
        xg = x.cpu().numpy()
        xd, free = direct(K, rhs, fixed)
        err = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
        Kf = K[free][:, free]
        true_res = np.linalg.norm(rhs[free] - Kf @ xg[free]) / np.linalg.norm(rhs[free])
        print(f"   {len(rhs)} unknowns: {its} iterations, relres {relres:.2e} (host {true_res:.2e}), err {err:.2e}",
              flush=True)
        assert its > 0
        assert true_res < 2e-12 and relres < 2e-12
        # both residuals are formed in FP64 with different summation orders: they agree to a
        # few percent, and absolutely to the rounding level of K x (~1e-15 relative)
        assert abs(true_res - relres) <= 0.02 * true_res + 1e-14
        assert err < 1e-8
        assert np.all(xg[fixed] == 0.0)
    finally:
        solver.close()


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
@pytest.mark.parametrize("name", ["tiny_p3", "medium_p2"])
def test_each_krylov_method_matches_direct_solve(name, method):
    import torch
    print(f"[gpu] linear solve {name} {method}", flush=True)
    M, fixed, ctx, K, rhs, dK, drhs = system(name)
    solver = GpuLinearSolver(ctx, fixed, rtol=1e-12, maxit=20000, method=method, restart=40)
    try:
        x = torch.empty_like(drhs)
        solver.factor(dK)
        its, relres = solver.solve(drhs, x)
        xg = x.cpu().numpy()
        xd, free = direct(K, rhs, fixed)
        err = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
        print(f"   {its} iterations, relres {relres:.2e}, err {err:.2e} {solver.last_message}", flush=True)
        assert solver.last_message == "" and relres < 2e-12 and err < 1e-8
        y = torch.empty_like(drhs)
        solver.solve(drhs, y)
        assert torch.equal(x, y)
    finally:
        solver.close()


@pytest.mark.parametrize("factor", ["fp32", "fp64"])
@pytest.mark.parametrize("name", ["tiny_p3", "medium_p2"])
def test_factor_precisions_match_direct_solve(name, factor, monkeypatch):
    import torch
    print(f"[gpu] linear solve {name}, {factor} factors", flush=True)
    monkeypatch.setenv("UVH_LIN_FACTOR", factor)
    M, fixed, ctx, K, rhs, dK, drhs = system(name)
    solver = GpuLinearSolver(ctx, fixed, rtol=1e-12, maxit=20000, method="bicgstab")
    try:
        x = torch.empty_like(drhs)
        solver.factor(dK)
        its, relres = solver.solve(drhs, x)
        xg = x.cpu().numpy()
        xd, free = direct(K, rhs, fixed)
        err = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
        print(f"   {its} iterations, relres {relres:.2e}, err {err:.2e} {solver.last_message}", flush=True)
        assert solver.last_message == "" and relres < 2e-12 and err < 1e-8
        assert np.all(xg[fixed] == 0.0)
    finally:
        solver.close()


def test_batched_solve_matches_direct_solves():
    """uvhb_lin_solve over members of three element types and two sizes: every active member
    meets the single solve's bars, an inactive member's output is left alone, a zero right-hand
    side gives zero in no iterations, and the batched solve is bit for bit reproducible."""
    import torch
    from uasource_amd.linsolve import GpuBatchLinearSolver
    names = ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2", "newton_p2"]
    print(f"[gpu] batched linear solve {names}", flush=True)
    systems = [system(nm) for nm in names]
    solvers = [GpuLinearSolver(sy[2], sy[1], rtol=1e-12, maxit=20000, method="bicgstab") for sy in systems]
    batch = None
    try:
        for sv, sy in zip(solvers, systems):
            sv.factor(sy[5])
        batch = GpuBatchLinearSolver(solvers, rtol=1e-12, maxit=20000)
        rhs = [sy[6] for sy in systems]
        rhs[1] = torch.zeros_like(rhs[1])
        xs = [torch.full_like(r, 7.0) for r in rhs]
        active = np.array([1, 1, 1, 0, 1], dtype=np.uint8)
        its, rel, conv = batch.solve(rhs, xs, active=active)
        torch.cuda.synchronize()
        print(f"   iterations {its.tolist()}, relres {[f'{r:.1e}' for r in rel]}, converged {conv.tolist()}",
              flush=True)
        assert batch.last_message == ""
        assert conv.tolist() == [1, 1, 1, 0, 1]
        assert its[1] == 0 and torch.count_nonzero(xs[1]).item() == 0
        assert its[3] == 0 and torch.all(xs[3] == 7.0)
        for m in (0, 2, 4):
            M, fixed, ctx, K, r, dK, drhs = systems[m]
            xg = xs[m].cpu().numpy()
            xd, free = direct(K, r, fixed)
            err = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
            true_res = np.linalg.norm(r[free] - K[free][:, free] @ xg[free]) / np.linalg.norm(r[free])
            print(f"   {names[m]}: {its[m]} iterations, relres {rel[m]:.2e} (host {true_res:.2e}), err {err:.2e}",
                  flush=True)
            assert its[m] > 0 and rel[m] < 2e-12 and true_res < 2e-12 and err < 1e-8
            assert np.all(xg[fixed] == 0.0)
            # the single solver on the same system takes about as many iterations (the
            # reductions differ in shape, so the rounding differs)
            y = torch.empty_like(drhs)
            its1, _ = solvers[m].solve(drhs, y)
            assert abs(its1 - its[m]) <= max(3, 0.25 * its1)
        ys = [torch.full_like(r, 7.0) for r in rhs]
        its2, rel2, conv2 = batch.solve(rhs, ys, active=active)
        assert its2.tolist() == its.tolist() and rel2.tolist() == rel.tolist()
        assert all(torch.equal(a, b) for a, b in zip(xs, ys))
    finally:
        if batch is not None:
            batch.close()
        for sv in solvers:
            sv.close()


def test_batched_solve_limit_order_and_steps():
    """The batched solve's iteration limit (every member stops at maxit with its best iterate and
    is reported as not converged), its call order (iterate/end before begin refused), and the
    solve in steps (begin, iterate replayed until stopped, end) equal to one uvhb_lin_solve
    bit for bit."""
    import torch
    from uasource_amd import uvh as U
    from uasource_amd.linsolve import GpuBatchLinearSolver
    names = ["tiny_p2", "medium_p2"]
    print(f"[gpu] batched linear solve: limit, call order, steps {names}", flush=True)
    systems = [system(nm) for nm in names]
    solvers = [GpuLinearSolver(sy[2], sy[1], rtol=1e-10, maxit=20000, method="bicgstab") for sy in systems]
    batch = None
    try:
        for sv, sy in zip(solvers, systems):
            sv.factor(sy[5])
        batch = GpuBatchLinearSolver(solvers, rtol=1e-10, maxit=20000)
        rhs = [sy[6] for sy in systems]
        with pytest.raises(U.UvhLibraryError):
            batch.iterate(4)
        xs = [torch.empty_like(r) for r in rhs]
        its, rel, conv = batch.solve(rhs, xs, maxit=3)
        print(f"   maxit 3: iterations {its.tolist()}, relres {rel.tolist()}, {batch.last_message}", flush=True)
        assert its.tolist() == [3, 3] and conv.tolist() == [0, 0] and batch.last_message != ""
        assert np.all(rel > 1e-10) and np.all(np.isfinite(rel))
        ref = [torch.empty_like(r) for r in rhs]
        its1, rel1, conv1 = batch.solve(rhs, ref)
        assert conv1.tolist() == [1, 1]
        ys = [torch.empty_like(r) for r in rhs]
        batch.begin(rhs)
        replays = 0
        while True:
            batch.iterate(8)
            torch.cuda.synchronize()
            replays += 1
            if batch.stopped():
                break
        its2, rel2, conv2 = batch.end(ys)
        print(f"   steps: {replays} calls of 8 iterations, iterations {its2.tolist()}", flush=True)
        assert its2.tolist() == its1.tolist() and rel2.tolist() == rel1.tolist() and conv2.tolist() == [1, 1]
        assert all(torch.equal(a, b) for a, b in zip(ref, ys))
    finally:
        if batch is not None:
            batch.close()
        for sv in solvers:
            sv.close()


def test_krylov_settings_are_checked():
    print("[gpu] linear solve settings", flush=True)
    M, fixed, ctx, K, rhs, dK, drhs = system("tiny_p2")
    solver = GpuLinearSolver(ctx, fixed)
    try:
        with pytest.raises(uvh.UvhLibraryError):
            solver.set_krylov("gmres", 1)
        with pytest.raises(uvh.UvhLibraryError):
            solver.set_krylov("gmres", 65)
        with pytest.raises(KeyError):
            solver.set_krylov("cg", 30)
    finally:
        solver.close()


def test_solve_is_reproducible_and_zero_rhs_gives_zero():
    import torch
    print("[gpu] linear solve reproducibility", flush=True)
    M, fixed, ctx, K, rhs, dK, drhs = system("medium_p2")
    solver = GpuLinearSolver(ctx, fixed)
    try:
        solver.factor(dK)
        a, b = torch.empty_like(drhs), torch.empty_like(drhs)
        ia, _ = solver.solve(drhs, a)
        ib, _ = solver.solve(drhs, b)
        assert ia == ib and torch.equal(a, b)
        z = torch.full_like(drhs, 7.0)
        its, relres = solver.solve(torch.zeros_like(drhs), z)
        assert its == 0 and torch.count_nonzero(z).item() == 0
    finally:
        solver.close()


def test_iteration_limit_and_call_order_are_enforced():
    import torch
    print("[gpu] linear solve limits", flush=True)
    M, fixed, ctx, K, rhs, dK, drhs = system("medium_p2")
    solver = GpuLinearSolver(ctx, fixed, raise_on_noconv=True)
    try:
        x = torch.empty_like(drhs)
        with pytest.raises(uvh.UvhLibraryError):   # solve before factor
            solver.solve(drhs, x)
        solver.factor(dK)
        with pytest.raises(uvh.UvhLibraryError):   # rtol must be positive
            solver.solve(drhs, x, rtol=0.0)
        with pytest.raises(LinearSolveError):      # two iterations cannot reach 1e-12
            solver.solve(drhs, x, rtol=1e-12, maxit=2)
        assert solver.last_iterations == 2 and solver.last_relres > 1e-12
    finally:
        solver.close()


@pytest.mark.parametrize("name", ["cpu_reference", "medium_p2"])
def test_newton_with_device_solver_matches_oracle_newton(o2lib, name):
    from test_newton import oracle_solution, rel
    M, F0, F1, fixed, ref = oracle_solution(o2lib, name)
    cv = default_ctrlvar()
    print(f"[gpu] newton {name} with the device linear solver (oracle: {ref.iterations} iterations)", flush=True)
    asm = newton.GpuAssembler(cv, M, F0, F1)
    solver = GpuLinearSolver(asm.ctx, fixed, rtol=cv.LinSolveTol)
    try:
        res = newton.solve(asm, cv, linear_solver=solver, fixed=fixed)
    finally:
        solver.close()
    print(f"   {res.iterations} iterations, {res.reason}, rForce {res.rForce:.2e}, "
          f"Krylov iterations {res.linear_iterations}", flush=True)
    assert ref.converged and res.converged, (ref.reason, res.reason)
    assert abs(res.iterations - ref.iterations) <= 1
    assert res.rForce < 1e-15
    for a, b in ((res.ub, ref.ub), (res.vb, ref.vb), (res.h, ref.h)):
        assert rel(a, b) < 1e-7
