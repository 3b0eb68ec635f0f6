"""Newton-Raphson uvh solve (uasource_amd.newton): the driver restates
SSTREAM_TransientImplicit.m:224-418. CPU tests run the driver on oracle O2; GPU tests run it on
the HIP assembly and require the same behaviour: both converge to Úa's default tolerance
(rForce < 1e-15), the iteration counts agree within one, and the converged u, v, h agree.
The precision form (FP32 Jacobian, FP32 residual until rForce < 1e-8 or its line search stalls,
then FP64) must reach the FP64 solution to 1e-6 (velocities relative to the speed norm)."""
import numpy as np
import pytest

from newton_oracle import OracleAssembler
from uasource_amd import newton
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, newton_problem

_ref = {}


def oracle_solution(o2lib, name):
    if name not in _ref:
        M = build_mesh(WORKLOADS[name])
        F0, F1, BCs, fixed = newton_problem(M, WORKLOADS[name])
        cv = default_ctrlvar()
        res = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, fixed=fixed)
        _ref[name] = (M, F0, F1, fixed, res)
    return _ref[name]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_oracle_newton_converges_quadratically(o2lib):
    M, F0, F1, fixed, res = oracle_solution(o2lib, "cpu_reference")
    assert res.converged and res.rForce < 1e-15, (res.reason, res.rForce)
    # quadratic convergence at the end: the last full steps square the residual
    last = [h[2] for h in res.history if h[1] == 1.0][-3:]
    assert len(last) == 3 and last[2] < 1e-3 * last[1] and last[1] < 1e-2 * last[0]
    # the fixed degrees of freedom did not move
    x0 = np.concatenate([F1.ub, F1.vb, F1.h])
    x1 = np.concatenate([res.ub, res.vb, res.h])
    assert np.array_equal(x0[fixed], x1[fixed])


def test_host_solvers_agree(o2lib):
    """The host solves give the same Newton iteration: minimum-degree direct (the default),
    COLAMD with partial pivoting (its fall-back), and the LU kept across steps as the GMRES
    preconditioner; the same iteration counts (within one), rForce < 1e-15 and the same state."""
    import scipy.sparse.linalg as spla
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    colamd = lambda K, rhs: spla.splu(K.tocsc(), permc_spec="COLAMD").solve(rhs)  # noqa: E731
    lagged = newton.LaggedLU()
    for solver in (colamd, lagged):
        res = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, linear_solver=solver, fixed=fixed)
        assert res.converged and res.rForce < 1e-15, res.reason
        assert abs(res.iterations - ref.iterations) <= 1
        for x, y in ((res.ub, ref.ub), (res.vb, ref.vb), (res.h, ref.h)):
            assert rel(x, y) < 1e-9
    # the LU was reused for some steps (GMRES on an older factorization)
    assert lagged.factorizations < ref.iterations and max(lagged.krylov) > 0


def test_checkpoint_and_resume_give_the_uninterrupted_run(o2lib):
    """A run stopped at its time limit and resumed from the checkpoint, one iteration per call,
    ends where the uninterrupted run ends, bit for bit."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    ck, calls = None, 0
    while True:
        res = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, fixed=fixed, time_limit=0.0, resume=ck)
        calls += 1
        if res.checkpoint is None:
            break
        ck = res.checkpoint
        assert res.reason == "time limit" and calls < 100
    assert calls == ref.iterations + 1 or calls == ref.iterations
    assert res.converged and res.iterations == ref.iterations and res.rForce == ref.rForce
    for x, y in ((res.ub, ref.ub), (res.vb, ref.vb), (res.h, ref.h)):
        assert np.array_equal(x, y)
    assert [h[:4] for h in res.history] == [h[:4] for h in ref.history]


def test_auto_line_search_takes_the_cauchy_step_where_newton_fails(o2lib):
    """rLineMinUa "-Auto-": with a linear solve that returns the reversed Newton step, the
    Newton direction gives no reduction; the Newton-only line search stops at once, while
    the Cauchy step with the mass matrix (rLineminUa.m:137-259) still reduces rForce at every
    iteration. Where the Newton step does well, both give the same iteration."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    cv.NRitmax = 3
    reverse = lambda K, rhs: -newton.superlu_solve(K, rhs)  # noqa: E731
    plain = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, linear_solver=reverse, fixed=fixed)
    assert plain.reason == "backtracking stagnated" and plain.iterations == 1
    cv.rLineMinUa = "-Auto-"
    auto = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, linear_solver=reverse, fixed=fixed)
    assert auto.cauchy_steps == auto.iterations >= 3
    r = [h[2] for h in auto.history]
    assert all(b < a for a, b in zip(r, r[1:]))
    c = default_ctrlvar()
    c.rLineMinUa = "-Auto-"
    same = newton.solve(OracleAssembler(o2lib, c, M, F0, F1), c, fixed=fixed)
    assert same.iterations == ref.iterations and same.rForce == ref.rForce


def test_steepest_descent_stage(o2lib):
    """The steepest-descent stage of the Auto line search on its own: its direction is frhs with
    the fixed degrees of freedom held, and what it returns is never worse than the start."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    asm = OracleAssembler(o2lib, cv, M, F0, F1)
    asm.set_fixed(fixed)
    N = asm.normalisation()
    K, rhs = asm.assemble_rk(False, False)
    r0 = asm.cost(0.0, False, N)[0]

    def cost(g, s):
        asm.set_step(s)
        return asm.cost(g, False, N)
    out = newton._CauchyM(M, fixed).steepest(cv, K, rhs, cost, r0, N)
    assert out is not None
    g, best, s, gD = out
    free = np.ones(3 * M.Nnodes, dtype=bool)
    free[fixed] = False
    assert np.all(s[~free] == 0) and np.array_equal(s[free], rhs[free])
    assert gD > 0 and best[0] <= r0 and (g == 0.0 or best[0] < r0)


def _rows(idx, n3):
    """Rows of the identity at ``idx`` (m x n3): L x = x[idx]."""
    import scipy.sparse as sp
    idx = np.asarray(idx)
    return sp.csr_matrix((np.ones(idx.size), (np.arange(idx.size), idx)), shape=(idx.size, n3))


def _infeasible_start(F1, fixed, seed=3):
    """F1 with the constrained values moved off their prescribed ones."""
    x = np.concatenate([F1.ub, F1.vb, F1.h])
    rng = np.random.default_rng(seed)
    x[fixed] += 1e-3 * (np.abs(x[fixed]) + 1.0) * rng.standard_normal(len(fixed))
    G = F1.copy()
    n = F1.h.size
    G.ub, G.vb, G.h = x[:n].copy(), x[n:2 * n].copy(), x[2 * n:].copy()
    return G


def test_constraints_as_L_give_the_eliminated_solution(o2lib):
    """Dirichlet conditions given as L x = cuvh (Lagrange multipliers and the saddle-point
    system of solveKApe.m), started off the constraints, reach the solution of the eliminated
    form; at convergence frhs = -R - L'luvh = 0, so the multipliers are the reactions,
    luvh = -R on the constrained rows."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    x0 = np.concatenate([F1.ub, F1.vb, F1.h])
    asm = OracleAssembler(o2lib, cv, M, F0, _infeasible_start(F1, fixed))
    res = newton.solve(asm, cv, L=_rows(fixed, x0.size), cuvh=x0[fixed])
    assert res.converged and res.rForce < 1e-15, (res.reason, res.rForce)
    # both runs stop at rForce < 1e-15 (|R| / |Fext0| < 3e-8), which fixes the state [u; v; h]
    # to ~1e-9 of its norm (vb, a hundredth of ub here, is compared at that scale too)
    x1 = np.concatenate([res.ub, res.vb, res.h])
    assert rel(x1, np.concatenate([ref.ub, ref.vb, ref.h])) < 1e-9
    # feasibility as SSTREAM_TransientImplicit.m:210 measures it: |L x - cuvh| / |cuvh|
    assert np.linalg.norm(x1[fixed] - x0[fixed]) < 1e-10 * np.linalg.norm(x0[fixed])
    R = asm._assemble(x1, True)["R"]
    assert rel(res.luvh, -R[fixed]) < 1e-8
    # the free rows are in equilibrium
    free = np.ones(x0.size, dtype=bool)
    free[fixed] = False
    assert np.linalg.norm(R[free]) < 1e-7 * np.linalg.norm(R[fixed])
    with pytest.raises(ValueError):
        newton.solve(asm, cv, L=_rows(fixed, x0.size), cuvh=x0[fixed], fixed=fixed)
    with pytest.raises(ValueError):
        newton.solve(asm, cv, L=_rows(fixed, x0.size)[:, :-1], cuvh=x0[fixed])


def _tied(M, F1, fixed):
    """The Dirichlet rows plus one tie, h at two neighbouring free nodes equal: (L, cuvh, a, b)."""
    import scipy.sparse as sp
    n = M.Nnodes
    x0 = np.concatenate([F1.ub, F1.vb, F1.h])
    isfixed = np.zeros(3 * n, dtype=bool)
    isfixed[fixed] = True
    a, b = next((int(e[0]), int(e[1])) for e in M.connectivity
                if not isfixed[2 * n + e[0]] and not isfixed[2 * n + e[1]] and abs(F1.h[e[0]] - F1.h[e[1]]) > 1e-3)
    tie = sp.csr_matrix(([1.0, -1.0], ([0, 0], [2 * n + a, 2 * n + b])), shape=(1, 3 * n))
    return sp.vstack([_rows(fixed, 3 * n), tie]).tocsr(), np.concatenate([x0[fixed], [0.0]]), a, b


def test_tie_constraint_and_resume_with_multipliers(o2lib):
    """A constraint that is not a boundary condition, h at two neighbouring nodes tied equal
    (a row with +1 and -1), on top of the Dirichlet rows: the solution meets every constraint to
    rounding, differs from the untied one, and a run stopped at its time limit and resumed from
    the checkpoint (state and multipliers) ends where the uninterrupted run ends, bit for bit."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    L, cuvh, a, b = _tied(M, F1, fixed)
    res = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, L=L, cuvh=cuvh)
    assert res.converged and res.rForce < 1e-15, (res.reason, res.rForce)
    x1 = np.concatenate([res.ub, res.vb, res.h])
    assert np.linalg.norm(L @ x1 - cuvh) < 1e-10 * np.linalg.norm(cuvh)
    assert abs(ref.h[a] - ref.h[b]) > 1e-6 and rel(res.h, ref.h) > 1e-12
    ck, calls = None, 0
    while True:
        r = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, L=L, cuvh=cuvh, time_limit=0.0, resume=ck)
        calls += 1
        if r.checkpoint is None:
            break
        ck = r.checkpoint
        assert "luvh" in ck and calls < 100
    assert r.iterations == res.iterations and r.rForce == res.rForce
    assert np.array_equal(r.h, res.h) and np.array_equal(r.luvh, res.luvh)


def test_constraints_with_the_lagged_lu(o2lib):
    """The lagged LU (GMRES preconditioned with an older factorization) on the saddle-point
    system of the constrained step reaches the direct solve's solution in as many Newton steps,
    with fewer factorizations than steps."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    L, cuvh, a, b = _tied(M, F1, fixed)
    direct = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, L=L, cuvh=cuvh)
    s = newton.LaggedLU()
    lagged = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, L=L, cuvh=cuvh, linear_solver=s)
    assert lagged.converged and lagged.iterations == direct.iterations
    assert s.factorizations < lagged.iterations
    assert rel(np.concatenate([lagged.ub, lagged.vb, lagged.h]), np.concatenate([direct.ub, direct.vb, direct.h])) < 1e-9
    assert rel(lagged.luvh, direct.luvh) < 1e-7


def test_line_search_accepts_full_step_and_backtracks():
    cv = default_ctrlvar()
    # full step below NewtonAcceptRatio * r0
    g, best, n = newton._line_search(cv, lambda g: (1.0, 0, 0), 1.0, (0.1, 0.0, 0.0))
    assert g == 1.0 and n == 0
    # f(g) = (1 - 2g)^2 + 0.2 g^2... a parabola with its minimum at g ~ 0.45: backtracking finds a reduction
    f = lambda g: ((1 - 2 * g) ** 2 + 0.2 * g * g, 0.0, 0.0)
    g, best, n = newton._line_search(cv, f, 1.0, f(1.0))
    assert 0.25 <= g < 1.0 and best[0] < 0.5 and n >= 1
    # no reduction anywhere: gamma = 0
    g, best, n = newton._line_search(cv, lambda g: (2.0, 0, 0), 1.0, (2.0, 0.0, 0.0))
    assert g == 0.0


def test_nan_work_is_judged_on_rforce_alone(o2lib, monkeypatch):
    """A line-search point without a work measure (rWork NaN, as after a Cauchy or
    steepest-descent point) counts as rWork = 0 in the exit tests
    (SSTREAM_TransientImplicit.m:275-281): with every accepted point reported so, the driver
    still stops with "converged" at the iteration where rForce meets the tolerance."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    real = newton._line_search

    def no_work(*a, **k):
        g, best, n = real(*a, **k)
        return g, (best[0], float("nan"), float("nan")), n

    monkeypatch.setattr(newton, "_line_search", no_work)
    cv = default_ctrlvar()
    res = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, fixed=fixed)
    assert res.converged and res.reason == "converged"
    assert abs(res.iterations - ref.iterations) <= 1 and res.rForce < cv.uvhDesiredWorkAndForceTolerances[1]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cpu_reference", "medium_p2", "medium_p3"])
def test_gpu_newton_matches_oracle_newton(o2lib, name):
    import sys
    M, F0, F1, fixed, ref = oracle_solution(o2lib, name)
    cv = default_ctrlvar()
    print(f"[gpu] newton {name} (oracle: {ref.iterations} iterations, {ref.reason})", flush=True)
    sys.stdout.flush()
    gpu = newton.solve(newton.GpuAssembler(cv, M, F0, F1), cv, fixed=fixed)
    print(f"   gpu: {gpu.iterations} iterations, {gpu.reason}, rForce {gpu.rForce:.2e}", flush=True)
    assert ref.converged and gpu.converged, (ref.reason, gpu.reason)
    assert abs(gpu.iterations - ref.iterations) <= 1
    assert gpu.rForce < 1e-15
    for a, b in ((gpu.ub, ref.ub), (gpu.vb, ref.vb), (gpu.h, ref.h)):
        assert rel(a, b) < 1e-9


@pytest.mark.gpu
def test_gpu_newton_auto_line_search_matches_oracle(o2lib):
    """rLineMinUa "-Auto-" on the GPU assembly: with the reversed Newton step only the Cauchy
    points reduce rForce; the GPU-driven run takes them like the oracle-driven one, and with
    the true step both converge the same way."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    print("[gpu] newton Auto line search cpu_reference", flush=True)
    cv = default_ctrlvar()
    cv.rLineMinUa = "-Auto-"
    cv.NRitmax = 3
    reverse = lambda K, rhs: -newton.superlu_solve(K, rhs)  # noqa: E731
    a = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, linear_solver=reverse, fixed=fixed)
    g = newton.solve(newton.GpuAssembler(cv, M, F0, F1), cv, linear_solver=reverse, fixed=fixed)
    print(f"   oracle {[h[2] for h in a.history]}, gpu {[h[2] for h in g.history]}", flush=True)
    assert g.cauchy_steps == a.cauchy_steps == g.iterations == a.iterations
    for x, y in zip(g.history, a.history):
        assert abs(x[2] - y[2]) <= 1e-9 * y[2]
    c = default_ctrlvar()
    c.rLineMinUa = "-Auto-"
    res = newton.solve(newton.GpuAssembler(c, M, F0, F1), c, fixed=fixed)
    assert res.converged and abs(res.iterations - ref.iterations) <= 1 and res.rForce < 1e-15


@pytest.mark.gpu
def test_gpu_newton_with_other_sliding_laws(o2lib):
    """Newton under Budd converges on the GPU assembly as on the oracle's (same iterations within
    one, rForce < 1e-15, same u, v, h); under Cornford (with the Auto line search; the
    Coulomb-type laws do not converge from this start, on either assembly) the first five
    iterations follow the oracle-driven ones."""
    M = build_mesh(WORKLOADS["cpu_reference"])
    F0, F1, _, fixed = newton_problem(M, WORKLOADS["cpu_reference"])
    rng = np.random.default_rng(3)
    F1.q = rng.uniform(0.5, 1.5, M.Nnodes)
    F1.muk = 0.02 * rng.uniform(0.3, 0.7, M.Nnodes)
    print("[gpu] newton with the Budd and Cornford sliding laws", flush=True)
    cv = default_ctrlvar()
    cv.SlidingLaw = "Budd"
    a = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, fixed=fixed)
    g = newton.solve(newton.GpuAssembler(cv, M, F0, F1), cv, fixed=fixed)
    print(f"   Budd: oracle {a.iterations}, gpu {g.iterations} iterations, rForce {g.rForce:.1e}", flush=True)
    assert a.converged and g.converged and abs(a.iterations - g.iterations) <= 1 and g.rForce < 1e-15
    for x, y in ((g.ub, a.ub), (g.vb, a.vb), (g.h, a.h)):
        assert rel(x, y) < 1e-7
    cv = default_ctrlvar()
    cv.SlidingLaw, cv.rLineMinUa, cv.NRitmax = "Cornford", "-Auto-", 5
    a = newton.solve(OracleAssembler(o2lib, cv, M, F0, F1), cv, fixed=fixed)
    g = newton.solve(newton.GpuAssembler(cv, M, F0, F1), cv, fixed=fixed)
    ra, rg = [h[2] for h in a.history], [h[2] for h in g.history]
    print(f"   Cornford: oracle {ra}, gpu {rg}", flush=True)
    assert len(ra) == len(rg) >= 5
    for x, y in zip(rg, ra):
        assert abs(x - y) <= 1e-6 * y


@pytest.mark.gpu
def test_gpu_newton_with_linear_constraints_matches_oracle(o2lib):
    """Linear constraints (Dirichlet rows and a tie) with Lagrange multipliers, from an
    infeasible start: the device-resident cost (R + L'luvh + gamma L'dl on the GPU) against
    the oracle-driven run."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    L, cuvh, a, b = _tied(M, F1, fixed)
    G1 = _infeasible_start(F1, fixed)
    print("[gpu] newton with linear constraints L (cpu_reference)", flush=True)
    o = newton.solve(OracleAssembler(o2lib, cv, M, F0, G1), cv, L=L, cuvh=cuvh)
    asm = newton.GpuAssembler(cv, M, F0, G1)
    g = newton.solve(asm, cv, L=L, cuvh=cuvh)
    print(f"   gpu: {g.iterations} iterations ({o.iterations} oracle), {g.reason}, rForce {g.rForce:.2e}", flush=True)
    assert o.converged and g.converged, (o.reason, g.reason)
    assert abs(g.iterations - o.iterations) <= 1 and g.rForce < 1e-15
    xo, xg = np.concatenate([o.ub, o.vb, o.h]), np.concatenate([g.ub, g.vb, g.h])
    assert rel(xg, xo) < 1e-9 and rel(g.luvh, o.luvh) < 1e-7
    assert np.linalg.norm(L @ xg - cuvh) < 1e-10 * np.linalg.norm(cuvh)
    assert asm.shift is None   # the assembler's cost is the unconstrained one again


@pytest.mark.gpu
def test_gpu_time_steps_with_resident_fields(o2lib):
    """Implicit time steps on one GpuAssembler (next_time_step: F0 takes the converged state in
    place, nothing is uploaded again; then new forcing, a new dt and a given first guess) equal
    each step solved on a fresh assembler built from the previous step's result, bit for bit."""
    M, F0, F1, fixed, ref = oracle_solution(o2lib, "cpu_reference")
    cv = default_ctrlvar()
    print("[gpu] newton three time steps, resident fields", flush=True)
    asm = newton.GpuAssembler(cv, M, F0, F1)
    r1 = newton.solve(asm, cv, fixed=fixed)
    asm.next_time_step()
    r2 = newton.solve(asm, cv, fixed=fixed)
    G0 = F0.copy()
    G0.ub, G0.vb, G0.h = r1.ub, r1.vb, r1.h
    G0.as_, G0.ab = F1.as_, F1.ab
    G1 = F1.copy()
    G1.ub, G1.vb, G1.h = r1.ub, r1.vb, r1.h
    fresh = newton.solve(newton.GpuAssembler(cv, M, G0, G1), cv, fixed=fixed)
    print(f"   step 2: {r2.iterations} iterations (fresh {fresh.iterations}), rForce {r2.rForce:.1e}", flush=True)
    assert r2.converged and r2.iterations == fresh.iterations
    for x, y in ((r2.ub, fresh.ub), (r2.vb, fresh.vb), (r2.h, fresh.h)):
        assert np.array_equal(x, y)
    # the second step moved the state on from the first (h evolves with dt)
    assert rel(r2.h, ref.h) > 1e-12
    # a third step with new forcing (a_s up by half), a shorter dt and a given first guess
    H1 = F1.copy()
    H1.as_ = 1.5 * F1.as_
    H1.ub, H1.vb, H1.h = r2.ub, r2.vb, r2.h
    asm.next_time_step(F1=H1, dt=0.5 * cv.dt)
    r3 = newton.solve(asm, cv, fixed=fixed)
    cv3 = cv.copy()
    cv3.dt = 0.5 * cv.dt
    H0 = F0.copy()
    H0.ub, H0.vb, H0.h, H0.as_, H0.ab = r2.ub, r2.vb, r2.h, F1.as_, F1.ab
    fresh3 = newton.solve(newton.GpuAssembler(cv3, M, H0, H1), cv3, fixed=fixed)
    print(f"   step 3: {r3.iterations} iterations (fresh {fresh3.iterations})", flush=True)
    assert r3.converged and r3.iterations == fresh3.iterations
    for x, y in ((r3.ub, fresh3.ub), (r3.vb, fresh3.vb), (r3.h, fresh3.h)):
        assert np.array_equal(x, y)
    assert rel(r3.h, r2.h) > 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["medium_p2", "medium_p3"])
def test_gpu_newton_precision_form_reaches_fp64_solution(o2lib, name):
    M, F0, F1, fixed, ref = oracle_solution(o2lib, name)
    cv = default_ctrlvar()
    print(f"[gpu] newton precision form {name}", flush=True)
    res = newton.solve(newton.GpuAssembler(cv, M, F0, F1), cv, fixed=fixed, fp32=True)
    # velocities against the speed scale ||(u, v)|| (v is near zero here, so ||dv||/||v|| says
    # little), thickness against ||h||
    speed = np.linalg.norm(np.concatenate([ref.ub, ref.vb]))
    errs = [np.linalg.norm(res.ub - ref.ub) / speed, np.linalg.norm(res.vb - ref.vb) / speed, rel(res.h, ref.h)]
    print(f"   fp32 form: {res.iterations} iterations ({ref.iterations} FP64), {res.reason}, "
          f"u/v/h rel err {errs[0]:.1e}/{errs[1]:.1e}/{errs[2]:.1e}", flush=True)
    assert res.converged and res.rForce < 1e-15
    assert res.history[-1][4], "the last step used an FP64 residual"
    assert max(errs) < 1e-6


def test_uv_system_in_the_uvh_layout(o2lib):
    """The index map that lets the uvh device solver take a uv step: the uv K (oracle O2's,
    4 blocks) placed into the uvh layout has the uv K as its upper-left 2 x 2 blocks and zeros in
    the thickness blocks, on the uvh pattern."""
    import scipy.sparse as sp
    from uasource_amd.fields import surface_from_floatation
    M = build_mesh(WORKLOADS["tiny_p2"])
    F0, F1, _, _ = newton_problem(M, WORKLOADS["tiny_p2"])
    cv = default_ctrlvar()
    G = F1.copy()
    G.s = surface_from_floatation(F1, cv)
    uv = o2lib.assemble_uv(cv, M, G)
    full = o2lib.assemble(cv, M, F0, F1)
    rp = full["mesh"].rowptr_n
    idx = newton.uv_to_uvh_index(rp)
    vals = np.where(idx >= 0, uv["Kval"][np.maximum(idx, 0)], 0.0)
    K = sp.csr_matrix((vals, full["K"].indices, full["K"].indptr), shape=full["K"].shape)
    n = M.Nnodes
    assert abs(K[:2 * n, :2 * n] - uv["K"]).max() == 0
    assert K[2 * n:, :].nnz == 0 or abs(K[2 * n:, :]).max() == 0
    assert abs(K[:, 2 * n:]).max() == 0
