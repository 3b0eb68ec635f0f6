"""Batched Newton solves (uasource_amd.batch_newton): five independent uvh problems of 3-, 6-
and 10-node elements (the Newton workloads of ``fields.newton_problem``, with two reseeded
copies of the 3-node one) stepped together, the assembly and residual-cost stages replayed
from HIP graphs.

* Each member converges as its single-mesh solve does (``newton.solve`` with a
  ``GpuAssembler`` and the device linear solver): iteration counts within one, rForce below
  1e-15, u, v and h within 1e-7 (relative) of the single-mesh solution.
* The graph replays give the eager run's results bit for bit: the same history (gamma,
  rForce, rWork per iteration) and the same states, after tens of replays of each graph with
  the eager solves and host reads between them.
"""
import dataclasses

import numpy as np
import pytest

from uasource_amd import newton
from uasource_amd.batch_newton import BatchNewton
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, newton_problem
from uasource_amd.linsolve import GpuLinearSolver

pytestmark = pytest.mark.gpu

_members = None


def members():
    global _members
    if _members is None:
        ws = [WORKLOADS["cpu_reference"], WORKLOADS["medium_p2"], WORKLOADS["medium_p3"],
              dataclasses.replace(WORKLOADS["cpu_reference"], name="cpu_reference_s5", seed=5),
              dataclasses.replace(WORKLOADS["cpu_reference"], name="cpu_reference_s6", seed=6, C_sigma=0.3)]
        _members = []
        for w in ws:
            M = build_mesh(w)
            F0, F1, _, fixed = newton_problem(M, w)
            _members.append((w.name, M, F0, F1, fixed))
    return _members


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run_batch(graphs):
    cv = default_ctrlvar()
    bn = BatchNewton(cv, [(M, F0, F1, fixed) for _, M, F0, F1, fixed in members()], graphs=graphs)
    try:
        res = bn.solve()
        return res, bn.n_graph_replays
    finally:
        bn.close()


def test_batch_members_converge_like_single_mesh_solves():
    print("[gpu] batched Newton, 5 members, graphs", flush=True)
    res, replays = run_batch(True)
    assert replays > 0
    cv = default_ctrlvar()
    for (name, M, F0, F1, fixed), r in zip(members(), res):
        asm = newton.GpuAssembler(cv, M, F0, F1)
        solver = GpuLinearSolver(asm.ctx, fixed, rtol=cv.LinSolveTol)
        try:
            single = newton.solve(asm, cv, linear_solver=solver, fixed=fixed)
        finally:
            solver.close()
            asm.ctx.close()
        print(f"   {name} (Nele {M.Nele}, nod {M.nod}): batched {r.iterations} iterations ({r.reason}, "
              f"rForce {r.rForce:.2e}), single {single.iterations} ({single.reason})", flush=True)
        assert r.converged and single.converged, (r.reason, single.reason)
        assert abs(r.iterations - single.iterations) <= 1
        assert r.rForce < 1e-15
        for a, b in ((r.ub, single.ub), (r.vb, single.vb), (r.h, single.h)):
            assert rel(a, b) < 1e-7
        x0 = np.concatenate([F1.ub, F1.vb, F1.h])
        assert np.array_equal(np.concatenate([r.ub, r.vb, r.h])[fixed], x0[fixed])
    print(f"   {replays} graph replays", flush=True)


def test_batch_graph_replays_equal_eager_run():
    print("[gpu] batched Newton, graphs against eager", flush=True)
    eager, n0 = run_batch(False)
    graph, n1 = run_batch(True)
    assert n0 == 0 and n1 >= 20
    for (name, *_), a, b in zip(members(), eager, graph):
        assert a.history == b.history, name
        assert a.linear_iterations == b.linear_iterations, name
        for u, v in ((a.ub, b.ub), (a.vb, b.vb), (a.h, b.h)):
            assert np.array_equal(u, v), name
    print(f"   {n1} replays, every member's history and state equal", flush=True)
