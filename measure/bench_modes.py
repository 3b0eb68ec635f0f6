"""Measurements of the batched, serving and Newton forms (``bench.py --batch / --serve /
--newton``). Each returns the dict that bench.py prints as one JSON line."""
from __future__ import annotations

import time

import numpy as np

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, batch_workloads, build_mesh, make_fields, newton_problem, serve_workloads


def _device_fields(torch, dev, F0, F1):
    keep = []

    class _F:
        pass
    f1, f0 = _F(), _F()
    for k in uvh.F1_KEYS:
        t = torch.from_numpy(np.ascontiguousarray(getattr(F1, k))).to(dev)
        keep.append(t)
        setattr(f1, k, t)
    for k in uvh.F0_KEYS:
        t = torch.from_numpy(np.ascontiguousarray(getattr(F0, k))).to(dev)
        keep.append(t)
        setattr(f0, k, t)
    return uvh.make_fields(f0, f1), keep


def _time(torch, fn, steps, stream):
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_batch(steps: int, warmup: int, n_members: int = 64, log=print) -> dict:
    """64 independent meshes (log-uniform 5k-80k elements, P1/P2 mixed, perturbed C and AGlen):
    batched assembly (one launch per stage and element type), FP32 element arithmetic with FP64
    accumulation and FP64, autotuned tiles, eager and HIP-graph replay; against the same work
    as 64 single-mesh assemblies."""
    import torch
    from uasource_amd.batch import UvhBatch, autotune
    dev = torch.device("cuda:0")
    cv = default_ctrlvar()
    ws = batch_workloads(n=n_members)
    t0 = time.perf_counter()
    members = []
    for w in ws:
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        members.append((M, F0, F1))
    b = UvhBatch(0)
    for M, _, _ in members:
        b.add_mesh(M)
    log(f"batch: {len(members)} meshes, {sum(m[0].Nele for m in members)} elements, set-up {time.perf_counter() - t0:.1f} s")
    fields, keep, R, K = [], [], [], []
    for (M, F0, F1), mv in zip(members, b.members):
        f, k = _device_fields(torch, dev, F0, F1)
        fields.append(f)
        keep += k
        R.append(torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev))
        K.append(torch.zeros(mv.nnz, dtype=torch.float64, device=dev))
    ctrl = uvh.make_ctrl(cv, members[0][2])
    s = torch.cuda.Stream(dev)
    sp = s.cuda_stream
    nele = sum(m[0].Nele for m in members)
    out = {}
    with torch.cuda.stream(s):
        for fp32 in (True, False):
            flags = uvh.UVH_FP32 if fp32 else 0
            tiles = autotune(b, ctrl, fields, R, K, flags, sp, torch, reps=3)
            for _ in range(warmup):
                b.assemble_device(ctrl, fields, flags, R, K, stream=sp)
            b.sync_status()
            ms = _time(torch, lambda: b.assemble_device(ctrl, fields, flags, R, K, stream=sp), steps, s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                b.assemble_device(ctrl, fields, flags, R, K, stream=sp)
            ms_graph = _time(torch, g.replay, steps, s)
            key = "fp32" if fp32 else "fp64"
            out[key] = {"ms": ms, "ms_graph": ms_graph, "elements_per_s": nele / (min(ms, ms_graph) * 1e-3),
                        "tiles": {f"nod{k[0]}": v[0] for k, v in tiles.items()}}
            out[key + "_R"] = [r.clone() for r in R]
            out[key + "_K"] = [k.clone() for k in K]
            log(f"batch {key}: {ms:.3f} ms eager, {ms_graph:.3f} ms graph, tiles {out[key]['tiles']}")
        # the same work as separate single-mesh assemblies (one context per member)
        ctxs = [mv for mv in b.members]

        def singles():
            for (M, F0, F1), mv, f, r, k in zip(members, ctxs, fields, R, K):
                mv.assemble_device(ctrl, f, 0, r, k, None, None, stream=sp)
        for _ in range(warmup):
            singles()
        ms_single = _time(torch, singles, steps, s)
    err = {"R": 0.0, "K": 0.0}
    for i, (M, F0, F1) in enumerate(members):
        r32, r64 = out["fp32_R"][i], out["fp64_R"][i]
        k32, k64 = out["fp32_K"][i], out["fp64_K"][i]
        err["R"] = max(err["R"], (torch.linalg.norm(r32 - r64) / torch.linalg.norm(r64)).item())
        err["K"] = max(err["K"], (torch.linalg.norm(k32 - k64) / torch.linalg.norm(k64)).item())
    b.close()
    single_rate = nele / (ms_single * 1e-3)
    return {
        "metric": "uvh elements assembled/s (R+K), batched ensemble",
        "value": out["fp32"]["elements_per_s"], "unit": "elements/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": min(out["fp32"]["ms"], out["fp32"]["ms_graph"]), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 element arithmetic, f64 accumulation",
        "data": "synthetic (64 seeded jittered-Delaunay meshes, perturbed C and AGlen)",
        "config": {"workload": "batch64", "members": len(members), "Nele_total": nele,
                   "types": sorted({f"P{1 if m[0].nod == 3 else 2 if m[0].nod == 6 else 3}" for m in members}),
                   "parallelism": "1 GPU"},
        "batched": {k: {kk: vv for kk, vv in v.items()} for k, v in out.items() if k in ("fp32", "fp64")},
        "single_mesh_fp64_elements_per_s": single_rate,
        "speedup_fp32_batched_vs_single_fp64": out["fp32"]["elements_per_s"] / single_rate,
        "fp32_max_rel_err_vs_fp64": err,
    }


def bench_batch_newton(n_members: int = 16, log=print) -> dict:
    """Batched Newton solves (``uasource_amd.batch_newton``): ``n_members`` independent uvh
    problems (the Newton workload recipe on 100 km x 50 km P1/P2 meshes of 2k-7k elements,
    reseeded per member), run to rForce < 1e-15 with the assembly and cost stages replayed
    from HIP graphs and the members' Newton steps in one batched Krylov solve; for comparison
    eagerly (the same iterates bit for bit, ``tests/test_gpu_batch_newton.py``) and with the
    members' steps solved one after another."""
    import dataclasses
    from uasource_amd.batch_newton import BatchNewton
    base = [WORKLOADS["cpu_reference"], WORKLOADS["medium_p2"]]
    mem = []
    for k in range(n_members):
        w = dataclasses.replace(base[k % 2], name=f"member{k}", seed=100 + k)
        M = build_mesh(w)
        F0, F1, _, fixed = newton_problem(M, w)
        mem.append((M, F0, F1, fixed))
    cv = default_ctrlvar()
    out = {}
    for graphs, batched, key in ((True, True, "graphs"), (False, True, "eager"), (True, True, "graphs"),
                                 (True, False, "graphs_member_solves")):
        bn = BatchNewton(cv, mem, graphs=graphs, batched_solve=batched)
        t = time.perf_counter()
        res = bn.solve()
        dt = time.perf_counter() - t
        replays = bn.n_graph_replays
        bn.close()
        out[key] = {"seconds": dt, "graph_replays": replays, "iterations": [r.iterations for r in res],
                    "converged": all(r.converged for r in res), "max_rForce": max(r.rForce for r in res),
                    "krylov_iterations": int(sum(sum(r.linear_iterations) for r in res))}
        log(f"batch newton ({key}): {dt:.2f} s, iterations {out[key]['iterations']}, replays {replays}")
    nele = sum(m[0].Nele for m in mem)
    return {
        "metric": "batched Newton uvh solves to rForce < 1e-15 (HIP-graph stages)", "value": out["graphs"]["seconds"],
        "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": out["graphs"]["seconds"] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded meshes, inflow conditions)",
        "config": {"workload": "batch_newton", "members": n_members, "Nele_total": nele,
                   "linear_solver": "block ILU(0) BiCGStab (GPU), all members batched", "parallelism": "1 GPU"},
        "graphs": out["graphs"], "eager": out["eager"], "graphs_member_solves": out["graphs_member_solves"],
        "note": "the first graphs run includes capture; the value is the second graphs run",
    }


def bench_serve(n_requests: int = 500, n_meshes: int = 24, slots: int = 8, log=print) -> dict:
    """500 R/K requests (Newton step : line search = 1:3) over meshes of 2k-300k elements
    (lognormal, P1/P2/P3 = 50/35/15 %), Poisson arrivals at fractions of the saturated rate."""
    import torch
    from uasource_amd.serve import UvhServer, make_requests
    meshes = []
    t0 = time.perf_counter()
    for w in serve_workloads(n_meshes, seed=17):
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        meshes.append((M, F0, F1))
    cv = default_ctrlvar()
    srv = UvhServer(cv, meshes, slots=slots, hbm_budget=32e9, keep_outputs=False)
    log(f"serve: {n_meshes} meshes, {sum(m[0].Nele for m in meshes)} elements, set-up {time.perf_counter() - t0:.1f} s")
    # every request list is made (its states pinned) before anything is timed; the sweep
    # replays one list (arrivals at rate 1, rescaled per load: same meshes, kinds and states)
    sat_reqs = make_requests(meshes, n_requests, float("inf"), seed=11, pinned=True)
    paced = make_requests(meshes, n_requests, 1.0, seed=12, pinned=True)
    warm = make_requests(meshes, 40, 1.0, seed=13, pinned=True)
    base_arrival = {id(r): r.arrival for r in sat_reqs + paced + warm}

    def at_rate(rs, rate):
        for r in rs:
            r.arrival = base_arrival[id(r)] / rate
            r.start = r.done = float("nan")
            r.bypassed = 0
        return rs
    # warm-up, discarded: one saturated run and one pass over the loads. About one HIP call
    # in 10^4 (a state upload or a kernel launch, normally microseconds) holds the host for
    # 6-13 ms, which stalls every request in the system; most of these come in the first
    # ~8k calls of a process. The server counts such stalls (UvhServer.stalls) and each
    # sweep point reports them.
    srv.run(sat_reqs)
    warm_rate = n_requests / max(r.done for r in sat_reqs)
    for frac in (0.3, 0.5, 0.7, 0.9):
        srv.run(at_rate(paced, frac * warm_rate))
    # service time of each (mesh, kind) alone: the median of five after one untimed launch
    from uasource_amd.serve import Request
    alone = {}
    for mi in range(n_meshes):
        st = [torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
              for a in (meshes[mi][2].ub, meshes[mi][2].vb, meshes[mi][2].h)]
        for newton in (False, True):
            r = Request(0, mi, newton, 0.0, *st)
            srv._launch(r, 0)
            torch.cuda.synchronize()
            ts = []
            for _ in range(5):
                t = time.perf_counter()
                srv._launch(r, 0)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t)
            alone[(mi, newton)] = float(np.median(ts))
    # saturation: everything at t = 0; one burst lasts ~25 ms, so the median of five
    spans = []
    for _ in range(5):
        srv.run(at_rate(sat_reqs, 1.0))
        spans.append(max(r.done for r in sat_reqs))
    makespan = float(np.median(spans))
    sat_rate = n_requests / makespan
    elements = sum(meshes[r.mesh][0].Nele for r in sat_reqs)
    log(f"serve: saturated {sat_rate:.0f} requests/s, {elements / makespan / 1e6:.1f} M elements/s")
    # size classes: the meshes in thirds by element count (latency percentiles of a mix whose
    # service times span two orders of magnitude say more per class than overall)
    by_size = sorted(range(n_meshes), key=lambda mi: meshes[mi][0].Nele)
    third = n_meshes // 3
    size_classes = {"small": set(by_size[:third]), "medium": set(by_size[third:2 * third]),
                    "large": set(by_size[2 * third:])}
    sweep = []
    for frac, policy in ((0.3, "fifo"), (0.5, "fifo"), (0.7, "fifo"), (0.9, "fifo"), (0.7, "sjf")):
        srv.policy = policy
        srv.run(at_rate(warm, frac * sat_rate))
        reqs = at_rate(paced, frac * sat_rate)
        srv.run(reqs)
        lat = np.array([r.latency for r in reqs])
        slow = np.array([r.latency / alone[(r.mesh, r.newton)] for r in reqs])
        svc = np.array([alone[(r.mesh, r.newton)] for r in reqs])   # the mix's own service times
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        el = sum(meshes[r.mesh][0].Nele for r in reqs)
        classes = {}
        for cname, members in size_classes.items():
            sel = [k for k, r in enumerate(reqs) if r.mesh in members]
            if not sel:
                continue
            cl, cs = lat[sel], slow[sel]
            classes[cname] = {"meshes_Nele": [int(meshes[mi][0].Nele) for mi in sorted(members)], "requests": len(sel),
                              "p50_ms": float(np.percentile(cl, 50) * 1e3), "p99_ms": float(np.percentile(cl, 99) * 1e3),
                              "p99_over_p50": float(np.percentile(cl, 99) / np.percentile(cl, 50)),
                              "slowdown_p50": float(np.percentile(cs, 50)), "slowdown_p99": float(np.percentile(cs, 99)),
                              "service_alone_p99_over_p50": float(np.percentile(svc[sel], 99) / np.percentile(svc[sel], 50))}
        row = {"load": frac, "policy": policy, "rate_per_s": frac * sat_rate,
               "p50_ms": float(np.percentile(lat, 50) * 1e3),
               "p99_ms": float(np.percentile(lat, 99) * 1e3), "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
               "slowdown_p50": float(np.percentile(slow, 50)), "slowdown_p99": float(np.percentile(slow, 99)),
               "service_alone_p50_ms": float(np.percentile(svc, 50) * 1e3),
               "service_alone_p99_ms": float(np.percentile(svc, 99) * 1e3),
               "service_alone_p99_over_p50": float(np.percentile(svc, 99) / np.percentile(svc, 50)),
               "elements_per_s": el / span, "peak_inflight": srv.peak_inflight,
               "host_stalls_ms": [round(x * 1e3, 2) for x in srv.stalls], "size_classes": classes}
        sweep.append(row)
        log(f"serve load {frac} ({policy}): p50 {row['p50_ms']:.3f} ms p99 {row['p99_ms']:.3f} ms, slowdown p50 "
            f"{row['slowdown_p50']:.2f} p99 {row['slowdown_p99']:.2f}, host stalls {row['host_stalls_ms']}")
        for cname, c in classes.items():
            log(f"   {cname}: p50 {c['p50_ms']:.3f} ms p99 {c['p99_ms']:.3f} ms (p99/p50 {c['p99_over_p50']:.2f}; "
                f"alone {c['service_alone_p99_over_p50']:.2f}), slowdown p50 {c['slowdown_p50']:.2f} "
                f"p99 {c['slowdown_p99']:.2f}")
    srv.close()
    srv.policy = "fifo"
    at70 = next(r for r in sweep if r["load"] == 0.7 and r["policy"] == "fifo")
    return {
        "metric": "uvh request latency p50/p99 at 70% of saturated load",
        "value": at70["p99_ms"], "unit": "ms (p99 latency)", "n_gpus": 1, "steps": n_requests, "warmup": 0,
        "ms_per_step": at70["p50_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded meshes and Poisson arrivals)",
        "config": {"workload": "serve500", "meshes": n_meshes, "slots": slots,
                   "Nele_range": [int(min(m[0].Nele for m in meshes)), int(max(m[0].Nele for m in meshes))],
                   "parallelism": "1 GPU"},
        "saturated_requests_per_s": sat_rate, "saturated_elements_per_s": elements / makespan,
        "saturated_makespans_ms": [float(s * 1e3) for s in spans], "sweep": sweep,
    }


def bench_newton(workload: str = "newton_p2", device_solver: bool = False, log=print) -> dict:
    """Full Newton uvh solve (Úa's default tolerance rForce < 1e-15) with the GPU assembly,
    FP64 and the precision form; the linear solve on the host (SciPy SuperLU) or, with
    ``device_solver``, on the GPU (block ILU(0) BiCGStab to LinSolveTol = 1e-8)."""
    from uasource_amd import newton
    from uasource_amd.linsolve import GpuLinearSolver
    w = WORKLOADS[workload]
    M = build_mesh(w)
    F0, F1, BCs, fixed = newton_problem(M, w)
    cv = default_ctrlvar()
    out = {}
    for fp32 in (False, True):
        t = time.perf_counter()
        asm = newton.GpuAssembler(cv, M, F0, F1)
        solver = GpuLinearSolver(asm.ctx, fixed, rtol=cv.LinSolveTol) if device_solver else newton.superlu_solve
        log(f"newton {'fp32' if fp32 else 'fp64'} on {workload} ({M.Nele} elements, {3 * M.Nnodes} unknowns)")
        res = newton.solve(asm, cv, linear_solver=solver, fixed=fixed, fp32=fp32, verbose=True, log=log)
        total = time.perf_counter() - t
        out["fp32" if fp32 else "fp64"] = {
            "converged": bool(res.converged), "iterations": int(res.iterations), "rForce": float(res.rForce),
            "seconds": total, "reason": res.reason,
            "assembly_s": res.t_assembly, "solve_s": res.t_solve, "linesearch_s": res.t_linesearch,
            "rk_assemblies": res.n_rk, "r_assemblies": res.n_r, "state": (res.ub, res.vb, res.h)}
        if device_solver:
            out["fp32" if fp32 else "fp64"]["krylov_iterations"] = res.linear_iterations
            solver.close()
        asm.ctx.close()
        log(f"newton {'fp32' if fp32 else 'fp64'}: {res.iterations} iterations, {res.reason}, {total:.1f} s "
            f"(assembly {res.t_assembly:.2f} s, solve {res.t_solve:.2f} s, line search {res.t_linesearch:.2f} s)"
            + (f", Krylov iterations {res.linear_iterations}" if device_solver else ""))
    u64, v64, h64 = out["fp64"].pop("state")
    u32, v32, h32 = out["fp32"].pop("state")
    speed = np.linalg.norm(np.concatenate([u64, v64]))
    err = {"u": float(np.linalg.norm(u32 - u64) / speed), "v": float(np.linalg.norm(v32 - v64) / speed),
           "h": float(np.linalg.norm(h32 - h64) / np.linalg.norm(h64))}
    return {
        "metric": "uvh Newton solve to rForce < 1e-15", "value": out["fp64"]["seconds"], "unit": "s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": out["fp64"]["seconds"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mesh, inflow conditions)",
        "config": {"workload": workload, "Nele": M.Nele, "Nnodes": M.Nnodes,
                   "linear_solver": "block ILU(0) BiCGStab (GPU), LinSolveTol 1e-8" if device_solver
                   else "SciPy SuperLU (host; minimum-degree order, COLAMD fall-back)", "parallelism": "1 GPU"},
        "fp64": out["fp64"], "fp32_form": out["fp32"], "fp32_form_rel_err_vs_fp64": err,
    }


def transient_start_state(workload: str = "plain_p2", device_solver: bool = True, log=print) -> dict:
    """The start of a transient run as Úa makes it (Ua2D.m:598-611): a uv (diagnostic) Newton
    solve on F1's geometry with the GPU uv assembly; its velocities become F0's and F1's.
    Returns the mesh, the Dirichlet set, CtrlVar, the balanced (G0, G1), the uv result and its
    time."""
    from uasource_amd import newton
    w = WORKLOADS[workload]
    M = build_mesh(w)
    F0, F1, BCs, fixed = newton_problem(M, w)
    cv = default_ctrlvar()
    ctx = uvh.UvhContext(0)
    t0 = time.perf_counter()
    log(f"transient start on {workload} ({M.Nele} elements): uv solve ({2 * M.Nnodes} unknowns)")
    try:
        G0, G1, ruv = newton.transient_start(cv, M, F0, F1, fixed,
                                             lambda c, m, F: newton.GpuUvAssembler(c, m, F, ctx=ctx), verbose=True,
                                             log=log, device_solver=device_solver)
    finally:
        ctx.close()
    t_uv = time.perf_counter() - t0
    log(f"uv: {ruv.iterations} iterations, {ruv.reason}, rForce {ruv.rForce:.3e}, {t_uv:.1f} s")
    return {"M": M, "fixed": fixed, "cv": cv, "G0": G0, "G1": G1, "uv": ruv, "t_uv": t_uv, "workload": workload}


def uvh_solve(start: dict, fp32: bool = False, device_solver: bool = True, log=print):
    """The implicit uvh Newton solve (SSTREAM_TransientImplicit.m) from a transient start's
    balanced state, FP64 or in the precision form; returns (NewtonResult, seconds)."""
    from uasource_amd import newton
    from uasource_amd.linsolve import GpuLinearSolver
    cv, fixed = start["cv"], start["fixed"]
    t1 = time.perf_counter()
    asm = newton.GpuAssembler(cv, start["M"], start["G0"], start["G1"])
    solver = GpuLinearSolver(asm.ctx, fixed, rtol=cv.LinSolveTol) if device_solver else newton.superlu_solve
    try:
        res = newton.solve(asm, cv, linear_solver=solver, fixed=fixed, verbose=True, log=log, fp32=fp32)
    finally:
        if device_solver:
            solver.close()
        asm.ctx.close()
    t = time.perf_counter() - t1
    log(f"uvh{' (precision form)' if fp32 else ''}: {res.iterations} iterations, {res.reason}, "
        f"rForce {res.rForce:.3e}, {t:.1f} s")
    return res, t


def solution_errors(a, b) -> dict:
    """Relative errors of the converged u, v, h of ``a`` against ``b``: normwise, and the
    largest entry over the largest magnitude; u and v against the speed scale of (u, v) (v is
    near zero in these problems, so its own norm says little), h against h."""
    speed = np.linalg.norm(np.concatenate([b.ub, b.vb]))
    smax = max(np.max(np.abs(b.ub)), np.max(np.abs(b.vb)))
    out = {}
    for k in ("ub", "vb", "h"):
        x, y = getattr(a, k), getattr(b, k)
        n, m = (speed, smax) if k != "h" else (np.linalg.norm(y), np.max(np.abs(y)))
        out[k] = {"normwise": float(np.linalg.norm(x - y) / n), "max": float(np.max(np.abs(x - y)) / m)}
    return out


def bench_newton_transient(workload: str = "plain_p2", device_solver: bool = True, log=print,
                           precision: bool = False) -> dict:
    """The transient start (``transient_start_state``), then the implicit uvh Newton solve from
    that balanced state with the GPU assembly; with ``device_solver`` both take the device
    block-ILU(0) Krylov solve (the uv steps on the uvh solver with the thickness held), else host
    SuperLU; to Úa's default tolerance rForce < 1e-15. FP64; with ``precision`` the uvh solve is
    repeated in the precision form from the same start, and its converged u, v, h are compared
    with the FP64 ones."""
    start = transient_start_state(workload, device_solver, log)
    ruv, t_uv, M = start["uv"], start["t_uv"], start["M"]
    res, t_uvh = uvh_solve(start, False, device_solver, log)
    hist = lambda r: [[float(v) for v in h[:4]] for h in r.history]  # noqa: E731
    out = {
        "metric": "uvh Newton solve to rForce < 1e-15 (transient start)", "value": t_uv + t_uvh, "unit": "s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": (t_uv + t_uvh) * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mesh, inflow conditions)",
        "config": {"workload": workload, "Nele": M.Nele, "Nnodes": M.Nnodes,
                   "linear_solver": ("block ILU(0) Krylov (GPU)" if device_solver else "SciPy SuperLU (host)")
                   + " for uv and uvh", "parallelism": "1 GPU"},
        "uv": {"converged": bool(ruv.converged), "iterations": int(ruv.iterations), "rForce": float(ruv.rForce),
               "seconds": t_uv, "history": hist(ruv)},
        "uvh": {"converged": bool(res.converged), "iterations": int(res.iterations), "rForce": float(res.rForce),
                "reason": res.reason, "seconds": t_uvh, "assembly_s": res.t_assembly, "solve_s": res.t_solve,
                "linesearch_s": res.t_linesearch, "history": hist(res),
                "krylov_iterations": res.linear_iterations if device_solver else None},
    }
    if precision:
        r32, t32 = uvh_solve(start, True, device_solver, log)
        out["uvh_precision_form"] = {"converged": bool(r32.converged), "iterations": int(r32.iterations),
                                     "rForce": float(r32.rForce), "reason": r32.reason, "seconds": t32,
                                     "history": [[float(v) for v in h[:4]] + [bool(h[4])] for h in r32.history],
                                     "errors_vs_fp64": solution_errors(r32, res)}
    return out


def bench_newton_long(workload: str, time_limit: float, checkpoint: str, line_search: str = "-Newton Step-",
                      log=print) -> dict:
    """FP64 Newton solve of a workload too large for one call (the full config-2 mesh: ~65 s
    per host factorization): runs for ``time_limit`` seconds, then saves the iteration's state
    to ``checkpoint`` (.npz of plain arrays); the next call resumes from it. The line reports the
    totals over all calls so far and ``"status": "converged"`` once done."""
    import os
    from uasource_amd import newton
    w = WORKLOADS[workload]
    M = build_mesh(w)
    F0, F1, BCs, fixed = newton_problem(M, w)
    cv = default_ctrlvar()
    cv.rLineMinUa = line_search
    ck, acc = None, {"seconds": 0.0, "assembly_s": 0.0, "solve_s": 0.0, "linesearch_s": 0.0, "rk": 0, "r": 0,
                     "calls": 0, "cauchy": 0}
    if os.path.exists(checkpoint):
        z = np.load(checkpoint)
        ck = {k: z[k] for k in ("iteration", "gamma", "r", "r0", "rForce", "rWork", "residual64", "backtrack_ok",
                                "ratios", "history", "x")}
        acc = {k: float(z["acc_" + k]) for k in acc}
        log(f"newton {workload}: resuming at iteration {int(ck['iteration'])}")
    t = time.perf_counter()
    asm = newton.GpuAssembler(cv, M, F0, F1)
    log(f"newton fp64 on {workload} ({M.Nele} elements, {3 * M.Nnodes} unknowns), time limit {time_limit:.0f} s")
    res = newton.solve(asm, cv, fixed=fixed, verbose=True, log=log, time_limit=time_limit, resume=ck)
    acc["seconds"] += time.perf_counter() - t
    acc["assembly_s"] += res.t_assembly
    acc["solve_s"] += res.t_solve
    acc["linesearch_s"] += res.t_linesearch
    acc["rk"] += res.n_rk
    acc["r"] += res.n_r
    acc["calls"] += 1
    acc["cauchy"] = acc.get("cauchy", 0.0) + res.cauchy_steps
    asm.ctx.close()
    if res.checkpoint is not None:
        c = res.checkpoint
        np.savez(checkpoint, **{k: np.asarray(c[k], dtype=float) for k in c},
                 **{"acc_" + k: np.float64(v) for k, v in acc.items()})
    elif os.path.exists(checkpoint):
        os.remove(checkpoint)
    hist = [[float(v) for v in h[:4]] for h in res.history]
    log(f"newton fp64: {res.iterations} iterations, {res.reason}, {acc['seconds']:.0f} s over {int(acc['calls'])} calls")
    return {
        "metric": "uvh Newton solve to rForce < 1e-15", "value": acc["seconds"], "unit": "s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": acc["seconds"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mesh, inflow conditions)",
        "config": {"workload": workload, "Nele": M.Nele, "Nnodes": M.Nnodes,
                   "linear_solver": "SciPy SuperLU (host, minimum degree)", "parallelism": "1 GPU"},
        "status": "converged" if res.converged else res.reason, "iterations": int(res.iterations),
        "rForce": float(res.rForce), "calls": int(acc["calls"]), "assembly_s": acc["assembly_s"],
        "solve_s": acc["solve_s"], "linesearch_s": acc["linesearch_s"], "rk_assemblies": int(acc["rk"]),
        "r_assemblies": int(acc["r"]), "line_search": line_search, "cauchy_steps": int(acc["cauchy"]),
        "history": hist,
    }


def bench_uv(steps: int, warmup: int, workload: str = "plain_p2", cpu_seconds: float = 10.0, log=print) -> dict:
    """uv (diagnostic SSTREAM) assembly, R and the 2 x 2-block K, of the config-2 mesh (same
    fields as the uvh workload, s from the floatation of h): FP64 and the precision form, with
    the oracle's C restatement of uvMatrixAssemblySSTREAM timed on the host beside it."""
    import os
    import torch
    from uasource_amd.fields import surface_from_floatation
    dev = torch.device("cuda:0")
    w = WORKLOADS[workload]
    M = build_mesh(w)
    _, F1, _ = make_fields(M, w)
    cv = default_ctrlvar()
    F = F1.copy()
    F.s = surface_from_floatation(F1, cv)
    ctx = uvh.UvhContext(0)
    ctx.set_mesh(M)
    keep = []

    class _F:
        pass
    fd = _F()
    for k in uvh.UV_KEYS:
        t = torch.from_numpy(np.ascontiguousarray(getattr(F, k), dtype=np.float64)).to(dev)
        keep.append(t)
        setattr(fd, k, t)
    fields = uvh.make_uv_fields(fd)
    ctrl, uvc = uvh.make_ctrl(cv, F), uvh.make_uv_ctrl(cv)
    nnzN = ctx.nnz // 9
    R = torch.zeros(2 * M.Nnodes, dtype=torch.float64, device=dev)
    K = torch.zeros(4 * nnzN, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream()
    out = {}
    for fp32 in (False, True):
        flags = uvh.UVH_NO_NAN_CHECK | (uvh.UVH_FP32 if fp32 else 0)
        run = lambda: ctx.assemble_uv_device(ctrl, uvc, fields, flags, R, K, stream=s.cuda_stream)  # noqa: E731
        for _ in range(warmup):
            run()
        ms = _time(torch, run, steps, s)
        ctx.profile(True)            # per-kernel times from a separate pass
        _time(torch, run, steps, s)
        kt = {k: 1e3 * v[0] / max(v[1], 1) for k, v in ctx.kernel_times().items()}
        ctx.profile(False)
        out["fp32" if fp32 else "fp64"] = {"ms": ms, "elements_per_s": M.Nele / (ms * 1e-3),
                                           "kernels_us": {k: round(v, 1) for k, v in kt.items()}}
        log(f"uv {'fp32' if fp32 else 'fp64'}: {ms:.3f} ms, {M.Nele / (ms * 1e-3) / 1e6:.1f} M el/s, kernels {kt}")
    # algorithmic bytes: 11 nodal inputs and 2 outputs (FP64), connectivity, K written once
    nbytes = M.Nnodes * (11 + 2) * 8 + M.Nele * M.nod * 4 + 4 * nnzN * 8
    ms64 = out["fp64"]["ms"]
    cpu = None
    try:
        from oracle import o2
        threads = min(16, os.cpu_count() or 1)
        o2.assemble_uv(cv, M, F, nthreads=threads)      # warm
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < cpu_seconds / 2 or n == 0:
            o2.assemble_uv(cv, M, F, nthreads=threads)
            n += 1
        cpu = {"value": n * M.Nele / (time.perf_counter() - t0), "unit": "elements/s", "cores": threads,
               "kind": "port", "sample": f"{n} full uv assemblies of {workload}"}
    except Exception as exc:   # the oracle may be absent where only the library travels
        log(f"uv CPU baseline skipped: {exc}")
    ctx.close()
    return {
        "metric": "uv elements assembled/s (R+K), diagnostic SSTREAM", "value": M.Nele / (ms64 * 1e-3),
        "unit": "elements/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms64,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered-Delaunay mesh, config-2 fields, s from floatation)",
        "config": {"workload": f"uv_{workload}", "Nele": M.Nele, "Nnodes": M.Nnodes, "nod": M.nod, "nip": M.nip,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": nbytes / (ms64 * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": nbytes / (ms64 * 1e-3) / 1e9 / 8000.0, "traffic": None},
        "fp64": out["fp64"], "fp32_form": out["fp32"], "cpu_baseline": cpu,
    }
