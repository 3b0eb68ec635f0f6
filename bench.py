"""Benchmark of the uvh R+K assembly on one MI355X.

A step is one full Newton-Raphson R+K assembly (``uvhAssembly`` with
``Ronly=false``) of the plain FP64 workload: 199,712 six-node triangles,
401,005 nodes, 7 integration points, with every input already resident in HBM.
The line printed reports elements assembled per second, the HBM-roofline
fraction of the assembly, per-kernel times from HIP events on the launch
stream, and the oracle's C restatement timed on this host's cores beside it.

    python bench.py [--steps K] [--warmup W] [--workload plain_p2] [--ronly] [--fp32]

Other forms of the path, one JSON line each: ``--batch`` (64 meshes in one persistent
launch per stage), ``--serve`` (500 requests, latency percentiles over a load sweep),
``--newton`` (a full Newton solve), ``--uv`` (the uv diagnostic assembly).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def algorithmic_bytes(nn: int, nele: int, nod: int, nnz: int, ronly: bool) -> int:
    """Compulsory DRAM bytes of one assembly: 19 FP64 nodal inputs read once
    (F1: h,u,v,S,B,rho,C,m,n,A,as,ab; F0: h,u,v,as,ab; x,y), 3 FP64 residual
    outputs per node, int32 connectivity, and K values written once."""
    b = nn * (19 * 8 + 3 * 8) + nele * nod * 4
    if not ronly:
        b += nnz * 8
    return b


def cpu_baseline(workload: str, M, F0, F1, cv, ronly: bool, budget_s: float):
    """Oracle O2 (C restatement, OpenMP) on this host: elements/s on the same workload."""
    from oracle import o2
    o2.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        avail = os.cpu_count() or 1
    threads = min(avail, threads) if threads > 0 else min(avail, 16)
    mesh = o2.O2Mesh(M)
    o2.assemble(cv, M, F0, F1, mesh=mesh, nthreads=threads, ronly=ronly)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        o2.assemble(cv, M, F0, F1, mesh=mesh, nthreads=threads, ronly=ronly)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 50:
            break
    times.sort()
    med = times[len(times) // 2]
    out = {"value": M.Nele / med, "unit": "elements/s", "cores": threads, "kind": "port",
           "sample": f"{workload}: {len(times)} full R{'' if ronly else '+K'} assemblies of {M.Nele} elements "
                     f"(median {med * 1e3:.1f} ms), oracle O2 C restatement of Úa with {threads} OpenMP threads"}

    def bounded(nthreads, r_only, seconds):
        ts, t_end = [], time.perf_counter() + seconds
        while not ts or (time.perf_counter() < t_end and len(ts) < 10):
            t0 = time.perf_counter()
            o2.assemble(cv, M, F0, F1, mesh=mesh, nthreads=nthreads, ronly=r_only)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return {"value": M.Nele / ts[len(ts) // 2], "runs": len(ts), "median_ms": ts[len(ts) // 2] * 1e3}
    # the same assembly on one core, and the R-only assembly on all of them (bounded samples)
    out["single_core"] = bounded(1, ronly, budget_s * 0.6)
    if not ronly:
        bounded(threads, True, 0.0)   # warm-up
        out["r_only"] = bounded(threads, True, budget_s * 0.3)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain_p2")
    ap.add_argument("--ronly", action="store_true")
    ap.add_argument("--fp32", action="store_true",
                    help="precision form: FP32 element arithmetic, FP64 accumulation (reports its error vs FP64)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", action="store_true", help="64-mesh batched assembly")
    ap.add_argument("--batch-newton", action="store_true", help="batched Newton solves with HIP-graph stages")
    ap.add_argument("--serve", action="store_true", help="request serving, latency over a load sweep")
    ap.add_argument("--newton", action="store_true",
                    help="full Newton solve (newton_p2 with the host solver; --workload with --device-solver)")
    ap.add_argument("--uv", action="store_true", help="uv (diagnostic) assembly of the config-2 mesh")
    ap.add_argument("--device-solver", action="store_true",
                    help="with --newton: linear solves on the GPU (block ILU(0) BiCGStab)")
    ap.add_argument("--checkpoint", default="", metavar="FILE",
                    help="with --newton: FP64 solve of --workload in calls of --time-limit seconds, resumed "
                         "from FILE (.npz) until it converges")
    ap.add_argument("--time-limit", type=float, default=900.0, help="with --checkpoint: seconds per call")
    ap.add_argument("--line-search", choices=("newton", "auto"), default="newton",
                    help='with --checkpoint: the Newton direction only, or rLineminUa\'s "-Auto-" with the '
                         "Cauchy step with the mass matrix")
    ap.add_argument("--transient-start", action="store_true",
                    help="with --newton: a uv solve first, then the uvh solve from that state (Ua2D.m:598-611); "
                         "with --fp32 the uvh solve again in the precision form, with its u, v, h errors")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes three times (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the uvh assembly is built and measured on a single MI355X "
              f"(--gpus {args.gpus} is not supported; independent meshes shard across GPUs with no exchange).",
              file=sys.stderr)
        return 2

    if args.batch or args.serve or args.newton or args.uv or args.batch_newton:
        from measure import bench_modes
        log = lambda m: print(m, file=sys.stderr, flush=True)
        if args.batch_newton:
            line = bench_modes.bench_batch_newton(log=log)
        elif args.batch:
            line = bench_modes.bench_batch(args.steps, args.warmup, log=log)
        elif args.serve:
            line = bench_modes.bench_serve(log=log)
        elif args.uv:
            line = bench_modes.bench_uv(args.steps, args.warmup, cpu_seconds=args.cpu_seconds, log=log)
        elif args.transient_start:
            line = bench_modes.bench_newton_transient(args.workload, device_solver=args.device_solver, log=log,
                                                      precision=args.fp32)
        elif args.checkpoint:
            line = bench_modes.bench_newton_long(args.workload, args.time_limit, args.checkpoint,
                                                 "-Auto-" if args.line_search == "auto" else "-Newton Step-", log=log)
        else:
            line = bench_modes.bench_newton("newton_p2" if args.workload == "plain_p2" and not args.device_solver
                                            else args.workload, device_solver=args.device_solver, log=log)
        print(json.dumps(line), flush=True)
        return 0

    import numpy as np
    import torch

    from uasource_amd import uvh
    from uasource_amd.ctrlvar import default_ctrlvar
    from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

    w = WORKLOADS[args.workload]
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    cv = default_ctrlvar()
    dev = torch.device("cuda:0")
    ctx = uvh.UvhContext(0)
    ctx.set_mesh(M)

    class Dev:
        pass
    d1, d0 = Dev(), Dev()
    for k in uvh.F1_KEYS:
        setattr(d1, k, torch.from_numpy(np.ascontiguousarray(getattr(F1, k))).to(dev))
    for k in uvh.F0_KEYS:
        setattr(d0, k, torch.from_numpy(np.ascontiguousarray(getattr(F0, k))).to(dev))
    fields = uvh.make_fields(d0, d1)
    ctrl = uvh.make_ctrl(cv, F1)
    R = torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev)
    K = None if args.ronly else torch.zeros(ctx.nnz, dtype=torch.float64, device=dev)
    flags = (uvh.UVH_RONLY if args.ronly else 0) | (uvh.UVH_FP32 if args.fp32 else 0)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def step():
        ctx.assemble_device(ctrl, fields, flags, R, K, None, None, stream=sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.sync_status()

    t_start = torch.cuda.Event(enable_timing=True)
    t_stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for _ in range(args.steps):
        step()
    t_stop.record(stream)
    torch.cuda.synchronize()
    ms_total = t_start.elapsed_time(t_stop)
    ctx.sync_status()
    ms_per_step = ms_total / args.steps

    precision = None
    if args.fp32:
        # error of the precision form against the FP64 path on the same inputs (the FP64 path
        # is the one held to 1e-12 of the oracle)
        R64 = torch.zeros_like(R)
        K64 = None if K is None else torch.zeros_like(K)
        T32, F32 = torch.zeros_like(R), torch.zeros_like(R)
        T64, F64 = torch.zeros_like(R), torch.zeros_like(R)
        ctx.assemble_device(ctrl, fields, flags, R, K, T32, F32, stream=sptr)
        ctx.assemble_device(ctrl, fields, flags & ~uvh.UVH_FP32, R64, K64, T64, F64, stream=sptr)
        ctx.sync_status()
        scale = max(torch.linalg.norm(T64).item(), torch.linalg.norm(F64).item())
        precision = {"R_rel_err": torch.linalg.norm(R - R64).item() / scale,
                     "Tint_rel_err": (torch.linalg.norm(T32 - T64) / torch.linalg.norm(T64)).item(),
                     "Fext_rel_err": (torch.linalg.norm(F32 - F64) / torch.linalg.norm(F64)).item(),
                     "reference": "FP64 path of this library on the same inputs"}
        if K is not None:
            precision["K_rel_err"] = (torch.linalg.norm(K - K64) / torch.linalg.norm(K64)).item()
        del R64, K64, T32, F32, T64, F64

    # per-kernel durations: HIP events around each launch, on the launch stream
    ctx.profile(True)
    for _ in range(args.steps):
        step()
    times = ctx.kernel_times()
    ctx.profile(False)
    # per assembly: a kernel may be launched several times per assembly (one launch per chunk
    # in the pipelined form), and launches on the two streams overlap, so the roofline uses the
    # assembly's duration between events on the caller's stream
    kernels = {name: tot / args.steps for name, (tot, cnt) in times.items()}
    launches = {name: cnt / args.steps for name, (tot, cnt) in times.items()}
    assembly_ms = ms_per_step
    dominant = max(kernels, key=kernels.get)

    # roofline of the dominant kernel: the algorithmic bytes of one assembly (all of which that
    # kernel's one launch per assembly works through) over its average launch time (HIP events
    # on its stream); the whole assembly's rate beside it
    nbytes = algorithmic_bytes(M.Nnodes, M.Nele, M.nod, ctx.nnz, args.ronly)
    dom_ms = kernels[dominant] / max(1.0, launches[dominant])
    achieved = nbytes / (dom_ms * 1e-3) / 1e9
    achieved_assembly = nbytes / (assembly_ms * 1e-3) / 1e9
    traffic = traffic_assembly = None
    traffic_source = None
    tfile = os.path.join(REPO, "profiles", f"traffic_{args.workload}{'_fp32' if args.fp32 else ''}"
                                           f"{'_ronly' if args.ronly else ''}.json")
    if os.path.exists(tfile):
        from measure.profiling import kernel_source_hash
        with open(tfile) as fh:
            tinfo = json.load(fh)
        # the committed counter measurement applies only to the build made from the same kernel
        # sources, with the same kernels in the assembly
        if tinfo.get("source_sha256") == kernel_source_hash() and sorted(tinfo.get("kernels", [])) == sorted(kernels):
            sym = tinfo["symbols"].get(dominant)
            if sym is not None:
                b = tinfo["per_kernel_bytes"][sym]
                traffic = b["fetch"] + b["write"]
            traffic_assembly = tinfo.get("hbm_bytes_per_assembly")
            traffic_source = os.path.relpath(tfile, REPO)

    # VALU use of the element kernel under the factored flop model of the element arithmetic,
    # nip (42 nod^2 + 90 nod + 150) per element for R+K (the arithmetic is not counted
    # instruction by instruction), against the vector peak of the element arithmetic's type
    valu = None
    ek = [k for k in kernels if k.startswith("element")]
    if ek and not args.ronly:
        flops = M.Nele * M.nip * (42 * M.nod ** 2 + 90 * M.nod + 150)
        tflops = flops / (sum(kernels[k] for k in ek) * 1e-3) / 1e12
        vpeak = 157.3 if args.fp32 else 78.6
        valu = {"kernel": "+".join(ek), "model_flops": flops, "achieved": tflops, "peak": vpeak,
                "unit": "TFLOP/s", "frac": tflops / vpeak,
                "model": "nip*(42*nod^2 + 90*nod + 150) flops per element"}

    line = {
        "metric": "uvh elements assembled/s (R+K), HBM GB/s fraction; p50/p99 request latency",
        "value": M.Nele * args.steps / (ms_total * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 element arithmetic, f64 accumulation" if args.fp32 else "f64",
        "data": "synthetic (seeded jittered-Delaunay mesh and field recipe)",
        "config": {"workload": args.workload, "Nele": M.Nele, "Nnodes": M.Nnodes, "nod": M.nod, "nip": M.nip,
                   "nnz_K": ctx.nnz, "mode": "R only" if args.ronly else "R+K", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dominant, "kernel_us": dom_ms * 1e3, "algorithmic_bytes": nbytes,
                     "traffic_source": traffic_source,
                     "assembly": {"achieved": achieved_assembly, "frac": achieved_assembly / HBM_PEAK_GBS,
                                  "us": assembly_ms * 1e3, "traffic": traffic_assembly,
                                  "note": "all kernels of one assembly, between events on its stream"}},
        "valu": valu,
        "kernels_us": {k: v * 1e3 for k, v in kernels.items()},
        "kernel_launches_per_assembly": launches,
        "precision": precision,
        "dominant_kernel": dominant,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, M, F0, F1, cv, args.ronly, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
