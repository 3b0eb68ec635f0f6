"""Summaries of rocprofv3 runs of ``bench.py``: per-kernel durations and HBM bytes.

Two subcommands, both reading rocprofv3 CSV output and writing short text under
``profiles/``:

``stats <trace_dir> <workload> [--fp32] [--ronly] [--command CMD] [--phases W,T,E]``
    turns ``rocprofv3 --kernel-trace --stats`` output (``run_kernel_stats.csv``) into
    ``profiles/kernel_stats_<workload>[_ronly].md``. With ``--phases`` the per-call trace
    (``run_kernel_trace.csv``) of every kernel called W+T+E times is split in dispatch order into
    ``bench.py``'s W warm-up calls, T timed calls and E calls bracketed by HIP events, and the
    average duration of each phase is listed as well.

``traffic <dir> <workload> [--fp32] [--ronly]``
    turns four ``rocprofv3 --pmc`` passes into ``profiles/traffic_<workload>[_ronly].json``
    (HBM bytes per assembly, per kernel). ``<dir>`` holds ``fetch/`` and ``write/`` (FETCH_SIZE
    and WRITE_SIZE over ``bench.py``) and ``cfetch/`` and ``cwrite/`` (the same counters over
    ``bench.py --copy-calibration 1073741824``, a plain device copy of 1 GiB). The copy reads
    and writes exactly 1 GiB, so counter-KiB x 1024 / 2^30 is the counter's bytes per true
    byte; every kernel's counter is divided by that factor. The passes are run as::

      rocprofv3 --pmc FETCH_SIZE -d <dir>/fetch -o run --output-format csv -- python3 bench.py --steps 3 --warmup 1 --no-cpu-baseline
      rocprofv3 --pmc WRITE_SIZE -d <dir>/write -o run --output-format csv -- python3 bench.py --steps 3 --warmup 1 --no-cpu-baseline
      rocprofv3 --pmc FETCH_SIZE -d <dir>/cfetch -o run --output-format csv -- python3 bench.py --copy-calibration 1073741824
      rocprofv3 --pmc WRITE_SIZE -d <dir>/cwrite -o run --output-format csv -- python3 bench.py --copy-calibration 1073741824
"""
from __future__ import annotations

import collections
import csv
import json
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(REPO, "profiles")

# kernel base name -> the name bench.py's per-kernel HIP-event timers use (without the _rk/_r suffix)
MODE_KERNELS = {"k_element": "element", "k_element_f32": "element", "k_rows": "rows", "k_rows_r": "rows",
                "k_fused": "fused"}
PLAIN_KERNELS = {"k_nodal_b": "nodal_b"}
COPY_BYTES = 1 << 30


ANON = "(anonymous namespace)::"


def signature_name(full: str) -> str:
    """Kernel name without its argument list; the "(anonymous namespace)::" qualifier of
    file-local kernels is kept."""
    head = full.replace(ANON, "\0")
    return head.split("(")[0].replace("\0", ANON).strip()


def base_name(full: str) -> str:
    return signature_name(full).split("<")[0].replace("void ", "").replace("uvh::", "").replace(ANON, "").strip()


def timer_name(full: str, ronly: bool) -> str | None:
    base = base_name(full)
    if base in MODE_KERNELS:
        fp32 = "float" in signature_name(full) and base in ("k_element", "k_rows", "k_rows_r")
        return MODE_KERNELS[base] + ("32" if fp32 else "") + ("_r" if ronly else "_rk")
    return PLAIN_KERNELS.get(base)


def per_kernel(path: str, counter: str) -> dict:
    """Average counter value (KiB) per dispatch, keyed by the kernel's full name."""
    with open(os.path.join(path, "run_counter_collection.csv")) as fh:
        rows = list(csv.DictReader(fh))
    tot = collections.defaultdict(float)
    cnt = collections.defaultdict(set)
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        name = signature_name(r["Kernel_Name"]).replace("void ", "").strip()
        tot[name] += float(r["Counter_Value"])
        cnt[name].add(r["Dispatch_Id"])
    return {k: tot[k] / len(cnt[k]) for k in tot}


KERNEL_SOURCES = ("uvh_kernels.hip", "uvh_api.hip", "uvh_internal.h")


def kernel_source_hash() -> str:
    """SHA-256 over the kernel sources of the library: a traffic record applies to the build
    made from exactly these sources."""
    import hashlib
    h = hashlib.sha256()
    for fn in KERNEL_SOURCES:
        with open(os.path.join(REPO, "uasource_amd", "csrc", fn), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def traffic(d: str, workload: str, ronly: bool, fp32: bool = False) -> dict:
    fetch = per_kernel(os.path.join(d, "fetch"), "FETCH_SIZE")
    write = per_kernel(os.path.join(d, "write"), "WRITE_SIZE")
    cf = per_kernel(os.path.join(d, "cfetch"), "FETCH_SIZE")
    cw = per_kernel(os.path.join(d, "cwrite"), "WRITE_SIZE")
    copy_kernel = max(cf, key=cf.get)
    fscale = cf[copy_kernel] * 1024 / COPY_BYTES
    wscale = cw[copy_kernel] * 1024 / COPY_BYTES
    out = {"workload": workload, "mode": "R only" if ronly else "R+K",
           "source_sha256": kernel_source_hash(),
           "copy_calibration": {"kernel": copy_kernel, "bytes": COPY_BYTES, "copies": 3,
                                "fetch_scale": fscale, "write_scale": wscale,
                                "note": "counter KiB x 1024 per true byte of a device copy of 1 GiB, "
                                        "averaged over its 3 dispatches"},
           "per_kernel_bytes": {}, "symbols": {}}
    total = 0.0
    names = set()
    for k in fetch:
        t = timer_name(k, ronly)
        if t is None or (t.startswith(("element", "rows")) and ("32" in t) != fp32):
            continue  # the precision-form run also makes one FP64 assembly for its error check
        fb = fetch[k] * 1024 / fscale
        wb = write.get(k, 0.0) * 1024 / wscale
        out["per_kernel_bytes"][k] = {"fetch": fb, "write": wb}
        out["symbols"][t] = k
        total += fb + wb
        names.add(t)
    out["hbm_bytes_per_assembly"] = total
    out["kernels"] = sorted(names)
    return out


def stats_md(trace_dir: str, workload: str, command: str) -> str:
    with open(os.path.join(trace_dir, "run_kernel_stats.csv")) as fh:
        rows = list(csv.DictReader(fh))
    lines = [f"# Kernel durations: uvh assembly, {workload}", "",
             "Command (repository root on the GPU host):", "", "```", command, "```", "",
             "| Kernel | Calls | Average (us) | Min (us) | Max (us) | Share |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = signature_name(r["Name"])
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f}% |")
    return "\n".join(lines) + "\n"


def phases_md(trace_dir: str, phases: tuple) -> str:
    """Average duration per bench phase of each kernel called exactly sum(phases) times."""
    with open(os.path.join(trace_dir, "run_kernel_trace.csv")) as fh:
        rows = list(csv.DictReader(fh))
    calls = collections.defaultdict(list)
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        calls[signature_name(r["Kernel_Name"])].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    w, t, e = phases
    lines = ["", "Per-call durations from the same run's kernel trace (`run_kernel_trace.csv`), split in",
             f"dispatch order into the {w} warm-up assemblies, the {t} timed ones and the {e} that `bench.py`",
             "brackets with HIP events for its per-kernel times (`kernels_us`, `roofline.kernel_us`). The",
             "first table's averages run over all calls, including any change of a kernel's duration over",
             "the first calls after start-up; the last column lists the final five calls.", "",
             "| Kernel | Calls | warm-up avg (us) | timed avg (us) | event-bracketed avg (us) | last 5 (us) |",
             "|---|---|---|---|---|---|"]
    for name, d in calls.items():
        if len(d) != w + t + e:
            continue
        avg = lambda xs: sum(xs) / len(xs) if xs else float("nan")
        last = ", ".join(f"{x:.0f}" for x in d[-5:])
        lines.append(f"| `{name}` | {len(d)} | {avg(d[:w]):.1f} | {avg(d[w:w + t]):.1f} | "
                     f"{avg(d[w + t:]):.1f} | {last} |")
    return "\n".join(lines) + "\n"


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) < 3 or argv[0] not in ("stats", "traffic"):
        print(__doc__)
        return 2
    cmd, d, workload = argv[:3]
    ronly = "--ronly" in argv
    suffix = ("_fp32" if "--fp32" in argv else "") + ("_ronly" if ronly else "")
    if cmd == "traffic":
        out = traffic(d, workload, ronly, "--fp32" in argv)
        fn = os.path.join(PROFILES, f"traffic_{workload}{suffix}.json")
        with open(fn, "w") as fh:
            json.dump(out, fh, indent=1)
        print(json.dumps(out, indent=1))
        return 0
    command = ""
    if "--command" in argv:
        command = argv[argv.index("--command") + 1]
    text = stats_md(d, workload, command)
    if "--phases" in argv:
        text += phases_md(d, tuple(int(x) for x in argv[argv.index("--phases") + 1].split(",")))
    fn = os.path.join(PROFILES, f"kernel_stats_{workload}{suffix}.md")
    with open(fn, "w") as fh:
        fh.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
