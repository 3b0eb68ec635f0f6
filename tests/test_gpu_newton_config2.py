"""The config-2 Newton solve at its stated size on the GPU: 199,712 6-node elements with 7
points, started as an Úa transient run starts (UaMain/Ua2D.m:598-611: a uv solve on F1's
geometry, whose velocities become F0's and F1's), then the implicit uvh solve
(SSTREAM_TransientImplicit.m) to Úa's default tolerance rForce < 1e-15, with the GPU assembly
and the device block-ILU(0) Krylov solve for both systems. The iteration counts must be
those of the oracle-driven run (O2 assembly, SuperLU steps) recorded in
tests/golden/newton_plain_p2_history.json within one, and the rForce history must follow it.
The precision form (FP32 element arithmetic, FP64 accumulation; FP64 residual and Jacobian from
rForce 1e-8 on) from the same start converges to the same tolerance, its u, v within 1e-9 of
the FP64 solution's speed scale and its h within 1e-9 of the FP64 h (normwise, relative)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_state = {}


def start_and_fp64():
    from measure import bench_modes
    if not _state:
        log = lambda m: print(m, flush=True)  # noqa: E731
        st = bench_modes.transient_start_state("plain_p2", device_solver=True, log=log)
        res, _ = bench_modes.uvh_solve(st, fp32=False, device_solver=True, log=log)
        _state.update(start=st, fp64=res)
    return _state["start"], _state["fp64"]


def test_config2_transient_newton_matches_the_oracle_driven_history():
    print("[gpu] config-2 Newton solve (transient start), 199,712 P2 elements", flush=True)
    ref = json.load(open(os.path.join(HERE, "golden", "newton_plain_p2_history.json")))
    st, res = start_and_fp64()
    for sysname, g in (("uv", st["uv"]), ("uvh", res)):
        o = ref[sysname]
        print(f"   {sysname}: GPU {g.iterations} iterations rForce {g.rForce:.3e}; "
              f"oracle {o['iterations']} rForce {o['rForce']:.3e}", flush=True)
        assert g.converged and g.rForce < 1e-15
        assert abs(g.iterations - o["iterations"]) <= 1
        # the histories follow each other: rForce at each common iteration within a factor 10
        # once both are below 1e-3 (the linear solves differ: direct against rtol 1e-8)
        for hg, ho in zip(g.history, o["history"]):
            if hg[2] < 1e-3 and ho[2] < 1e-3 and ho[2] > 1e-14:
                assert 0.1 < hg[2] / ho[2] < 10.0, (sysname, hg, ho)


def test_config2_precision_form_newton_meets_the_fp64_solution():
    from measure import bench_modes
    print("[gpu] config-2 Newton solve, precision form", flush=True)
    st, res64 = start_and_fp64()
    r32, _ = bench_modes.uvh_solve(st, fp32=True, device_solver=True, log=lambda m: print(m, flush=True))
    err = bench_modes.solution_errors(r32, res64)
    print(f"   {r32.iterations} iterations ({res64.iterations} FP64), rForce {r32.rForce:.3e}; u, v, h errors "
          + ", ".join(f"{k} {v['normwise']:.1e} (max {v['max']:.1e})" for k, v in err.items()), flush=True)
    assert r32.converged and r32.rForce < 1e-15
    assert all(v["normwise"] < 1e-9 for v in err.values())
