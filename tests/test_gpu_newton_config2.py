"""The config-2 Newton solve at its stated size on the GPU: 199,712 6-node elements with 7
points, started as an Úa transient run starts (UaMain/Ua2D.m:598-611: a uv solve on F1's
geometry, whose velocities become F0's and F1's), then the implicit uvh solve
(SSTREAM_TransientImplicit.m) to Úa's default tolerance rForce < 1e-15, with the GPU assembly
and the device block-ILU(0) Krylov solve for both systems. The iteration counts must be
those of the oracle-driven run (O2 assembly, SuperLU steps) recorded in
tests/golden/newton_plain_p2_history.json within one, and the rForce history must follow it."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_config2_transient_newton_matches_the_oracle_driven_history():
    from measure import bench_modes
    print("[gpu] config-2 Newton solve (transient start), 199,712 P2 elements", flush=True)
    ref = json.load(open(os.path.join(HERE, "golden", "newton_plain_p2_history.json")))
    out = bench_modes.bench_newton_transient("plain_p2", device_solver=True, log=lambda m: print(m, flush=True))
    for sysname in ("uv", "uvh"):
        g, o = out[sysname], ref[sysname]
        print(f"   {sysname}: GPU {g['iterations']} iterations rForce {g['rForce']:.3e}; "
              f"oracle {o['iterations']} rForce {o['rForce']:.3e}", flush=True)
        assert g["converged"] and g["rForce"] < 1e-15
        assert abs(g["iterations"] - o["iterations"]) <= 1
        # the histories follow each other: rForce at each common iteration within a factor 10
        # once both are below 1e-3 (the linear solves differ: direct against rtol 1e-8)
        for hg, ho in zip(g["history"], o["history"]):
            if hg[2] < 1e-3 and ho[2] < 1e-3 and ho[2] > 1e-14:
                assert 0.1 < hg[2] / ho[2] < 10.0, (sysname, hg, ho)
