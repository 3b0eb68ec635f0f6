"""Serving (uasource_amd.serve): continuous batching of R and R+K requests over several
resident meshes. Every request completes; each result is bit for bit the result of the same
request served alone (the kernels are deterministic and requests on one mesh never overlap);
in-flight outputs never exceed the HBM budget; with room for several, requests overlap."""
import numpy as np
import pytest

from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import batch_workloads, build_mesh, make_fields
from uasource_amd.serve import UvhServer, make_requests, percentiles

pytestmark = pytest.mark.gpu

_meshes = None


def meshes():
    global _meshes
    if _meshes is None:
        ws = batch_workloads(n=6, seed=21, lo=800, hi=12000, types=((3, 2), (6, 5), (10, 8)), weights=(0.5, 0.35, 0.15),
                             name="serve", C_sigma=0.2, A_sigma=0.0)
        _meshes = []
        for w in ws:
            M = build_mesh(w)
            F0, F1, _ = make_fields(M, w)
            _meshes.append((M, F0, F1))
    return _meshes


@pytest.mark.parametrize("rate", [float("inf"), 2000.0])
def test_requests_in_the_mix_equal_requests_alone(rate):
    print(f"[gpu] serve 60 requests rate={rate}", flush=True)
    srv = UvhServer(default_ctrlvar(), meshes(), slots=6, hbm_budget=8e9)
    try:
        reqs = make_requests(meshes(), 60, rate, seed=3, pinned=not np.isfinite(rate))
        srv.run(reqs)
        assert all(np.isfinite(r.done) and r.done >= r.start >= r.arrival for r in reqs)
        for r in reqs:
            assert np.all(np.isfinite(r.R)), f"request {r.rid}: non-finite R"
            alone = srv.serve_alone(r)
            assert np.array_equal(alone.R, r.R), f"request {r.rid}: R in the mix differs from alone"
            if r.newton:
                assert np.array_equal(alone.K, r.K), f"request {r.rid}: K in the mix differs from alone"
            else:
                assert r.K is None
        if not np.isfinite(rate):
            assert srv.peak_inflight > 1, "requests that fit together did not overlap"
        print("   " + str(percentiles(reqs)) + f", peak in flight {srv.peak_inflight}", flush=True)
    finally:
        srv.close()


def test_admission_respects_the_hbm_budget():
    print("[gpu] serve under a tight HBM budget", flush=True)
    ms = meshes()
    srv = UvhServer(default_ctrlvar(), ms, slots=6, hbm_budget=float("inf"))
    try:
        need = max(m.out_bytes(True) for m in srv.meshes)
        srv.budget = 1.5 * need      # room for about one Newton step at a time
        reqs = make_requests(ms, 40, float("inf"), seed=4, newton_fraction=0.5)
        srv.run(reqs)
        assert srv.peak_used <= srv.budget
        assert all(np.isfinite(r.done) for r in reqs)
        srv.budget = 1.0
        with pytest.raises(ValueError):
            srv.run(make_requests(ms, 2, float("inf"), seed=5, newton_fraction=1.0))
    finally:
        srv.close()


def test_a_request_with_a_nan_reports_the_error_and_the_others_are_unaffected():
    """A request whose state holds a NaN thickness completes with the reference's error
    (uvhMatrixAssembly:NaNinR, or ' NaN in K ' for a Newton step: uvhMatrixAssembly.m:575-584),
    read from its own status word; every other request in the mix has no error and equals
    its alone result bit for bit."""
    print("[gpu] serve with one NaN request", flush=True)
    ms = meshes()
    srv = UvhServer(default_ctrlvar(), ms, slots=6, hbm_budget=8e9)
    try:
        reqs = make_requests(ms, 30, float("inf"), seed=7, newton_fraction=0.5)
        bad = {5, 17}
        for k in bad:
            reqs[k].h = reqs[k].h.copy()
            reqs[k].h[3] = np.nan
        srv.run(reqs)
        for r in reqs:
            if r.rid in bad:
                assert r.error is not None, f"request {r.rid}: the NaN went unreported"
                want = "" if r.newton else "uvhMatrixAssembly:NaNinR"
                assert r.error.identifier in (want, "uvhMatrixAssembly:NaNinR"), r.error
            else:
                assert r.error is None, f"request {r.rid}: unexpected {r.error}"
                alone = srv.serve_alone(r)
                assert alone.error is None and np.array_equal(alone.R, r.R)
                if r.newton:
                    assert np.array_equal(alone.K, r.K)
    finally:
        srv.close()
