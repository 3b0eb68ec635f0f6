"""The measured configurations at their stated sizes, on the GPU, through the C ABI.

* Precision form, 500,000 10-node elements with 16 points (``precision_p3``, a Gaussian bump in
  the bed so the grounding line crosses the domain): the FP64 path against the FP64 oracle O2
  to the plain bar (R 1e-12 of max(|Tint|, |Fext|), every K block 1e-12 normwise), and the
  FP32-element / FP64-accumulate form against the same oracle to its bars (R 1e-6, K blocks
  1e-5); the errors are printed per configuration.
* Batching, 64 meshes of 5k-80k mixed 3- and 6-node elements with perturbed C and AGlen: every
  member of the batched FP64 assembly equals its single-mesh assembly to 1e-14, a seeded sample
  of members meets the oracle to 1e-12, and the batched FP32 form stays within 1e-6 (R) and
  1e-5 (K) of the FP64 single-mesh results.
* Serving, 500 R / R+K requests (1:3) on 24 resident meshes of 2k-300k elements (3/6/10-node
  mix), arrivals faster than the device serves them: every request completes without error,
  and a seeded sample of 40 equals the same request served alone, bit for bit.
"""
import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, batch_workloads, build_mesh, make_fields, serve_workloads

from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def block_errors(Nn, nnzN, gpu_K, ref_K, colind):
    """Largest normwise relative error over the nine Nn x Nn blocks of K (CSR values of the
    library's layout: block row bi occupies [bi 3 nnzN, (bi + 1) 3 nnzN)), in chunks."""
    num = np.zeros(9)
    den = np.zeros(9)
    n = gpu_K.size
    step = 1 << 25
    for a in range(0, n, step):
        b = min(n, a + step)
        bi = np.arange(a, b, dtype=np.int64) // (3 * nnzN)
        bj = colind[a:b].astype(np.int64) // Nn
        k = 3 * bi + bj
        d = gpu_K[a:b] - ref_K[a:b]
        num += np.bincount(k, weights=d * d, minlength=9)
        den += np.bincount(k, weights=ref_K[a:b] * ref_K[a:b], minlength=9)
    return float(np.max(np.sqrt(num[den > 0] / den[den > 0])))


def device_members(mem, batch, torch, dev):
    """Device copies of every member's fields and output buffers."""
    fields, R, K, keep = [], [], [], []
    for (M, F0, F1), mv in zip(mem, batch.members):
        f = {}
        for src, keys in ((F1, uvh.F1_KEYS), (F0, uvh.F0_KEYS)):
            ns = type("F", (), {})()
            for k in keys:
                t = torch.from_numpy(np.ascontiguousarray(getattr(src, k))).to(dev)
                keep.append(t)
                setattr(ns, k, t)
            f[id(src)] = ns
        fields.append(uvh.make_fields(f[id(F0)], f[id(F1)]))
        R.append(torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev))
        K.append(torch.zeros(mv.nnz, dtype=torch.float64, device=dev))
    return fields, R, K, keep


def test_config3_precision_p3_full_size(o2lib):
    w = WORKLOADS["precision_p3"]
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    cv = default_ctrlvar()
    say(f"[gpu] config 3 precision_p3: Nele={M.Nele} Nn={M.Nnodes} nod={M.nod} nip={M.nip}")
    ref = o2lib.assemble(cv, M, F0, F1)
    assert ref["status"] == 0
    refK = ref["K"].data
    scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        rowptr, colind = ctx.pattern()
        assert np.array_equal(rowptr, ref["K"].indptr.astype(np.int64))
        assert np.array_equal(colind, ref["K"].indices.astype(np.int32))
        nnzN = ctx.nnz // 9
        errs = {}
        for fp32 in (False, True):
            gpu = ctx.assemble(cv, F0, F1, fp32=fp32)
            errs[fp32] = (np.linalg.norm(gpu["R"] - ref["R"]) / scale,
                          block_errors(M.Nnodes, nnzN, gpu["Kval"], refK, colind))
            del gpu
        say(f"   FP64 path vs oracle: R {errs[False][0]:.2e}, K (worst block) {errs[False][1]:.2e}")
        say(f"   FP32 form vs oracle: R {errs[True][0]:.2e}, K (worst block) {errs[True][1]:.2e}")
        assert errs[False][0] <= 1e-12 and errs[False][1] <= 1e-12
        assert errs[True][0] <= 1e-6 and errs[True][1] <= 1e-5
    finally:
        ctx.close()


def test_config4_batch64_full_size(o2lib):
    import torch
    from uasource_amd.batch import UvhBatch
    from test_gpu_parity import assert_parity
    ws = batch_workloads()   # 64 members, 5k-80k elements, P1/P2, C sigma 0.3, AGlen sigma 0.2
    mem = []
    for w in ws:
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        mem.append((M, F0, F1))
    say(f"[gpu] config 4 batch64: {len(mem)} members, {sum(m[0].Nele for m in mem)} elements")
    dev = torch.device("cuda:0")
    b = UvhBatch(0)
    try:
        for M, _, _ in mem:
            b.add_mesh(M)
        cv = default_ctrlvar()
        fields, R, K, keep = device_members(mem, b, torch, dev)
        ctrl = uvh.make_ctrl(cv, mem[0][2])
        stream = torch.cuda.current_stream().cuda_stream
        out = {}
        for fp32 in (False, True):
            b.assemble_device(ctrl, fields, uvh.UVH_FP32 if fp32 else 0, R, K, stream=stream)
            assert b.sync_status() == [0] * len(mem)
            out[fp32] = [(r.cpu().numpy(), k.cpu().numpy()) for r, k in zip(R, K)]
        ctx = uvh.UvhContext(0)
        worst = {False: 0.0, True: 0.0}
        try:
            for i, (M, F0, F1) in enumerate(mem):
                ctx.set_mesh(M)
                single = ctx.assemble(cv, F0, F1, want_tf=True)
                scale = max(np.linalg.norm(single["Tint"]), np.linalg.norm(single["Fext"]))
                kn = np.linalg.norm(single["Kval"])
                for fp32 in (False, True):
                    eR = np.linalg.norm(out[fp32][i][0] - single["R"]) / scale
                    eK = np.linalg.norm(out[fp32][i][1] - single["Kval"]) / kn
                    worst[fp32] = max(worst[fp32], eR, eK)
                    tol = (1e-6, 1e-5) if fp32 else (1e-14, 1e-14)
                    assert eR <= tol[0] and eK <= tol[1], f"member {i} fp32={fp32}: R {eR:.2e} K {eK:.2e}"
        finally:
            ctx.close()
        say(f"   batched vs single-mesh: FP64 {worst[False]:.2e}, FP32 form {worst[True]:.2e}")
        rng = np.random.default_rng(40)
        for i in sorted(rng.choice(len(mem), 4, replace=False)):
            M, F0, F1 = mem[i]
            ref = o2lib.assemble(cv, M, F0, F1)
            assert_parity(b.members[i], M, {"R": out[False][i][0], "Kval": out[False][i][1]}, ref)
            say(f"   member {i} (Nele {M.Nele}, nod {M.nod}) meets the oracle")
    finally:
        b.close()


def test_config5_serve500_full_size():
    from uasource_amd.serve import UvhServer, make_requests
    meshes = []
    for w in serve_workloads():
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        meshes.append((M, F0, F1))
    sizes = [m[0].Nele for m in meshes]
    say(f"[gpu] config 5 serve500: {len(meshes)} meshes of {min(sizes)}-{max(sizes)} elements")
    srv = UvhServer(default_ctrlvar(), meshes, slots=8, hbm_budget=32e9)
    try:
        reqs = make_requests(meshes, 500, 20000.0, seed=11)
        rng = np.random.default_rng(41)
        sample = set(rng.choice(len(reqs), 40, replace=False).tolist())
        for r in reqs:
            r.keep = r.rid in sample
        srv.run(reqs)
        assert all(np.isfinite(r.done) and r.done >= r.start >= r.arrival for r in reqs)
        assert all(r.error is None for r in reqs)
        assert srv.peak_inflight > 1
        for r in reqs:
            if r.rid not in sample:
                continue
            alone = srv.serve_alone(r)
            assert alone.error is None and np.array_equal(alone.R, r.R), f"request {r.rid}: R differs from alone"
            if r.newton:
                assert np.array_equal(alone.K, r.K), f"request {r.rid}: K differs from alone"
        lat = np.array([r.latency for r in reqs])
        say(f"   500 requests: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, p99 {np.percentile(lat, 99) * 1e3:.2f} ms, "
            f"peak in flight {srv.peak_inflight}; 40 sampled requests equal their alone results")
    finally:
        srv.close()
