"""Batched assembly (uvhb_*): one launch per stage and element type over many meshes of
mixed element types gives each member the R and K of a single-mesh assembly of that member
to rounding (1e-14 normwise; the batch kernels are compiled separately, so the compiler may
fuse a different product of an a*b + c*d into an FMA) and the oracle's to the 1e-12 bar;
the persistent-workgroup schedule gives the results of the default one to the same bounds;
repeated batched assemblies agree bit for bit; NaN status is reported per member; a captured
HIP graph of the batched assembly replays to the eager result; the autotuner's tile choice
changes no result."""
import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.batch import UvhBatch, autotune, tile_candidates
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import batch_workloads, build_mesh, make_fields

pytestmark = pytest.mark.gpu

_members = None


def members():
    global _members
    if _members is None:
        ws = batch_workloads(n=7, seed=9, lo=300, hi=6000, types=((3, 2), (6, 5), (10, 8)), weights=(0.4, 0.4, 0.2),
                             C_sigma=0.2)
        _members = []
        for w in ws:
            M = build_mesh(w)
            F0, F1, _ = make_fields(M, w)
            _members.append((M, F0, F1))
    return _members


def device_setup(batch, torch, dev):
    fields, R, K, keep = [], [], [], []
    for (M, F0, F1), mv in zip(members(), batch.members):
        class F:
            pass
        f1, f0 = F(), F()
        for k in uvh.F1_KEYS:
            t = torch.from_numpy(np.ascontiguousarray(getattr(F1, k))).to(dev)
            keep.append(t)
            setattr(f1, k, t)
        for k in uvh.F0_KEYS:
            t = torch.from_numpy(np.ascontiguousarray(getattr(F0, k))).to(dev)
            keep.append(t)
            setattr(f0, k, t)
        fields.append(uvh.make_fields(f0, f1))
        R.append(torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev))
        K.append(torch.zeros(mv.nnz, dtype=torch.float64, device=dev))
    return fields, R, K, keep


@pytest.fixture(scope="module")
def batch():
    import torch
    b = UvhBatch(0)
    for M, _, _ in members():
        b.add_mesh(M)
    yield b
    b.close()


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("ronly", [False, True])
def test_batch_members_match_single_mesh_assembly(batch, fp32, ronly):
    import torch
    dev = torch.device("cuda:0")
    print(f"[gpu] batch of {len(batch.members)} members fp32={fp32} ronly={ronly}", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    ctrl = uvh.make_ctrl(cv, members()[0][2])
    flags = (uvh.UVH_FP32 if fp32 else 0) | (uvh.UVH_RONLY if ronly else 0)
    stream = torch.cuda.current_stream().cuda_stream
    batch.assemble_device(ctrl, fields, flags, R, None if ronly else K, stream=stream)
    assert batch.sync_status() == [0] * len(batch.members)
    first = [(r.clone(), k.clone()) for r, k in zip(R, K)]
    batch.assemble_device(ctrl, fields, flags, R, None if ronly else K, stream=stream)
    batch.sync_status()
    for i, (r, k) in enumerate(first):   # bit for bit from run to run
        assert torch.equal(r, R[i]) and (ronly or torch.equal(k, K[i])), f"member {i}: not reproducible"
    ctx = uvh.UvhContext(0)
    try:
        for i, (M, F0, F1) in enumerate(members()):
            ctx.set_mesh(M)
            ref = ctx.assemble(cv, F0, F1, ronly=ronly, fp32=fp32, want_tf=True)
            scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
            tol = 1e-6 if fp32 else 1e-14   # FP32: different roundings of FP32 element arithmetic
            errR = np.linalg.norm(R[i].cpu().numpy() - ref["R"]) / scale
            assert errR <= tol, f"member {i}: R differs by {errR:.2e}"
            if not ronly:
                kb = K[i].cpu().numpy()
                errK = np.linalg.norm(kb - ref["Kval"]) / np.linalg.norm(ref["Kval"])
                assert errK <= tol, f"member {i}: K differs by {errK:.2e}"
    finally:
        ctx.close()


def test_batch_members_meet_the_oracle(batch, o2lib):
    import torch
    from test_gpu_parity import assert_parity
    dev = torch.device("cuda:0")
    print("[gpu] batch members against the oracle", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    batch.assemble_device(uvh.make_ctrl(cv, members()[0][2]), fields, 0, R, K,
                          stream=torch.cuda.current_stream().cuda_stream)
    batch.sync_status()
    for i, ((M, F0, F1), mv) in enumerate(zip(members(), batch.members)):
        ref = o2lib.assemble(cv, M, F0, F1)
        assert_parity(mv, M, {"R": R[i].cpu().numpy(), "Kval": K[i].cpu().numpy()}, ref)


def test_batch_nan_is_reported_per_member(batch):
    import torch
    dev = torch.device("cuda:0")
    print("[gpu] batch NaN status", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    # member 2's h gets a NaN: only member 2 reports it
    bad = 2
    hbad = torch.from_numpy(members()[bad][2].h.copy()).to(dev)
    hbad[3] = float("nan")
    keep.append(hbad)
    fields[bad].h = hbad.data_ptr()
    stream = torch.cuda.current_stream().cuda_stream
    batch.assemble_device(uvh.make_ctrl(cv, members()[0][2]), fields, 0, R, K, stream=stream)
    with pytest.raises(uvh.UaError):
        batch.sync_status()
    st = (uvh.C.c_int32 * len(batch.members))()
    batch.lib.uvhb_sync_status(batch.h, st)
    assert [i for i, v in enumerate(st) if v != 0] == [bad]


def test_batch_graph_replay_matches_eager(batch):
    import torch
    dev = torch.device("cuda:0")
    print("[gpu] batch graph replay", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    ctrl = uvh.make_ctrl(cv, members()[0][2])
    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        batch.assemble_device(ctrl, fields, 0, R, K, stream=s.cuda_stream)   # eager; uploads the member table
    s.synchronize()
    eager_R = [r.clone() for r in R]
    eager_K = [k.clone() for k in K]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        batch.assemble_device(ctrl, fields, 0, R, K, stream=s.cuda_stream)
    for rep in range(4):
        for r, k in zip(R, K):
            r.fill_(-1.0)
            k.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(len(R)):
            assert torch.equal(R[i], eager_R[i]) and torch.equal(K[i], eager_K[i]), f"replay {rep} member {i}"
        _ = (R[0] * 2).sum().item()   # eager work between replays


def test_autotuner_choice_changes_no_result(batch):
    import torch
    dev = torch.device("cuda:0")
    print("[gpu] batch autotune", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    ctrl = uvh.make_ctrl(cv, members()[0][2])
    stream = torch.cuda.current_stream().cuda_stream
    batch.assemble_device(ctrl, fields, 0, R, K, stream=stream)
    batch.sync_status()
    before = [k.clone() for k in K]
    chosen = autotune(batch, ctrl, fields, R, K, 0, stream, torch, reps=2)
    for (nod, fp32), (epb, ms) in chosen.items():
        nip = next(m.nip for m in batch.members if m.nod == nod)
        assert epb in tile_candidates(nod, nip, fp32)
        assert batch.get_tile(nod, fp32) == epb
    batch.assemble_device(ctrl, fields, 0, R, K, stream=stream)
    batch.sync_status()
    for a, b in zip(before, K):
        assert torch.equal(a, b)


@pytest.mark.parametrize("fp32", [False, True])
def test_persistent_schedule_matches_default(batch, fp32, monkeypatch):
    import torch
    dev = torch.device("cuda:0")
    print(f"[gpu] batch persistent schedule fp32={fp32}", flush=True)
    cv = default_ctrlvar()
    fields, R, K, keep = device_setup(batch, torch, dev)
    ctrl = uvh.make_ctrl(cv, members()[0][2])
    flags = uvh.UVH_FP32 if fp32 else 0
    stream = torch.cuda.current_stream().cuda_stream
    batch.assemble_device(ctrl, fields, flags, R, K, stream=stream)
    batch.sync_status()
    ref = [(r.clone(), k.clone()) for r, k in zip(R, K)]
    monkeypatch.setenv("UVH_BATCH", "persistent")
    pb = UvhBatch(0)
    try:
        for M, _, _ in members():
            pb.add_mesh(M)
        for nod in {m.nod for m in batch.members}:
            pb.set_tile(nod, fp32, batch.get_tile(nod, fp32))
        for rep in range(2):
            for r, k in zip(R, K):
                r.fill_(-1.0)
                k.fill_(-1.0)
            pb.assemble_device(ctrl, fields, flags, R, K, stream=stream)
            assert pb.sync_status() == [0] * len(pb.members)
            for i, (r0, k0) in enumerate(ref):
                errR = ((R[i] - r0).norm() / r0.norm()).item()
                errK = ((K[i] - k0).norm() / k0.norm()).item()
                tol = 1e-6 if fp32 else 1e-14   # FP32: roundings of separately compiled FP32 arithmetic
                assert errR <= tol and errK <= tol, f"member {i} rep {rep}: R {errR:.2e} K {errK:.2e}"
    finally:
        pb.close()
