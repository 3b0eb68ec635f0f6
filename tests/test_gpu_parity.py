"""GPU parity of the HIP assembly (through the C ABI) against the oracle.

Bar: the sparsity pattern is bit-exact; every K block and Tint, Fext agree with the
FP64 oracle to 1e-12 normwise (relative), and R to 1e-12 of max(|Tint|, |Fext|)
(R = Tint - Fext cancels); single K entries are held to 1e-10 of their block's
largest entry.
"""
import sys

import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

pytestmark = pytest.mark.gpu

TOL = 1e-12
# Entries are also checked one by one against their block's largest entry. Úa forms the
# element Jacobian from absolute coordinates (up to 4e5 m on the 400 km domains), which
# costs ~3 digits to cancellation in any implementation, so single entries of a block
# may differ by ~1e-12 of the block maximum even where the norms agree to 1e-13.
ENTRY_TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_mesh_cache = {}


def mesh_fields(name):
    if name not in _mesh_cache:
        w = WORKLOADS[name]
        M = build_mesh(w)
        F0, F1, BCs = make_fields(M, w)
        _mesh_cache[name] = (M, F0, F1, BCs)
    M, F0, F1, BCs = _mesh_cache[name]
    return M, F0.copy(), F1.copy(), BCs


@pytest.fixture(scope="module")
def ctx():
    c = uvh.UvhContext(0)
    yield c
    c.close()


def assert_parity(ctx, M, gpu, ref, ronly=False):
    Nn = M.Nnodes
    scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
    errR = np.linalg.norm(gpu["R"] - ref["R"]) / scale
    assert errR <= TOL, f"R rel err {errR:.3e}"
    if gpu.get("Tint") is not None:
        for k in ("Tint", "Fext"):
            for bi in range(3):
                a, b = gpu[k][bi * Nn:(bi + 1) * Nn], ref[k][bi * Nn:(bi + 1) * Nn]
                nb = np.linalg.norm(b)
                if nb > 0:
                    assert np.linalg.norm(a - b) <= TOL * nb, (k, bi, np.linalg.norm(a - b) / nb)
    if ronly:
        return
    rowptr, colind = ctx.pattern()
    Kref = ref["K"]
    assert np.array_equal(rowptr, Kref.indptr.astype(np.int64)), "CSR row pointers differ"
    assert np.array_equal(colind, Kref.indices.astype(np.int32)), "CSR column indices differ"
    Kg = gpu["Kval"]
    Kr = Kref.data
    rows = np.repeat(np.arange(3 * Nn), np.diff(rowptr)) // Nn
    cols = colind // Nn
    worst, worst_ent, fails = 0.0, 0.0, []
    for bi in range(3):
        for bj in range(3):
            sel = (rows == bi) & (cols == bj)
            a, b = Kg[sel], Kr[sel]
            nb = np.linalg.norm(b)
            if nb == 0:
                assert np.all(a == 0)
                continue
            rel = np.linalg.norm(a - b) / nb
            ent = np.max(np.abs(a - b)) / np.max(np.abs(b))
            worst, worst_ent = max(worst, rel), max(worst_ent, ent)
            if rel > TOL:
                fails.append(f"block ({bi},{bj}) normwise {rel:.3e}")
            if ent > ENTRY_TOL:
                fails.append(f"block ({bi},{bj}) entrywise {ent:.3e}")
    say(f"   R {errR:.2e}, worst K block normwise {worst:.2e}, entrywise {worst_ent:.2e}")
    assert not fails, "; ".join(fails)


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "cpu_reference", "medium_p2", "medium_p3"])
def test_parity_against_oracle(ctx, o2lib, name):
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    say(f"[gpu] parity {name}: Nele={M.Nele} nod={M.nod} nip={M.nip}")
    ctx.set_mesh(M)
    gpu = ctx.assemble(cv, F0, F1, want_tf=True)
    ref = o2lib.assemble(cv, M, F0, F1)
    assert_parity(ctx, M, gpu, ref)


@pytest.mark.parametrize("name", ["tiny_p2", "cpu_reference"])
def test_parity_against_matlab_order_oracle(ctx, name):
    from oracle import uvh_oracle as O1
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    say(f"[gpu] O1 parity {name}")
    ctx.set_mesh(M)
    gpu = ctx.assemble(cv, F0, F1, want_tf=True)
    ref = O1.uvh_matrix_assembly(cv, M, F0, F1, BCs)
    assert_parity(ctx, M, gpu, ref)


@pytest.mark.parametrize("name", ["tiny_p1", "medium_p2", "medium_p3"])
def test_ronly_and_zero_fields(ctx, o2lib, name):
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    ctx.set_mesh(M)
    for zf in (False, True):
        say(f"[gpu] ronly {name} zero_fields={zf}")
        gpu = ctx.assemble(cv, F0, F1, ronly=True, zero_fields=zf, want_tf=True)
        ref = o2lib.assemble(cv, M, F0, F1, ronly=True, zero_fields=zf)
        assert gpu["Kval"] is None
        assert_parity(ctx, M, gpu, ref, ronly=True)
        full = ctx.assemble(cv, F0, F1, zero_fields=zf)
        assert np.array_equal(full["R"], gpu["R"]), "R-only and R+K residuals differ"


def test_deterministic_bitwise(ctx):
    M, F0, F1, BCs = mesh_fields("medium_p2")
    cv = default_ctrlvar()
    ctx.set_mesh(M)
    say("[gpu] determinism medium_p2")
    a = ctx.assemble(cv, F0, F1, want_tf=True)
    b = ctx.assemble(cv, F0, F1, want_tf=True)
    for k in ("R", "Kval", "Tint", "Fext"):
        assert np.array_equal(a[k], b[k]), k


def test_nan_in_h_raises_reference_error(ctx):
    M, F0, F1, BCs = mesh_fields("tiny_p2")
    cv = default_ctrlvar()
    ctx.set_mesh(M)
    F1.h = F1.h.copy()
    F1.h[5] = np.nan
    say("[gpu] NaN error path")
    with pytest.raises(uvh.UaError) as ei:
        ctx.assemble(cv, F0, F1)
    assert ei.value.identifier in ("uvhMatrixAssembly:NaNinR", "")
    # the context stays usable
    F1.h[5] = 900.0
    ctx.assemble(cv, F0, F1)


def test_uvhAssembly_dropin_signature(o2lib):
    from uasource_amd.uvh import uvhAssembly, UaRunInfo
    M, F0, F1, BCs = mesh_fields("cpu_reference")
    cv = default_ctrlvar()
    RunInfo = UaRunInfo()
    say("[gpu] uvhAssembly drop-in")
    UserVar, RunInfo, R, K = uvhAssembly(None, RunInfo, cv, M, F0, F1, None, BCs)
    ref = o2lib.assemble(cv, M, F0, F1)
    scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
    assert np.linalg.norm(R - ref["R"]) <= TOL * scale
    d = (K - ref["K"])
    assert abs(d).max() <= TOL * abs(ref["K"]).max()
    assert RunInfo.CPU.Assembly.uvh > 0
    cv.uvhMatrixAssembly.Ronly = True
    _, _, R2, K2 = uvhAssembly(None, RunInfo, cv, M, F0, F1, None, BCs)
    assert K2 is None and np.array_equal(R2, R)


def test_device_path_matches_host_path(ctx):
    import torch
    M, F0, F1, BCs = mesh_fields("medium_p2")
    cv = default_ctrlvar()
    ctx.set_mesh(M)
    say("[gpu] device path")
    host = ctx.assemble(cv, F0, F1, want_tf=True)
    dev = torch.device("cuda:0")
    tens = {}
    for k in uvh.F1_KEYS:
        tens["1" + k] = torch.from_numpy(np.ascontiguousarray(getattr(F1, k))).to(dev)
    for k in uvh.F0_KEYS:
        tens["0" + k] = torch.from_numpy(np.ascontiguousarray(getattr(F0, k))).to(dev)

    class FS:
        pass
    f1, f0 = FS(), FS()
    for k in uvh.F1_KEYS:
        setattr(f1, k, tens["1" + k])
    for k in uvh.F0_KEYS:
        setattr(f0, k, tens["0" + k])
    fields = uvh.make_fields(f0, f1)
    ctrl = uvh.make_ctrl(cv, F1)
    R = torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev)
    K = torch.zeros(ctx.nnz, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ctx.assemble_device(ctrl, fields, 0, R, K, None, None, stream=stream)
    ctx.sync_status()
    assert np.array_equal(R.cpu().numpy(), host["R"])
    assert np.array_equal(K.cpu().numpy(), host["Kval"])


def test_full_size_plain_p2_parity(ctx, o2lib):
    """The measured workload itself (199,712 six-node elements, 7 points) against O2."""
    M, F0, F1, BCs = mesh_fields("plain_p2")
    cv = default_ctrlvar()
    say(f"[gpu] full-size parity plain_p2 Nele={M.Nele}")
    ctx.set_mesh(M)
    gpu = ctx.assemble(cv, F0, F1, want_tf=True)
    ref = o2lib.assemble(cv, M, F0, F1, nthreads=16)
    assert_parity(ctx, M, gpu, ref)


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2", "medium_p3", "plain_p2"])
def test_fused_path_matches_oracle_and_two_pass(o2lib, monkeypatch, name):
    """A/B check in the manner of the reference's CtrlVar.Parallel.isTest: the fused
    LDS-window path (UVH_ASSEMBLY=fused: one kernel, no element rows in HBM) meets the oracle
    like the default two-pass path, and the two agree to rounding (they sum the diagonal
    blocks and residual entries of vertices in different orders, every other entry exactly
    alike); each path's R alone equals its R of R+K bit for bit, and the fused path is
    bitwise reproducible."""
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M, F0, F1, nthreads=16)
    monkeypatch.setenv("UVH_ASSEMBLY", "fused")
    c1 = uvh.UvhContext(0)
    monkeypatch.delenv("UVH_ASSEMBLY")
    c2 = uvh.UvhContext(0)
    try:
        say(f"[gpu] fused and two-pass paths {name}")
        c2.set_mesh(M)
        c1.set_mesh(M)
        c1.profile(True)
        fused = c1.assemble(cv, F0, F1, want_tf=True)
        assert "fused_rk" in c1.kernel_times(), "UVH_ASSEMBLY=fused did not select the fused kernel"
        two = c2.assemble(cv, F0, F1, want_tf=True)
        assert_parity(c2, M, two, ref)
        assert_parity(c1, M, fused, ref)
        scale = max(np.linalg.norm(two["Tint"]), np.linalg.norm(two["Fext"]))
        assert np.linalg.norm(fused["R"] - two["R"]) <= 1e-14 * scale
        assert np.linalg.norm(fused["Kval"] - two["Kval"]) <= 1e-14 * np.linalg.norm(two["Kval"])
        for c, out in ((c1, fused), (c2, two)):
            r_only = c.assemble(cv, F0, F1, ronly=True)
            assert np.array_equal(r_only["R"], out["R"])
        again = c1.assemble(cv, F0, F1, want_tf=True)
        for k in ("R", "Kval", "Tint", "Fext"):
            assert np.array_equal(again[k], fused[k]), k
    finally:
        c1.close()
        c2.close()


@pytest.mark.parametrize("name", ["tiny_p1", "medium_p2", "medium_p3"])
def test_gather_scatter_matches_lds_scatter(monkeypatch, name):
    """The gather form of the row scatter (UVH_SCATTER=gather: one thread per node-graph
    entry, no LDS) sums in the same order as the default LDS scatter: bit-identical R and K."""
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    ref = uvh.UvhContext(0)
    monkeypatch.setenv("UVH_SCATTER", "gather")
    c2 = uvh.UvhContext(0)
    monkeypatch.delenv("UVH_SCATTER")
    try:
        say(f"[gpu] gather scatter {name}")
        ref.set_mesh(M)
        c2.set_mesh(M)
        for fp32 in (False, True):
            a = ref.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
            b = c2.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
            for k in ("R", "Kval", "Tint", "Fext"):
                assert np.array_equal(a[k], b[k]), (k, fp32)
    finally:
        ref.close()
        c2.close()


@pytest.mark.parametrize("name", ["tiny_p1", "medium_p2", "medium_p3"])
@pytest.mark.parametrize("threads", [64, 192])
def test_element_tile_threads_change_no_result(monkeypatch, name, threads):
    """Element tiles of one to three waves (UVH_ELEMENT_THREADS) compute every element row in
    the same thread-local order as the 256-thread default: bit-identical R and K."""
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    ref = uvh.UvhContext(0)
    monkeypatch.setenv("UVH_ELEMENT_THREADS", str(threads))
    c2 = uvh.UvhContext(0)
    monkeypatch.delenv("UVH_ELEMENT_THREADS")
    try:
        say(f"[gpu] element tile threads={threads} {name}")
        ref.set_mesh(M)
        c2.set_mesh(M)
        for fp32 in (False, True):
            for ronly in (False, True):
                a = ref.assemble(cv, F0, F1, want_tf=True, fp32=fp32, ronly=ronly)
                b = c2.assemble(cv, F0, F1, want_tf=True, fp32=fp32, ronly=ronly)
                for k in ("R", "Tint", "Fext") + (() if ronly else ("Kval",)):
                    assert np.array_equal(a[k], b[k]), (k, fp32, ronly)
    finally:
        ref.close()
        c2.close()


@pytest.mark.parametrize("name,chunks", [("medium_p2", 5), ("medium_p3", 3), ("plain_p2", 8)])
def test_pipelined_assembly_matches_single_pass(monkeypatch, name, chunks):
    """The pipelined assembly (UVH_CHUNKS: element tiles in chunks on the caller's stream, the
    node rows of each finished chunk on a second stream) runs the same kernels on the same
    data: bit-identical R and K, FP64 and FP32."""
    M, F0, F1, BCs = mesh_fields(name)
    cv = default_ctrlvar()
    monkeypatch.setenv("UVH_CHUNKS", "1")
    ref = uvh.UvhContext(0)
    monkeypatch.setenv("UVH_CHUNKS", str(chunks))
    c2 = uvh.UvhContext(0)
    monkeypatch.delenv("UVH_CHUNKS")
    try:
        say(f"[gpu] pipelined {name} chunks={chunks}")
        ref.set_mesh(M)
        c2.set_mesh(M)
        for fp32 in (False, True):
            a = ref.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
            b = c2.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
            for k in ("R", "Kval", "Tint", "Fext"):
                assert np.array_equal(a[k], b[k]), (k, fp32)
    finally:
        ref.close()
        c2.close()


@pytest.mark.parametrize("where", ["middle", "end"])
def test_node_in_no_element(o2lib, where):
    """A node that belongs to no element (mid-numbering, and last, i.e. last in its row block):
    empty K rows, zero R rows, no NaN status, parity with the oracle elsewhere; the same for
    the uv assembly."""
    from test_oracle_invariants import mesh_with_isolated_node
    print(f"[gpu] node in no element ({where})", flush=True)
    w = WORKLOADS["tiny_p2"]
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    pos = M.Nnodes // 2 if where == "middle" else M.Nnodes
    M2, G0, G1 = mesh_with_isolated_node(M, F0, F1, pos)
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M2, G0, G1)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M2)
        gpu = ctx.assemble(cv, G0, G1, want_tf=True)   # raises on a (false) NaN status
        assert_parity(ctx, M2, gpu, ref)
        n2 = M2.Nnodes
        assert all(gpu["R"][b * n2 + pos] == 0.0 for b in range(3))
        G = G1.copy()
        G.s = ref["mesh"].b + G1.h
        uv_ref = o2lib.assemble_uv(cv, M2, G, mesh=ref["mesh"])
        uv = ctx.assemble_uv(cv, G)
        scale = max(np.linalg.norm(uv_ref["Tint"]), np.linalg.norm(uv_ref["Fext"]))
        assert np.linalg.norm(uv["R"] - uv_ref["R"]) <= 1e-12 * scale
        assert np.linalg.norm(uv["Kval"] - uv_ref["Kval"]) <= 1e-12 * np.linalg.norm(uv_ref["Kval"])
    finally:
        ctx.close()


@pytest.mark.parametrize("tri", [3, 6, 10])
def test_single_element_mesh(o2lib, tri):
    """One element: a single tile, a single row block and a dense pattern."""
    from uasource_amd.mesh import CreateMUA
    print(f"[gpu] single {tri}-node element", flush=True)
    M = CreateMUA(np.array([[1000.0, 2000.0], [4000.0, 2500.0], [1500.0, 5000.0]]), np.array([[0, 1, 2]]),
                  TriNodes=tri)
    F0, F1, _ = make_fields(M, WORKLOADS["tiny_p2"])
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M, F0, F1)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        for fp32 in (False, True):
            gpu = ctx.assemble(cv, F0, F1, want_tf=True, fp32=fp32)
            if not fp32:
                assert_parity(ctx, M, gpu, ref)
            else:
                scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
                assert np.linalg.norm(gpu["R"] - ref["R"]) <= 1e-6 * scale
    finally:
        ctx.close()


@pytest.mark.parametrize("tri,degree,grid", [(3, 4, (61, 31)), (6, 8, (41, 21)), (10, 10, (31, 16))])
def test_reference_default_quadrature_many_tiles(o2lib, tri, degree, grid):
    """The reference's default rules (QuadratureRuleDegree.m:15,21,28: degree 4/8/10, 6/16/25
    points for 3/6/10-node elements) on meshes of many tiles (at 25 points a tile holds about
    ten 10-node elements), against the oracle to the FP64 bar, with R only and the
    precision form beside."""
    from uasource_amd.mesh import CreateMUA, jittered_delaunay
    xy, el = jittered_delaunay(100e3, 50e3, grid[0], grid[1], 0.1, seed=60 + tri)
    M = CreateMUA(xy, el, TriNodes=tri, QuadratureRuleDegree=degree)
    say(f"[gpu] default quadrature nod={tri} nip={M.nip} Nele={M.Nele}")
    F0, F1, _ = make_fields(M, WORKLOADS["medium_p2"])
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M, F0, F1)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        gpu = ctx.assemble(cv, F0, F1, want_tf=True)
        assert_parity(ctx, M, gpu, ref)
        r_only = ctx.assemble(cv, F0, F1, ronly=True)
        assert np.array_equal(r_only["R"], gpu["R"])
        lo = ctx.assemble(cv, F0, F1, fp32=True)
        scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
        assert np.linalg.norm(lo["R"] - ref["R"]) <= 1e-6 * scale
        assert np.linalg.norm(lo["Kval"] - ref["Kval"]) <= 1e-5 * np.linalg.norm(ref["Kval"])
    finally:
        ctx.close()
