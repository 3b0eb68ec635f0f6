"""uv (diagnostic SSTREAM) assembly on the GPU (uvh_uv_*) against the uv oracle
(oracle o2.assemble_uv, UaMain/uvMatrixAssemblySSTREAM.m), to the bars of the uvh assembly:
pattern bit-exact; R within 1e-12 of max(||Tint||, ||Fext||); K within 1e-12 normwise (the
oracle's K is the reference's (K + K')/2; the library's is symmetric to rounding, checked to
1e-14); R-only gives the R of the R+K call bit for bit; runs repeat bit for bit; the
precision form (FP32 element arithmetic, FP64 accumulation) within 1e-6 (R) and 1e-5 (K); the
Cmin/AGlenmin floors and Picard (no viscosity-derivative terms) as in the oracle; the
ZeroFields flag is refused by the C ABI, and CtrlVar.uvAssembly.ZeroFields through the Python
mirror (which zeroes the velocities but for the fixed values of BCs) matches the oracle."""
import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

pytestmark = pytest.mark.gpu

_cases = {}


def case(o2lib, name):
    if name not in _cases:
        w = WORKLOADS[name]
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        cv = default_ctrlvar()
        ref = o2lib.assemble(cv, M, F0, F1)       # the floatation pre-pass leaves b in mesh.b
        F = F1.copy()
        F.s = ref["mesh"].b + F1.h
        _cases[name] = (M, F, ref["mesh"])
    return _cases[name]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p2", "tiny_p3", "medium_p2", "medium_p3", "plain_p2"])
def test_uv_matches_oracle(o2lib, name):
    print(f"[gpu] uv assembly {name}", flush=True)
    M, F, mesh = case(o2lib, name)
    cv = default_ctrlvar()
    ref = o2lib.assemble_uv(cv, M, F, mesh=mesh)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        rowptr, colind = ctx.uv_pattern()
        rp, ci = o2lib.uv_csr_pattern(mesh)
        assert np.array_equal(rowptr, rp) and np.array_equal(colind, ci)
        out = ctx.assemble_uv(cv, F, want_tf=True)
        scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
        errR = np.linalg.norm(out["R"] - ref["R"]) / scale
        errK = rel(out["Kval"], ref["Kval"])
        K = ctx.uv_K_matrix(out["Kval"])
        asym = abs(K - K.T).sum() / abs(K).sum()
        print(f"   R {errR:.2e}, K {errK:.2e}, asymmetry {asym:.1e}", flush=True)
        assert errR <= 1e-12 and errK <= 1e-12 and asym <= 1e-14
        assert rel(out["Tint"], ref["Tint"]) <= 1e-12
        # R only: the same R; a second run: the same bits
        r_only = ctx.assemble_uv(cv, F, ronly=True)
        assert np.array_equal(r_only["R"], out["R"])
        again = ctx.assemble_uv(cv, F)
        assert np.array_equal(again["R"], out["R"]) and np.array_equal(again["Kval"], out["Kval"])
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["medium_p2", "medium_p3"])
def test_uv_precision_form(o2lib, name):
    print(f"[gpu] uv precision form {name}", flush=True)
    M, F, mesh = case(o2lib, name)
    cv = default_ctrlvar()
    ref = o2lib.assemble_uv(cv, M, F, mesh=mesh)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        out = ctx.assemble_uv(cv, F, fp32=True)
        scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
        errR = np.linalg.norm(out["R"] - ref["R"]) / scale
        errK = rel(out["Kval"], ref["Kval"])
        print(f"   R {errR:.2e}, K {errK:.2e}", flush=True)
        assert errR <= 1e-6 and errK <= 1e-5
    finally:
        ctx.close()


def test_uv_floors_picard_and_refusals(o2lib):
    print("[gpu] uv floors, Picard, ZeroFields", flush=True)
    M, F, mesh = case(o2lib, "tiny_p2")
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        c = default_ctrlvar()
        c.Cmin, c.AGlenmin = 2e-2, 3e-9
        G = F.copy()
        G.C = np.where(np.arange(M.Nnodes) % 2 == 0, -1.0, F.C)
        G.AGlen = np.where(np.arange(M.Nnodes) % 3 == 0, 0.0, F.AGlen)
        for cv in (c, default_ctrlvar()):
            cv.Picard = cv is not c
            ref = o2lib.assemble_uv(cv, M, G, mesh=mesh)
            out = ctx.assemble_uv(cv, G, nan_check=False)
            scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
            assert np.linalg.norm(out["R"] - ref["R"]) / scale <= 1e-12
            assert rel(out["Kval"], ref["Kval"]) <= 1e-12
        # ZeroFields is the caller's business for the uv system
        ctrl = uvh.make_ctrl(default_ctrlvar(), F)
        keep = []
        fields = uvh.make_uv_fields(F, keep)
        R = np.zeros(2 * M.Nnodes)
        st = ctx.lib.uvh_uv_assemble(ctx.h, uvh.C.byref(ctrl), uvh.C.byref(uvh.make_uv_ctrl(default_ctrlvar())),
                                     uvh.C.byref(fields), uvh.UVH_ZEROFIELDS | uvh.UVH_RONLY, uvh._ptr(R), None,
                                     None, None)
        assert st == uvh.UVH_E_UNSUPPORTED
    finally:
        ctx.close()


def test_uv_zero_fields_through_the_mirror(o2lib):
    print("[gpu] uv ZeroFields", flush=True)
    from uasource_amd.fields import BoundaryConditions
    M, F, mesh = case(o2lib, "medium_p2")
    x = M.coordinates[:, 0]
    inflow = np.flatnonzero(x <= x.min() + 1.0)
    BCs = BoundaryConditions(ubFixedNode=inflow, ubFixedValue=np.full(inflow.size, 100.0),
                             vbFixedNode=inflow, vbFixedValue=np.zeros(inflow.size))
    cv = default_ctrlvar()
    cv.uvAssembly.ZeroFields = True
    ref = o2lib.assemble_uv(cv, M, F, mesh=mesh, BCs=BCs)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        out = ctx.assemble_uv(cv, F, want_tf=True, BCs=BCs)
        scale = max(np.linalg.norm(ref["Tint"]), np.linalg.norm(ref["Fext"]))
        assert np.linalg.norm(out["R"] - ref["R"]) / scale <= 1e-12
        assert rel(out["Kval"], ref["Kval"]) <= 1e-12
    finally:
        ctx.close()
