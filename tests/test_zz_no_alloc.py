"""The allocation-free entry point (uvh_assemble_device_noalloc), through the C ABI: its work
buffers live in the library's file-scope scratch array, so a call allocates no device memory
(the device's free memory is the same before and after); its R and K equal those of
uvh_assemble_device bit for bit (the same kernels in the same order), for R+K, R only and the
precision form, on two calls in a row of which the second has the smaller mesh (the array
keeps the first call's contents); NaN status reaches uvh_sync_status; a mesh whose work
buffers exceed the array is refused with UVH_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

pytestmark = pytest.mark.gpu


def bind(lib):
    f = lib.uvh_assemble_device_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    f.restype = C.c_int
    return f


class Dev:
    """Device copies of F0, F1 and output buffers for one mesh."""

    def __init__(self, torch, M, F0, F1, nnz):
        dev = torch.device("cuda:0")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)  # noqa: E731
        ns1, ns0 = type("F", (), {})(), type("F", (), {})()
        self.keep = []
        for k in uvh.F1_KEYS:
            setattr(ns1, k, t(getattr(F1, k)))
            self.keep.append(getattr(ns1, k))
        for k in uvh.F0_KEYS:
            setattr(ns0, k, t(getattr(F0, k)))
            self.keep.append(getattr(ns0, k))
        self.ns1 = ns1
        self.fields = uvh.make_fields(ns0, ns1)
        self.R = torch.zeros(3 * M.Nnodes, dtype=torch.float64, device=dev)
        self.K = torch.zeros(nnz, dtype=torch.float64, device=dev)
        self.T = torch.zeros_like(self.R)
        self.F = torch.zeros_like(self.R)


def run(ctx, noalloc, ctrl, d, flags, stream):
    ronly = bool(flags & uvh.UVH_RONLY)
    args = (ctx.h, C.byref(ctrl), C.byref(d.fields), flags, d.R.data_ptr(), None if ronly else d.K.data_ptr(),
            d.T.data_ptr(), d.F.data_ptr(), stream)
    st = noalloc(*args) if noalloc is not None else ctx.lib.uvh_assemble_device(*args)
    return st


def test_noalloc_matches_the_allocating_entry_point_and_allocates_nothing():
    import torch
    print("[gpu] allocation-free entry point", flush=True)
    lib = uvh.load_library()
    noalloc = bind(lib)
    cv = default_ctrlvar()
    stream = torch.cuda.current_stream().cuda_stream
    cases = []
    for name in ("medium_p2", "tiny_p2"):   # the second call has the smaller mesh
        w = WORKLOADS[name]
        M = build_mesh(w)
        F0, F1, _ = make_fields(M, w)
        cases.append((name, M, F0, F1))
    ctxs = []
    try:
        refs = {}
        for name, M, F0, F1 in cases:
            ctx = uvh.UvhContext(0)
            ctxs.append(ctx)
            ctx.set_mesh(M)
            d = Dev(torch, M, F0, F1, ctx.nnz)
            ctrl = uvh.make_ctrl(cv, F1)
            for flags in (0, uvh.UVH_RONLY, uvh.UVH_FP32):
                assert run(ctx, None, ctrl, d, flags, stream) == uvh.UVH_OK
                ctx.sync_status()
                refs[(name, flags)] = (d.R.clone(), d.K.clone(), d.T.clone(), d.F.clone())
            cases_d = (ctx, d, ctrl)
            refs[name] = cases_d
        # warm-up call (module and kernels loaded), then the free device memory must not move
        for name, M, F0, F1 in cases:
            ctx, d, ctrl = refs[name]
            assert run(ctx, noalloc, ctrl, d, 0, stream) == uvh.UVH_OK
        torch.cuda.synchronize()
        for rep in range(2):
            for name, M, F0, F1 in cases:
                ctx, d, ctrl = refs[name]
                for flags in (0, uvh.UVH_RONLY, uvh.UVH_FP32):
                    for buf in (d.R, d.K, d.T, d.F):
                        buf.fill_(-7.0)
                    torch.cuda.synchronize()
                    # only the call itself between the two readings (torch's own work may allocate)
                    free0 = torch.cuda.mem_get_info(0)[0]
                    assert run(ctx, noalloc, ctrl, d, flags, stream) == uvh.UVH_OK, ctx.lib.uvh_last_error(ctx.h)
                    ctx.sync_status()
                    free1 = torch.cuda.mem_get_info(0)[0]
                    assert free1 == free0, f"the allocation-free entry point allocated {free0 - free1} bytes"
                    R, K, T, F = refs[(name, flags)]
                    assert torch.equal(d.R, R) and torch.equal(d.T, T) and torch.equal(d.F, F), (name, flags, rep)
                    if not flags & uvh.UVH_RONLY:
                        assert torch.equal(d.K, K), (name, flags, rep)
        # a NaN in h reaches uvh_sync_status through the scratch status word
        ctx, d, ctrl = refs["tiny_p2"]
        h = d.ns1.h
        saved = h[3].item()
        h[3] = float("nan")
        assert run(ctx, noalloc, ctrl, d, 0, stream) == uvh.UVH_OK
        with pytest.raises(uvh.UaError):
            ctx.sync_status()
        h[3] = saved
        assert run(ctx, noalloc, ctrl, d, 0, stream) == uvh.UVH_OK
        ctx.sync_status()
    finally:
        for c in ctxs:
            c.close()


def test_noalloc_refuses_a_mesh_beyond_its_scratch():
    import torch
    print("[gpu] allocation-free entry point: size limit", flush=True)
    lib = uvh.load_library()
    noalloc = bind(lib)
    w = WORKLOADS["plain_p2"]   # 199,712 6-node elements: ~0.57 GB of element rows
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    ctx = uvh.UvhContext(0)
    try:
        ctx.set_mesh(M)
        d = Dev(torch, M, F0, F1, ctx.nnz)
        ctrl = uvh.make_ctrl(default_ctrlvar(), F1)
        st = run(ctx, noalloc, ctrl, d, 0, torch.cuda.current_stream().cuda_stream)
        assert st == uvh.UVH_E_ARG
        assert "allocation-free entry point holds at most" in ctx.lib.uvh_last_error(ctx.h).decode()
    finally:
        ctx.close()
