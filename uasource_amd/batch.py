"""Batched uvh assembly of many independent meshes on one GPU (``uvhb_*`` in include/uvh.h),
and the tile autotuner.

A batch holds ensemble members (each its own mesh, fields and outputs) and assembles all of
them with, per element type, one element launch and one row launch over a flattened
(member, block) work list: one workgroup per item, or persistent workgroups that take items
from an atomic counter when ``UVH_BATCH=persistent`` is set at creation. Each member's R and K are those of a single-mesh ``uvhAssembly``
(``UaMain/uvhAssembly.m:1``) on that member to rounding, and bit for bit the same from run to
run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import uvh

EXPORTED_BATCH = ("uvhb_create", "uvhb_destroy", "uvhb_last_error", "uvhb_add_mesh", "uvhb_size", "uvhb_member",
                  "uvhb_set_tile", "uvhb_get_tile", "uvhb_assemble_device", "uvhb_sync_status", "uvhb_cost_device")

_bound = False


def _bind(lib):
    global _bound
    if _bound:
        return lib
    vp, i32 = C.c_void_p, C.c_int32
    lib.uvhb_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.uvhb_create.restype = C.c_int
    lib.uvhb_destroy.argtypes = [vp]
    lib.uvhb_destroy.restype = None
    lib.uvhb_last_error.argtypes = [vp]
    lib.uvhb_last_error.restype = C.c_char_p
    lib.uvhb_add_mesh.argtypes = [vp, i32, i32, i32, vp, vp, vp, i32, vp, vp, C.POINTER(i32)]
    lib.uvhb_add_mesh.restype = C.c_int
    lib.uvhb_size.argtypes = [vp]
    lib.uvhb_size.restype = C.c_int
    lib.uvhb_member.argtypes = [vp, i32]
    lib.uvhb_member.restype = vp
    lib.uvhb_set_tile.argtypes = [vp, i32, i32, i32]
    lib.uvhb_set_tile.restype = C.c_int
    lib.uvhb_get_tile.argtypes = [vp, i32, i32]
    lib.uvhb_get_tile.restype = C.c_int
    lib.uvhb_assemble_device.argtypes = [vp, C.POINTER(uvh.UvhCtrl), C.POINTER(uvh.UvhFields), C.c_uint32,
                                         C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), vp]
    lib.uvhb_assemble_device.restype = C.c_int
    lib.uvhb_sync_status.argtypes = [vp, C.POINTER(i32)]
    lib.uvhb_sync_status.restype = C.c_int
    lib.uvhb_cost_device.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), vp, vp,
                                     vp]
    lib.uvhb_cost_device.restype = C.c_int
    _bound = True
    return lib


class MemberView(uvh.UvhContext):
    """A batch member's context (owned by the batch): pattern, CSC map, single-mesh calls."""

    def __init__(self, lib, handle, MUA, device):  # noqa: super().__init__ would create a context
        self.lib = lib
        self.h = C.c_void_p(handle)
        self.device = device
        self.nn, self.nele, self.nod, self.nip = MUA.Nnodes, MUA.Nele, MUA.nod, MUA.nip
        neq, nnz = C.c_int64(), C.c_int64()
        self._check(self.lib.uvh_pattern_info(self.h, C.byref(neq), C.byref(nnz)), "uvh_pattern_info")
        self.neq, self.nnz = neq.value, nnz.value
        self._pattern = None
        self._csc_map = None
        self.mesh_key = uvh.mesh_key(MUA)

    def close(self):  # the batch owns the context
        self.h = None


class UvhBatch:
    def __init__(self, device: int = 0):
        self.lib = _bind(uvh.load_library())
        h = C.c_void_p()
        st = self.lib.uvhb_create(C.byref(h), int(device))
        if st != uvh.UVH_OK:
            raise uvh.UvhLibraryError(f"uvhb_create failed with status {st}")
        self.h, self.device = h, device
        self.members: list[MemberView] = []

    def close(self):
        if self.h:
            for m in self.members:
                m.close()
            self.lib.uvhb_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st == uvh.UVH_OK:
            return
        msg = self.lib.uvhb_last_error(self.h).decode()
        if st == uvh.UVH_E_NAN_R:
            raise uvh.UaError("uvhMatrixAssembly:NaNinR", msg)
        if st == uvh.UVH_E_NAN_K:
            raise uvh.UaError("", msg)  # no identifier (uvhMatrixAssembly.m:577)
        if st == uvh.UVH_E_UNSUPPORTED:
            raise uvh.UvhUnsupported("uvhAssembly:unsupported", msg)
        raise uvh.UvhLibraryError(f"{what}: status {st}: {msg}")

    def add_mesh(self, MUA) -> int:
        conn = np.ascontiguousarray(MUA.connectivity, dtype=np.int32)
        x = np.ascontiguousarray(MUA.coordinates[:, 0], dtype=np.float64)
        y = np.ascontiguousarray(MUA.coordinates[:, 1], dtype=np.float64)
        pts = np.ascontiguousarray(MUA.points, dtype=np.float64)
        wts = np.ascontiguousarray(MUA.weights, dtype=np.float64)
        idx = C.c_int32()
        st = self.lib.uvhb_add_mesh(self.h, MUA.Nele, MUA.nod, MUA.Nnodes, uvh._ptr(conn), uvh._ptr(x), uvh._ptr(y),
                                    MUA.nip, uvh._ptr(pts), uvh._ptr(wts), C.byref(idx))
        self._check(st, "uvhb_add_mesh")
        self.members.append(MemberView(self.lib, self.lib.uvhb_member(self.h, idx.value), MUA, self.device))
        return idx.value

    def set_tile(self, nod: int, fp32: bool, epb: int):
        self._check(self.lib.uvhb_set_tile(self.h, nod, int(fp32), epb), "uvhb_set_tile")

    def get_tile(self, nod: int, fp32: bool) -> int:
        return self.lib.uvhb_get_tile(self.h, nod, int(fp32))

    def assemble_device(self, ctrl, fields, flags, R, K=None, stream=None):
        """``fields``: list of UvhFields; ``R``/``K``: lists of device tensors (K may be None
        with UVH_RONLY). Enqueue only."""
        n = len(self.members)
        farr = (uvh.UvhFields * n)(*fields)
        rarr = (C.c_void_p * n)(*[uvh._ptr(r) for r in R])
        karr = None if K is None else (C.c_void_p * n)(*[uvh._ptr(k) for k in K])
        st = self.lib.uvhb_assemble_device(self.h, C.byref(ctrl), farr, int(flags), rarr, karr, stream)
        self._check(st, "uvhb_assemble_device")

    def cost_device(self, R, d, w, rr, d2, stream=None):
        """Per member: rr[m] = |w R|^2, d2[m] = -(w R) . d (device tensors; ``w`` may be None).
        Enqueue only; capturable once a first call has set up the tables."""
        n = len(self.members)
        ra = (C.c_void_p * n)(*[uvh._ptr(r) for r in R])
        da = (C.c_void_p * n)(*[uvh._ptr(x) for x in d])
        wa = None if w is None else (C.c_void_p * n)(*[uvh._ptr(x) for x in w])
        st = self.lib.uvhb_cost_device(self.h, ra, da, wa, uvh._ptr(rr), uvh._ptr(d2), stream)
        self._check(st, "uvhb_cost_device")

    def sync_status(self):
        n = len(self.members)
        arr = (C.c_int32 * n)()
        st = self.lib.uvhb_sync_status(self.h, arr)
        self._check(st, "uvhb_sync_status")
        return list(arr)


# ------------------------------------------------------------------ autotuner

def tile_candidates(nod: int, nip: int, fp32: bool) -> list[int]:
    """Elements per workgroup to try: the largest tile under LDS caps of 40, 64, 80 and
    156 KiB (4, 2, 2 and 1 workgroups per CU), and the sizes that fill the 256 threads of the
    row stage in whole rounds."""
    per_el = max((21 * nod + 1) * 8, (nip * 36 + 1) * (4 if fp32 else 8))
    tables = (3 * nip * nod + nip + 1) * 8 + (3 * nip * nod * 4 + 16 if fp32 else 0)
    cap = min(256 // nip, 256 // nod)
    out = set()
    for lds in (40 * 1024, 64 * 1024, 80 * 1024, 156 * 1024):
        e = min(cap, (lds - tables) // per_el)
        if e >= 1:
            out.add(int(e))
    # row-stage items per element: one per block row (3 nod); FP64 3- and 6-node elements take
    # the x and y block rows in one item (2 nod), FP32 ones all nine blocks in one (nod)
    items = {3 * nod}
    if nod <= 6:
        items.add(nod if fp32 else 2 * nod)
    for per in sorted(items):
        for rounds in (1, 2, 3, 4):
            e = (256 * rounds) // per
            if 1 <= e <= cap and tables + e * per_el <= 156 * 1024:
                out.add(int(e))
    return sorted(out)


def autotune(batch: UvhBatch, ctrl, fields, R, K, flags, stream, torch, reps: int = 3, verbose=False) -> dict:
    """Times the batched assembly for every tile candidate of every element type in the batch
    (the other types held at their current tile) and keeps the fastest. Returns
    {(nod, fp32): (epb, ms)}."""
    fp32 = bool(flags & uvh.UVH_FP32)
    types = sorted({(m.nod, m.nip) for m in batch.members})
    chosen = {}
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    for nod, nip in types:
        best = None
        for epb in tile_candidates(nod, nip, fp32):
            try:
                batch.set_tile(nod, fp32, epb)
            except uvh.UvhLibraryError:
                continue  # the library's exact LDS count rejects it
            batch.assemble_device(ctrl, fields, flags, R, K, stream=stream)   # warm-up (lists, LDS attributes)
            batch.sync_status()
            times = []
            for _ in range(reps):
                ev0.record(s)
                batch.assemble_device(ctrl, fields, flags, R, K, stream=stream)
                ev1.record(s)
                torch.cuda.synchronize()
                times.append(ev0.elapsed_time(ev1))
            t = sorted(times)[len(times) // 2]
            if verbose:
                print(f"  autotune nod={nod} fp32={fp32} epb={epb}: {t:.3f} ms", flush=True)
            if best is None or t < best[1]:
                best = (epb, t)
        batch.set_tile(nod, fp32, best[0])
        chosen[(nod, fp32)] = best
    return chosen
