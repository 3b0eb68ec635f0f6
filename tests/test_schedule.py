"""The fused-assembly schedule (host side, no GPU): replaying oracle O2's element blocks
through the schedule's window slots, ranks, zero lists and flush lists -- exactly the
data flow of the fused kernel -- reproduces O2's assembled K and R, every owned row is
written once, window slots never overlap while live, and no slot is read before it is
zeroed."""
import ctypes as C
import os

import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

TABLES = ["rowptr_n", "colind_n", "pos", "inc_ptr", "inc", "patch_batch_ptr", "batch_elem", "ea_slot", "ea_rank",
          "batch_maxrank", "zero_ptr", "zero_slot", "zero_len", "flush_ptr", "flush_node", "flush_slot"]
DTYPES = {"rowptr_n": np.int64, "pos": np.uint8, "ea_rank": np.uint8, "batch_maxrank": np.uint8}


def host_lib():
    if not os.path.exists(uvh.LIB_PATH):
        pytest.skip("libuvh.so not built")
    lib = C.CDLL(uvh.LIB_PATH)
    lib.uvhh_build.restype = C.c_void_p
    lib.uvhh_build.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                               C.c_int64, C.c_int32, C.c_char_p, C.c_int32]
    lib.uvhh_size.restype = C.c_int64
    lib.uvhh_size.argtypes = [C.c_void_p, C.c_int32]
    lib.uvhh_copy.restype = None
    lib.uvhh_copy.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.uvhh_free.argtypes = [C.c_void_p]
    return lib


def build_schedule(M, epb, window, patches):
    lib = host_lib()
    conn = np.ascontiguousarray(M.connectivity, dtype=np.int32)
    x = np.ascontiguousarray(M.coordinates[:, 0])
    y = np.ascontiguousarray(M.coordinates[:, 1])
    err = C.create_string_buffer(256)
    h = lib.uvhh_build(M.Nele, M.nod, M.Nnodes, conn.ctypes.data, x.ctypes.data, y.ctypes.data, epb, window,
                       patches, err, 256)
    assert h, err.value.decode()
    out = {}
    for k, name in enumerate(TABLES):
        n = lib.uvhh_size(h, k)
        a = np.zeros(n, dtype=DTYPES.get(name, np.int32))
        if n:
            lib.uvhh_copy(h, k, a.ctypes.data)
        out[name] = a
    out["npatch"] = lib.uvhh_size(h, 16)
    out["peak_window"] = lib.uvhh_size(h, 17)
    out["patch_work"] = lib.uvhh_size(h, 18)
    lib.uvhh_free(h)
    return out


def replay(M, S, Kel, Rel, epb, window):
    """Kernel data flow on the CPU: returns (Kval, Tint, Fext) on the CSR pattern."""
    nod, Nn = M.nod, M.Nnodes
    rp, ci = S["rowptr_n"], S["colind_n"]
    nnzN = int(rp[-1])
    deg = np.diff(rp)
    Kval = np.full(9 * nnzN, np.nan)
    T = np.full(3 * Nn, np.nan)
    F = np.full(3 * Nn, np.nan)
    written = np.zeros(Nn, dtype=int)
    Kel = Kel.reshape(M.Nele, 3, 3, nod, nod)   # [e][bi][bj][i][j]
    Rel = Rel.reshape(M.Nele, nod, 6)           # Tx Fx Ty Fy Th Fh
    pos = S["pos"].reshape(M.Nele, nod, nod)
    for p in range(S["npatch"]):
        lds = np.full(window, np.nan)
        live = {}
        for b in range(S["patch_batch_ptr"][p], S["patch_batch_ptr"][p + 1]):
            for z in range(S["zero_ptr"][b], S["zero_ptr"][b + 1]):
                s0, ln = S["zero_slot"][z], S["zero_len"][z]
                assert 0 <= s0 and s0 + ln <= window
                for (a0, a1) in live.values():
                    assert s0 + ln <= a0 or a1 <= s0, "slot overlaps a live row"
                lds[s0:s0 + ln] = 0.0
                live[s0] = (s0, s0 + ln)
            els = S["batch_elem"][b * epb:(b + 1) * epb]
            slots = S["ea_slot"][b * epb * nod:(b + 1) * epb * nod].reshape(epb, nod)
            ranks = S["ea_rank"][b * epb * nod:(b + 1) * epb * nod].reshape(epb, nod)
            for r in range(int(S["batch_maxrank"][b]) + 1):
                for k in range(epb):
                    e = els[k]
                    if e < 0:
                        continue
                    for a in range(nod):
                        if slots[k, a] < 0 or ranks[k, a] != r:
                            continue
                        node = M.connectivity[e, a]
                        s = slots[k, a]
                        d = deg[node]
                        for bi in range(3):
                            for bj in range(3):
                                idx = s + bi * 3 * d + bj * d + pos[e, a, :]
                                lds[idx] += Kel[e, bi, bj, a, :]
                        lds[s + 9 * d + np.arange(3)] += Rel[e, a, [0, 2, 4]]
                        lds[s + 9 * d + 3 + np.arange(3)] += Rel[e, a, [1, 3, 5]]
            for f in range(S["flush_ptr"][b], S["flush_ptr"][b + 1]):
                node, s = S["flush_node"][f], S["flush_slot"][f]
                d = deg[node]
                written[node] += 1
                vals = lds[s:s + 9 * d]
                assert not np.any(np.isnan(vals)), "row read before it was zeroed"
                for bi in range(3):
                    start = bi * 3 * nnzN + 3 * rp[node]
                    Kval[start:start + 3 * d] = vals[bi * 3 * d:(bi + 1) * 3 * d]
                T[[node, Nn + node, 2 * Nn + node]] = lds[s + 9 * d:s + 9 * d + 3]
                F[[node, Nn + node, 2 * Nn + node]] = lds[s + 9 * d + 3:s + 9 * d + 6]
                del live[s]
        assert not live, "rows left in the window at the end of a patch"
    assert np.all(written == 1), "every node row is written exactly once"
    return Kval, T, F


@pytest.mark.parametrize("name,epb,window,patches", [
    ("tiny_p1", 8, 4000, 3),
    ("tiny_p2", 4, 3000, 4),
    ("tiny_p3", 2, 6000, 3),
    ("cpu_reference", 16, 6000, 8),
    ("medium_p2", 32, 12000, 16),
])
def test_schedule_replay_reproduces_assembly(name, epb, window, patches, o2lib):
    w = WORKLOADS[name]
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M, F0, F1)
    S = build_schedule(M, epb, window, patches)
    assert S["peak_window"] <= window
    assert S["patch_work"] >= M.Nele
    Kval, T, F = replay(M, S, ref["mesh"].Kel, ref["mesh"].Rel, epb, window)
    assert np.array_equal(S["rowptr_n"], ref["mesh"].rowptr_n)
    assert np.linalg.norm(Kval - ref["Kval"]) <= 1e-13 * np.linalg.norm(ref["Kval"])
    assert np.linalg.norm(T - ref["Tint"]) <= 1e-13 * np.linalg.norm(ref["Tint"])
    assert np.linalg.norm(F - ref["Fext"]) <= 1e-13 * np.linalg.norm(ref["Fext"])


def test_schedule_ranks_are_a_total_order_per_node():
    w = WORKLOADS["medium_p2"]
    M = build_mesh(w)
    epb = 32
    S = build_schedule(M, epb, 12000, 16)
    nod = M.nod
    for b in range(len(S["batch_maxrank"])):
        els = S["batch_elem"][b * epb:(b + 1) * epb]
        slots = S["ea_slot"][b * epb * nod:(b + 1) * epb * nod].reshape(epb, nod)
        ranks = S["ea_rank"][b * epb * nod:(b + 1) * epb * nod].reshape(epb, nod)
        seen = {}
        for k in range(epb):
            if els[k] < 0:
                continue
            for a in range(nod):
                if slots[k, a] >= 0:
                    node = int(M.connectivity[els[k], a])
                    seen.setdefault(node, []).append(int(ranks[k, a]))
        for node, rs in seen.items():
            assert sorted(rs) == list(range(len(rs))), (b, node, rs)
        assert max([max(v) for v in seen.values()] + [0]) == S["batch_maxrank"][b]


def test_full_size_schedule_fits_and_ghost_overhead_is_small():
    w = WORKLOADS["plain_p2"]
    M = build_mesh(w)
    S = build_schedule(M, 32, 12288, 256)
    assert S["peak_window"] <= 12288
    overhead = S["patch_work"] / M.Nele - 1
    print(f"plain_p2: {S['npatch']} patches, peak window {S['peak_window']} doubles, ghost overhead {overhead:.3f}")
    assert overhead < 0.3


@pytest.mark.parametrize("name", ["tiny_p1", "tiny_p3", "medium_p2"])
def test_csc_map_matches_scipy_transpose(name):
    """K in CSC (MATLAB's sparse layout) is the CSR values permuted by the map, on the same
    index arrays; checked against scipy's own CSR -> CSC conversion of random values."""
    import scipy.sparse as sp
    lib = host_lib()
    lib.uvhh_csc_map.restype = None
    lib.uvhh_csc_map.argtypes = [C.c_void_p, C.c_void_p]
    M = build_mesh(WORKLOADS[name])
    conn = np.ascontiguousarray(M.connectivity, dtype=np.int32)
    x = np.ascontiguousarray(M.coordinates[:, 0])
    y = np.ascontiguousarray(M.coordinates[:, 1])
    err = C.create_string_buffer(256)
    h = lib.uvhh_build(M.Nele, M.nod, M.Nnodes, conn.ctypes.data, x.ctypes.data, y.ctypes.data, 0, 0, 0, err, 256)
    assert h, err.value.decode()
    try:
        rp = np.zeros(lib.uvhh_size(h, 0), dtype=np.int64)
        lib.uvhh_copy(h, 0, rp.ctypes.data)
        ci = np.zeros(lib.uvhh_size(h, 1), dtype=np.int32)
        lib.uvhh_copy(h, 1, ci.ctypes.data)
        nnzN, Nn = int(rp[-1]), M.Nnodes
        # the full 3Nn x 3Nn pattern as uvh_get_pattern lays it out
        rowptr = np.zeros(3 * Nn + 1, dtype=np.int64)
        colind = np.zeros(9 * nnzN, dtype=np.int32)
        for bi in range(3):
            for k in range(Nn):
                start = bi * 3 * nnzN + 3 * rp[k]
                deg = rp[k + 1] - rp[k]
                rowptr[bi * Nn + k] = start
                for bj in range(3):
                    colind[start + bj * deg:start + (bj + 1) * deg] = ci[rp[k]:rp[k + 1]] + bj * Nn
        rowptr[-1] = 9 * nnzN
        cmap = np.zeros(9 * nnzN, dtype=np.int64)
        lib.uvhh_csc_map(h, cmap.ctypes.data)
    finally:
        lib.uvhh_free(h)
    vals = np.random.default_rng(5).standard_normal(9 * nnzN)
    K = sp.csr_matrix((vals, colind, rowptr), shape=(3 * Nn, 3 * Nn))
    Kc = K.tocsc()
    Kc.sort_indices()
    assert np.array_equal(Kc.indptr, rowptr) and np.array_equal(Kc.indices, colind)
    assert np.array_equal(Kc.data, vals[cmap])
    assert np.array_equal(np.sort(cmap), np.arange(9 * nnzN))  # a permutation
