"""The fused-assembly schedule (host side, no GPU): replaying oracle O2's element blocks
through the schedule -- the phases of the fused kernel, in its order: zero the rows that
start in a batch, add the previous batch's fixups, write out the rows that ended in the
previous batch, add the batch's row items -- reproduces O2's assembled K and R; every owned
row is written once; window slots never overlap while live; no slot is read before it is
zeroed; every value that takes more than one add in one phase takes at most two adds in all
(so the order of those adds cannot change the rounded sum), and the diagonal blocks and
residual entries of nodes with three or more elements take at most one add per phase."""
import ctypes as C
import os

import numpy as np
import pytest

from uasource_amd import uvh
from uasource_amd.ctrlvar import default_ctrlvar
from uasource_amd.fields import WORKLOADS, build_mesh, make_fields

MESH_TABLES = ["rowptr_n", "colind_n", "pos", "inc_ptr", "inc"]
SCHED_TABLES = ["patch_batch_ptr", "batch_elem", "item_ptr", "item_ea", "item_slot", "item_stage", "fix_ptr",
                "fix_node", "fix_slot", "fix_item_ptr", "fix_item", "zero_ptr", "zero_slot", "zero_len", "flush_ptr",
                "flush_node", "flush_slot"]
DTYPES = {"rowptr_n": np.int64, "pos": np.uint8}
STAGE_VALUES = 15


def host_lib():
    if not os.path.exists(uvh.LIB_PATH):
        pytest.skip("libuvh.so not built")
    lib = C.CDLL(uvh.LIB_PATH)
    lib.uvhh_build.restype = C.c_void_p
    lib.uvhh_build.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                               C.c_int64, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    lib.uvhh_size.restype = C.c_int64
    lib.uvhh_size.argtypes = [C.c_void_p, C.c_int32]
    lib.uvhh_copy.restype = None
    lib.uvhh_copy.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.uvhh_free.argtypes = [C.c_void_p]
    lib.uvhh_fused_layout.restype = None
    lib.uvhh_fused_layout.argtypes = [C.c_int32, C.c_int32, C.c_void_p]
    return lib


def build_schedule(M, epb, window, patch_elems=0, band_cells=0, expect_ok=True):
    lib = host_lib()
    conn = np.ascontiguousarray(M.connectivity, dtype=np.int32)
    x = np.ascontiguousarray(M.coordinates[:, 0])
    y = np.ascontiguousarray(M.coordinates[:, 1])
    err = C.create_string_buffer(256)
    h = lib.uvhh_build(M.Nele, M.nod, M.Nnodes, conn.ctypes.data, x.ctypes.data, y.ctypes.data, epb, window,
                       patch_elems, band_cells, err, 256)
    if not expect_ok:
        assert not h
        return err.value.decode()
    assert h, err.value.decode()
    out = {}
    for k, name in enumerate(MESH_TABLES + SCHED_TABLES):
        n = lib.uvhh_size(h, k)
        a = np.zeros(n, dtype=DTYPES.get(name, np.int32))
        if n:
            lib.uvhh_copy(h, k, a.ctypes.data)
        out[name] = a
    d = np.zeros(lib.uvhh_size(h, 36), dtype=np.uint8)
    lib.uvhh_copy(h, 36, d.ctypes.data)
    out["diag"] = d
    for k, name in zip(range(30, 36), ["npatch", "peak_window", "patch_work", "epb", "max_stage", "max_items"]):
        out[name] = lib.uvhh_size(h, k)
    lib.uvhh_free(h)
    return out


def fused_layout(nod, nip):
    lib = host_lib()
    o = np.zeros(3, dtype=np.int64)
    lib.uvhh_fused_layout(nod, nip, o.ctypes.data)
    return int(o[0]), int(o[1]), int(o[2])


class Window:
    """The LDS window of one workgroup, with per-value add counts per phase and in total."""

    def __init__(self, n):
        self.v = np.full(n, np.nan)
        self.total = np.zeros(n, dtype=int)      # adds since the slot was zeroed
        self.phase = np.zeros(n, dtype=int)      # adds in the current phase
        self.concurrent = np.zeros(n, dtype=bool)  # took two or more adds in one phase
        self.live = {}

    def zero(self, s0, ln):
        self.v[s0:s0 + ln] = 0.0
        self.total[s0:s0 + ln] = 0
        self.concurrent[s0:s0 + ln] = False

    def new_phase(self):
        self.concurrent |= self.phase > 1
        self.phase[:] = 0

    def add(self, idx, val):
        idx = np.atleast_1d(idx)
        assert not np.any(np.isnan(self.v[idx])), "a slot is added to before it was zeroed"
        np.add.at(self.v, idx, val)
        np.add.at(self.total, idx, 1)
        np.add.at(self.phase, idx, 1)


def replay(M, S, Kel, Rel, epb, window):
    """The fused kernel's data flow on the CPU: returns (Kval, Tint, Fext) on the CSR pattern."""
    nod, Nn = M.nod, M.Nnodes
    rp = S["rowptr_n"]
    nnzN = int(rp[-1])
    deg = np.diff(rp)
    ndeg_e = np.diff(S["inc_ptr"])
    Kval = np.full(9 * nnzN, np.nan)
    T = np.full(3 * Nn, np.nan)
    F = np.full(3 * Nn, np.nan)
    written = np.zeros(Nn, dtype=int)
    Kel = Kel.reshape(M.Nele, 3, 3, nod, nod)   # [e][bi][bj][i][j]
    Rel = Rel.reshape(M.Nele, nod, 6)           # Tx Fx Ty Fy Th Fh
    pos = S["pos"].reshape(M.Nele, nod, nod)

    def ordered_positions(node, s):
        d = deg[node]
        dc = S["diag"][node]
        return [s + (3 * bi + bj) * d + dc for bi in range(3) for bj in range(3)] + [s + 9 * d + k for k in range(6)]

    for p in range(S["npatch"]):
        W = Window(window)
        b0, b1 = S["patch_batch_ptr"][p], S["patch_batch_ptr"][p + 1]
        stage = None
        for b in range(b0, b1 + 1):
            # phase A: zero the rows that start in b; the fixups of b - 1
            W.new_phase()
            if b < b1:
                for z in range(S["zero_ptr"][b], S["zero_ptr"][b + 1]):
                    s0, ln = S["zero_slot"][z], S["zero_len"][z]
                    assert 0 <= s0 and s0 + ln <= window
                    for (a0, a1) in W.live.values():
                        assert s0 + ln <= a0 or a1 <= s0, "slot overlaps a live row"
                    W.zero(s0, ln)
                    W.live[s0] = (s0, s0 + ln)
            if b > b0:
                for g in range(S["fix_ptr"][b - 1], S["fix_ptr"][b]):
                    node, s = S["fix_node"][g], S["fix_slot"][g]
                    items = S["fix_item"][S["fix_item_ptr"][g]:S["fix_item_ptr"][g + 1]]
                    assert len(items) >= 2 and ndeg_e[node] >= 3
                    tot = stage[items].sum(axis=0)     # in item order, one value per slot
                    posn = ordered_positions(node, s)
                    # staging layout per block row bi: diag bj = 0..2, T, F
                    for bi in range(3):
                        for k in range(5):
                            v = bi * 5 + k
                            idx = posn[3 * bi + k] if k < 3 else posn[9 + (bi if k == 3 else 3 + bi)]
                            W.add(idx, tot[v])
            # phase B: write out the rows whose last element was in b - 1
            W.new_phase()
            if b > b0:
                for fidx in range(S["flush_ptr"][b - 1], S["flush_ptr"][b]):
                    node, s = S["flush_node"][fidx], S["flush_slot"][fidx]
                    d = deg[node]
                    written[node] += 1
                    vals = W.v[s:s + 9 * d + 6]
                    assert not np.any(np.isnan(vals)), "row read before it was zeroed"
                    span = slice(s, s + 9 * d + 6)
                    # a value added twice in one phase (in either order) is a sum of two terms
                    assert np.all(W.total[span][W.concurrent[span]] <= 2), "a sum of more terms was added unordered"
                    # the ordered values (diagonal block, T, F of a node with >= 3 elements) never
                    # took two adds in one phase
                    if ndeg_e[node] >= 3:
                        assert not np.any(W.concurrent[ordered_positions(node, s)])
                    for bi in range(3):
                        start = bi * 3 * nnzN + 3 * rp[node]
                        Kval[start:start + 3 * d] = vals[bi * 3 * d:(bi + 1) * 3 * d]
                    T[[node, Nn + node, 2 * Nn + node]] = vals[9 * d:9 * d + 3]
                    F[[node, Nn + node, 2 * Nn + node]] = vals[9 * d + 3:9 * d + 6]
                    del W.live[s]
            if b == b1:
                break
            # phase C: the row items of b
            W.new_phase()
            els = S["batch_elem"][b * epb:(b + 1) * epb]
            i0, i1 = S["item_ptr"][b], S["item_ptr"][b + 1]
            nst = max([S["item_stage"][k] for k in range(i0, i1)] + [-1]) + 1
            stage = np.full((max(nst, 1), STAGE_VALUES), np.nan)
            for k in range(i0, i1):
                ea = S["item_ea"][k]
                el, a = divmod(int(ea), nod)
                e = els[el]
                assert e >= 0
                node = M.connectivity[e, a]
                s, st = S["item_slot"][k], S["item_stage"][k]
                d = deg[node]
                for bi in range(3):
                    for bj in range(3):
                        for j in range(nod):
                            val = Kel[e, bi, bj, a, j]
                            if st >= 0 and j == a:
                                stage[st, bi * 5 + bj] = val
                            else:
                                W.add(s + bi * 3 * d + bj * d + pos[e, a, j], val)
                    if st >= 0:
                        stage[st, bi * 5 + 3] = Rel[e, a, 2 * bi]
                        stage[st, bi * 5 + 4] = Rel[e, a, 2 * bi + 1]
                    else:
                        W.add(s + 9 * d + bi, Rel[e, a, 2 * bi])
                        W.add(s + 9 * d + 3 + bi, Rel[e, a, 2 * bi + 1])
            W.new_phase()
        assert not W.live, "rows left in the window at the end of a patch"
    assert np.all(written == 1), "every node row is written exactly once"
    return Kval, T, F


@pytest.mark.parametrize("name,epb,window,patch_elems,band", [
    ("tiny_p1", 8, 4000, 20, 3),
    ("tiny_p2", 4, 3000, 12, 2),
    ("tiny_p3", 2, 6000, 16, 2),
    ("cpu_reference", 16, 6000, 250, 8),
    ("medium_p2", 11, 5384, 240, 6),
    ("medium_p3", 4, 6000, 96, 4),
])
def test_schedule_replay_reproduces_assembly(name, epb, window, patch_elems, band, o2lib):
    w = WORKLOADS[name]
    M = build_mesh(w)
    F0, F1, _ = make_fields(M, w)
    cv = default_ctrlvar()
    ref = o2lib.assemble(cv, M, F0, F1)
    S = build_schedule(M, epb, window, patch_elems, band)
    assert S["peak_window"] <= window
    assert S["patch_work"] >= M.Nele
    Kval, T, F = replay(M, S, ref["mesh"].Kel, ref["mesh"].Rel, epb, window)
    assert np.array_equal(S["rowptr_n"], ref["mesh"].rowptr_n)
    assert np.linalg.norm(Kval - ref["Kval"]) <= 1e-13 * np.linalg.norm(ref["Kval"])
    assert np.linalg.norm(T - ref["Tint"]) <= 1e-13 * np.linalg.norm(ref["Tint"])
    assert np.linalg.norm(F - ref["Fext"]) <= 1e-13 * np.linalg.norm(ref["Fext"])


def test_fixup_groups_are_the_nodes_with_several_ordered_items():
    """Within a batch, a node with three or more incident elements and several items has every
    item staged, in batch order, in one fixup group; every other item adds directly."""
    w = WORKLOADS["medium_p2"]
    M = build_mesh(w)
    epb = 12
    S = build_schedule(M, epb, 5384, 240, 6)
    nod = M.nod
    ndeg_e = np.diff(S["inc_ptr"])
    for b in range(len(S["item_ptr"]) - 1):
        els = S["batch_elem"][b * epb:(b + 1) * epb]
        by_node = {}
        for k in range(S["item_ptr"][b], S["item_ptr"][b + 1]):
            el, a = divmod(int(S["item_ea"][k]), nod)
            by_node.setdefault(int(M.connectivity[els[el], a]), []).append(k)
        groups = {int(S["fix_node"][g]): list(S["fix_item"][S["fix_item_ptr"][g]:S["fix_item_ptr"][g + 1]])
                  for g in range(S["fix_ptr"][b], S["fix_ptr"][b + 1])}
        for node, ks in by_node.items():
            stages = [int(S["item_stage"][k]) for k in ks]
            if ndeg_e[node] >= 3 and len(ks) >= 2:
                assert groups[node] == stages and all(x >= 0 for x in stages)
            else:
                assert node not in groups and all(x == -1 for x in stages)


def test_full_size_schedule_fits_the_default_layout():
    """The config-2 mesh (199,712 6-node elements, 7 points) at the kernel's default layout and
    patch count: the window and the ordered staging fit, and the ghost overhead is small."""
    w = WORKLOADS["plain_p2"]
    M = build_mesh(w)
    epb, stage_items, window = fused_layout(6, 7)
    S = build_schedule(M, epb, window)
    assert S["peak_window"] <= window and S["max_stage"] <= stage_items
    overhead = S["patch_work"] / M.Nele - 1
    print(f"plain_p2: epb {epb}, {S['npatch']} patches, peak window {S['peak_window']} of {window} doubles, "
          f"staging {S['max_stage']} of {stage_items}, ghost overhead {overhead:.3f}")
    assert overhead < 0.5


def test_non_conforming_mesh_has_no_schedule():
    """Three elements on one edge: a pair of nodes shares three elements, so the fused schedule
    is refused (the two-pass kernels assemble such meshes)."""
    from uasource_amd.mesh import CreateMUA
    xy = np.array([[0.0, 0.0], [1000.0, 0.0], [500.0, 800.0], [500.0, -800.0], [600.0, 900.0]])
    tri = np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]])
    M = CreateMUA(xy, tri, TriNodes=3)
    msg = build_schedule(M, 4, 4000, expect_ok=False)
    assert "more than two elements" in msg


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
    h = lib.uvhh_build(M.Nele, M.nod, M.Nnodes, conn.ctypes.data, x.ctypes.data, y.ctypes.data, 0, 0, 0, 0, err, 256)
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
