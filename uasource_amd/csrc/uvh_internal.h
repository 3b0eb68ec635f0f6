// Internal declarations shared by the host mesh builder, the kernels and the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct uvh_ctx;

namespace uvh {

constexpr int kMaxNod = 10;
constexpr int kMaxNip = 32;
// Row scatter: lanes per node (16 for 3- and 6-node elements, 32 for 10-node elements, whose
// element rows are 90 values long) and nodes per 256-thread workgroup (consecutive numbers).
constexpr int row_lanes(int nod) { return nod >= 10 ? 32 : 16; }
constexpr int row_nodes_per_block(int nod) { return 4 * (64 / row_lanes(nod)); }
// Bytes per row of the incidence-ordered column table: nod padded to a multiple of 8.
constexpr int pos_inc_stride(int nod) { return (nod + 7) & ~7; }

// Shape-function values and local derivatives at the integration points, plus the
// weights, for one element type and quadrature rule. Filled on the host from the
// quadrature points (the same forms as UaMain/shape_fun.m:52-82 and
// UaMain/shape_der.m:54-120) and copied to the device once per mesh.
struct ShapeTables {
    double N[kMaxNip][kMaxNod];
    double dxi[kMaxNip][kMaxNod];
    double deta[kMaxNip][kMaxNod];
    double w[kMaxNip];
};

// FP32 copies of the shape tables for the precision form's element rows.
struct ShapeTables32 {
    float N[kMaxNip][kMaxNod];
    float dxi[kMaxNip][kMaxNod];
    float deta[kMaxNip][kMaxNod];
};

// Per-mesh host data: the node graph that fixes the CSR pattern of K and the
// tables the kernels use to place element contributions.
struct MeshHost {
    int32_t nele = 0, nod = 0, nn = 0, nip = 0;
    std::vector<int32_t> conn;        // nele*nod, element-major, 0-based
    std::vector<double> x, y;         // nn
    std::vector<int64_t> rowptr_n;    // nn+1: node adjacency (sorted, self included)
    std::vector<int32_t> colind_n;    // nnzN
    std::vector<int32_t> inc_ptr;     // nn+1: node -> incident (element, local node)
    std::vector<int32_t> inc;         // e*nod + a, in element order per node
    std::vector<uint8_t> pos;         // nele*nod*nod: pos[(e*nod+a)*nod+b] = index of conn[e][b] in adj(conn[e][a])
    std::vector<uint8_t> diag;        // nn: index of node k in adj(k)
    std::vector<int32_t> inc_pos;     // nele*nod: position of (e, a) in the incidence order (inverse of inc)
    std::vector<uint8_t> pos_inc;     // nele*nod*pos_inc_stride(nod): pos rows in the incidence order
    // gather form of the row scatter: nzrow[x] is the node of node-graph entry x; for node n,
    // incidence t (t0 <= t < t1) and column c < deg(n), imap[imap_ptr[n] + (t - t0)*deg + c]
    // is the local index in element(t) of node adj(n)[c], or 0xFF if that element lacks it
    std::vector<int32_t> nzrow;       // nnzN
    std::vector<int64_t> imap_ptr;    // nn+1
    std::vector<uint8_t> imap;
    int32_t maxdeg = 0;               // largest node degree
    bool has_isolated = false;        // some node belongs to no element (its rows are empty)
    int64_t max_block_deg = 0;        // largest sum of degrees over row_nodes_per_block(nod) consecutive nodes
    ShapeTables tab{};
    int64_t nnzN() const { return rowptr_n.empty() ? 0 : rowptr_n.back(); }
};

// Fused-assembly schedule (see uvh_schedule.cpp): patches of elements, batches of `epb`
// elements per patch, LDS window slots (offsets in doubles) of the owned node rows.
struct Schedule {
    int epb = 0;
    int64_t window = 0;                    // LDS window capacity in doubles
    int32_t npatch = 0;
    std::vector<int32_t> patch_batch_ptr;  // npatch+1
    std::vector<int32_t> batch_elem;       // nbatch*epb, -1 padded
    // row items of a batch: (element slot, local node) of owned nodes, el*nod + a
    std::vector<int32_t> item_ptr;         // nbatch+1
    std::vector<int32_t> item_ea, item_slot;
    std::vector<int32_t> item_stage;       // staging index of the ordered values, or -1 (direct adds)
    // fixup groups of a batch: nodes with several items whose ordered values are summed in order
    std::vector<int32_t> fix_ptr;          // nbatch+1
    std::vector<int32_t> fix_node, fix_slot;
    std::vector<int32_t> fix_item_ptr;     // nfix+1
    std::vector<int32_t> fix_item;         // staging indices in item order
    std::vector<int32_t> zero_ptr;         // nbatch+1: rows whose first element is in the batch
    std::vector<int32_t> zero_slot, zero_len;
    std::vector<int32_t> flush_ptr;        // nbatch+1: rows whose last element is in the batch
    std::vector<int32_t> flush_node, flush_slot;
    int64_t peak_window = 0;               // largest window use over the patches (doubles)
    int64_t patch_work = 0;                // element visits (>= nele: ghosts are visited twice)
    int32_t max_stage = 0;                 // most staged items in one batch
    int32_t max_items = 0;                 // most row items in one batch
    int32_t nbatch() const { return int32_t(item_ptr.empty() ? 0 : item_ptr.size() - 1); }
};

// Row slot size in doubles: 9*deg K values, then Tx, Ty, Th, Fx, Fy, Fh.
int64_t slot_doubles(int64_t deg);
// Ordered values per staged item: the 3 diagonal entries of its block row's 3 blocks, then T, F,
// for each of the 3 block rows.
constexpr int kStageValues = 15;
// "" if every pair of distinct nodes shares at most two elements (see uvh_schedule.cpp).
std::string check_two_contributions(const MeshHost &m);
// Builds the schedule with patches of about `patch_elems` elements in bands `band_cells`
// element sizes high; patches whose window would exceed `window` doubles are split across
// their sweep direction. Needs rowptr_n, inc_ptr, inc, pos and diag.
std::string build_schedule(MeshHost &m, int epb, int64_t window, int patch_elems, int band_cells, Schedule &S);
// Default patch shape of the fused kernel for `nod`-node elements: elements per patch, band height.
void fused_patch_shape(int nod, int &patch_elems, int &band_cells);

// Builds rowptr_n/colind_n, inc_ptr/inc and pos from conn; returns "" or an error message.
std::string build_mesh_host(MeshHost &m);
// CSC value order of K: csc[k] = csr[map[k]] with CSC colptr/rowind equal to CSR rowptr/colind
// (the pattern is structurally symmetric). map has 9 * nnzN entries.
void build_csc_map(const MeshHost &m, int64_t *map);
// Fills the shape tables for nod nodes at the nip points pts (nip x 2) with weights wts.
std::string fill_shape_tables(ShapeTables &t, int nod, int nip, const double *pts, const double *wts);

// The schedule as the fused kernel reads it: per batch one header, and records that carry
// everything a thread needs in one load (no chains of dependent index loads).
struct BatchHdr {             // ranges of the batch's items, zero, fixup and write-out records
    int32_t item0, item1, zero0, zero1, fix0, fix1, flush0, flush1;
};
struct alignas(16) PatchItem {
    int32_t slot, stage, ea, deg;   // window slot, staging index or -1, el*nod + a, node degree
    uint8_t pos[16];                // column of each element node in the node's row
};
struct alignas(16) PatchFix {       // one node's staged items: stage0 .. stage0 + count - 1
    int32_t slot, stage0, count, deg_dcol;   // deg | diagonal column << 16
};
struct alignas(8) PatchZero { int32_t slot, len; };
struct alignas(16) PatchFlush {
    int32_t node, slot, deg, dcol;
    int64_t rp3, pad;               // 3 * rowptr_n[node]: the row's offset in each block row of K
};
struct PackedSchedule {
    std::vector<BatchHdr> hdr;
    std::vector<int32_t> batch_node;    // nbatch*epb*nod: node of (element slot, local node), -1 padded
    std::vector<PatchItem> items;
    std::vector<PatchFix> fixes;
    std::vector<PatchZero> zeros;
    std::vector<PatchFlush> flushes;
};
// Packs a schedule for the kernel.
void pack_schedule(const MeshHost &m, const Schedule &S, PackedSchedule &P);

// Argument blocks of the fused kernel, kept in device memory and read through one pointer
// each: as kernel arguments the ~30 pointers would not fit in scalar registers.
struct MeshArgs {             // fixed per mesh
    const int32_t *patch_batch_ptr;
    const BatchHdr *hdr;
    const int32_t *batch_node;
    const PatchItem *items;
    const PatchFix *fixes;
    const PatchZero *zeros;
    const PatchFlush *flushes;
    const double *x, *y;
    const ShapeTables *tab;
    double *b;                // nodal lower surface, written by the pre-pass
    int *status;
};

// LDS budget of the fused kernel per workgroup: two workgroups share a CU (160 KiB).
constexpr int64_t kFusedLdsBytes = 80 * 1024;

// LDS layout of the fused kernel (in doubles) and its elements per batch: shape tables,
// nodal staging, coefficient records, ordered staging (stage_items x kStageValues), window.
struct FusedLayout {
    int table_doubles, nodal_doubles, rec_doubles, epb, stage_items;
    int64_t window_doubles;
    int64_t lds_bytes() const {
        return (int64_t(table_doubles) + nodal_doubles + rec_doubles + int64_t(stage_items) * kStageValues +
                window_doubles) * int64_t(sizeof(double));
    }
};
FusedLayout fused_layout(int nod, int nip);

// Scalars of the assembly, device-ready (cos/sin of alpha precomputed).
struct Params {
    double dt, theta, beta0, tau_mult, uz, eps0, eta0, czero, kH, Hh0, g, rhow, ca, sa, thick_min;
    int32_t tau_opt, zero_fields, reset_thickness, enforce_b_above;
    // uv (diagnostic) assembly: uv != 0 clips C and AGlen at cmin and aglenmin and skips the
    // thickness rows; dvisk scales the viscosity-derivative terms (NRviscosity; 1 for uvh)
    double cmin, aglenmin, dvisk;
    int32_t uv, pad;
    // sliding law (enum uvh_sliding_law) and the lower bound of the effective pressure
    double nzero;
    int32_t law, pad2;
};

// Device pointers of the nodal fields, in the order of uvh_fields (dadh, q, muk and V0 may be
// null; q, muk and V0 are read only by the sliding laws that use them).
struct FieldPtrs {
    const double *h, *u, *v, *S, *B, *rho, *C, *m, *n, *A, *as1, *ab1;
    const double *h0, *u0, *v0, *as0, *ab0;
    const double *dadh;
    const double *q, *muk, *V0;
};

// One mesh of a batched assembly, as the batched kernels read it (device memory). The
// mesh tables and work buffers are fixed when the batch is built; the fields and the
// outputs are set per call.
struct BatchMember {
    const int32_t *conn;
    const double *x, *y;
    const ShapeTables *tab;
    const ShapeTables32 *tab32;
    const int32_t *inc_ptr, *inc_pos;
    const uint8_t *pos_inc, *diag;
    const int64_t *rowptr_n;
    double *b;
    void *EB, *RE;
    int64_t nnzN;
    int32_t nele, nn, nod, nip, epb, pad;
    FieldPtrs f;
    double *R, *Tint, *Fext, *Kval;
};

// Per-assembly argument block of the fused kernel (device memory; see MeshArgs).
struct CallArgs {             // per assembly
    FieldPtrs f;
    Params p;
    double *R, *Tint, *Fext, *Kval;
};

// One member of a batched residual cost (device memory): R, the step d and the weights w
// (null: all 1) of n entries; its chunks are partial[first .. first + nchunk - 1].
struct CostMember {
    const double *R, *d, *w;
    int64_t n;
    int32_t first, nchunk;
};

// Read access to a context's mesh (null before uvh_set_mesh) and device, for the parts of
// the library outside uvh_api.hip.
const MeshHost *ctx_mesh(const uvh_ctx *ctx);
int ctx_device(const uvh_ctx *ctx);

}  // namespace uvh
