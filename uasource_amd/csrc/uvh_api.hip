// C ABI of libuvh (declared in include/uvh.h): context, mesh upload, assembly entry points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/uvh.h"
#include "uvh_internal.h"

namespace uvh {
hipError_t launch_nodal_b(int nn, const FieldPtrs &f, const Params &p, double *b, int *status, hipStream_t s);
hipError_t launch_nodal_s(int nn, const double *sfc, const double *h, double *b, int *status, hipStream_t s);
hipError_t launch_element(int nod, int nele, int nip, int epb, int nthr, bool fp32, bool full, const int32_t *conn,
                          const double *x, const double *y, const FieldPtrs &f, const double *b,
                          const ShapeTables *tab, const ShapeTables32 *tab32, const Params &p,
                          const int32_t *inc_pos, bool ronly, void *EB, void *RE, int lb0, int lb1, int *status,
                          hipStream_t s);
hipError_t launch_rows(int nod, int nn, int64_t block_vals, bool fp32, const int32_t *inc_ptr, const uint8_t *pos,
                       const int64_t *rowptr_n, const uint8_t *diag, int64_t nnzN, const void *EB, const void *RE,
                       double *R, double *Tint, double *Fext, double *Kval, int check_nan, int *status,
                       const int32_t *blocks, int nblk, hipStream_t s, int nbk = 3);
int element_epb_max(int nod, int nip, bool fp32, size_t lds_cap, int nthr);
hipError_t launch_rows_gather(int nod, int nn, int64_t nnzN, bool fp32, const int32_t *nzrow, const int64_t *rowptr_n,
                              const int32_t *inc_ptr, const int64_t *imap_ptr, const uint8_t *imap,
                              const uint8_t *diag, const void *EB, const void *RE, double *R, double *Tint,
                              double *Fext, double *Kval, int check_nan, int *status, hipStream_t s);
size_t element_lds_bytes(int nod, int nip, int epb, bool fp32);
hipError_t launch_batch(const BatchMember *mem, int nmem, const int2 *nodal_items, int n_nodal,
                        const int2 *elem_items, const int *eb, const int2 *row_items, const int *rb,
                        const size_t *elem_lds, const size_t *row_lds, bool fp32, bool ronly, int check_nan,
                        int *ctr, int *status, const Params &p, bool flat, hipStream_t s);
int element_epb_default(int nod, int nip, bool fp32, int nthr);
int element_epb_uvh(int nod, int nip, bool fp32, int nthr);
size_t element_row_stride(int nod, bool fp32);
hipError_t launch_fused(int nod, int nip, int npatch, const FusedLayout &L, const MeshArgs *MA, const CallArgs *CA,
                        int64_t nnzN, int nn, bool sl, bool ronly, int check_nan, hipStream_t s);
hipError_t launch_cost(const CostMember *mem, int nmem, const int2 *items, int nitems, double2 *partial, double *rr,
                       double *d2, hipStream_t s);
constexpr int kCostChunk = 32768;
}  // namespace uvh

namespace {

// Chunks of the pipelined R+K assembly (UVH_CHUNKS overrides; 1 = element kernel, then rows).
constexpr int kDefaultChunks = 1;

struct KernelTimer {
    const char *name;
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct PendingEvent {
    int slot;
    hipEvent_t a, b;
};

template <class T>
hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
}

}  // namespace

struct uvh_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    std::string err;
    bool has_mesh = false;
    uvh::MeshHost mesh;
    // device mesh
    int32_t *d_conn = nullptr;
    double *d_x = nullptr, *d_y = nullptr;
    uvh::ShapeTables *d_tab = nullptr;
    uvh::ShapeTables32 *d_tab32 = nullptr;
    int64_t *d_rowptr_n = nullptr;
    int32_t *d_inc_ptr = nullptr, *d_inc = nullptr;
    uint8_t *d_pos = nullptr, *d_diag = nullptr, *d_pos_inc = nullptr;
    int32_t *d_inc_pos = nullptr;
    int32_t *d_nzrow = nullptr;            // gather form of the row scatter
    int64_t *d_imap_ptr = nullptr;
    uint8_t *d_imap = nullptr;
    bool gather_rows = false;              // UVH_SCATTER=gather
    // pipelined R+K assembly: element tiles in `nchunks` chunks on the caller's stream, the node
    // rows of each chunk (the node blocks whose incident elements all lie in chunks <= c) on
    // `aux_stream` as soon as their chunk is done; [FP64, FP32] plans (the tiles differ)
    int nchunks = 1;                       // UVH_CHUNKS
    hipStream_t aux_stream = nullptr;
    std::vector<hipEvent_t> chunk_ev;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool plan_ok[2] = {false, false};
    std::vector<int> plan_tiles[2];        // nchunks+1 tile boundaries
    std::vector<int> plan_boff[2];         // nchunks+1 offsets into the block list
    int32_t *d_plan_blocks[2] = {nullptr, nullptr};
    // fused-assembly schedule (UVH_ASSEMBLY=fused): FP64 R+K and R assemblies take the fused
    // kernel when the mesh has a schedule (every node in an element, at most two elements per
    // pair of nodes); the precision form and the uv assembly take the two-pass kernels
    bool fused_wanted = false;
    bool fused_ok = false;      // this mesh has a fused schedule
    std::string fused_why;      // why not, when it has none
    uvh::Schedule sched;
    uvh::FusedLayout layout{};
    uvh::PackedSchedule psched;       // host copy of the packed schedule (kept for its sizes)
    std::vector<void *> d_sched;      // the packed schedule's device arrays
    uvh::MeshArgs *d_margs = nullptr;
    // per-call argument blocks: a ring of pinned host copies and device copies; a slot is
    // reused only after the event recorded behind its upload has completed
    static constexpr int kRing = 8;
    uvh::CallArgs *h_cargs = nullptr, *d_cargs = nullptr;
    hipEvent_t ring_ev[kRing] = {};
    bool ring_used[kRing] = {};
    int ring_pos = 0, last_slot = -1;
    uvh::CallArgs last_args{};
    // work
    double *d_b = nullptr, *d_RE = nullptr, *d_EB = nullptr;
    double *d_zero = nullptr;   // uv assembly: zero nodal field (lazily allocated)
    int epb[2] = {0, 0};    // elements per block of the element kernel, uvh R+K: [FP64, FP32]
    int epb_wide[2] = {0, 0};   // the same for the uv assembly and uvh R-only calls
    // threads per element block (UVH_ELEMENT_THREADS=64/128/192: one- to three-wave tiles, each
    // with as many elements as its stages fit; measured slower than 256 on 6- and 10-node meshes)
    int el_threads[2] = {256, 256};
    int epb_override = 0;   // UVH_EPB: elements per element block, where the tile fits
    int bepb[2] = {0, 0};   // elements per 256-thread block as a batch member (uvhb_set_tile)
    // one-pass element rows (all nine blocks per thread) for 3- and 6-node elements: faster with
    // FP32 arithmetic; with FP64 the three-pass rows keep the registers for two waves per SIMD.
    // UVH_ROWS=full or =split forces one form for both.
    int rows_mode = 0;      // 0 automatic, 1 one-pass, 2 three-pass
    int *d_status = nullptr;
    int *h_status = nullptr;    // pinned copy of d_status for uvh_sync_status
    int *last_status = nullptr; // status word of the last call (d_status, or the scratch array's)
    int scratch_rc = UVH_E_HIP; // whether the scratch array's address was resolved at uvh_create
    // One assembly at a time per context: calls share the work buffers and the status word. A
    // call enqueued on another stream than the previous one waits for the previous one's
    // kernels (done_ev, recorded at the end of every call outside graph capture).
    hipEvent_t done_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool inflight = false;
    // host-path staging
    double *d_fields = nullptr;  // 18 * nn
    double *d_R = nullptr, *d_T = nullptr, *d_F = nullptr, *d_K = nullptr;
    // profiling
    bool profiling = false;
    std::vector<KernelTimer> timers;
    std::vector<PendingEvent> pending;

    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char *what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return UVH_E_HIP;
    }
    void free_mesh() {
        hipFree(d_conn); hipFree(d_x); hipFree(d_y); hipFree(d_tab); hipFree(d_tab32); hipFree(d_rowptr_n);
        d_tab32 = nullptr;
        hipFree(d_inc_ptr); hipFree(d_inc); hipFree(d_pos); hipFree(d_b); hipFree(d_RE); hipFree(d_EB);
        hipFree(d_fields); hipFree(d_R); hipFree(d_T); hipFree(d_F); hipFree(d_K);
        hipFree(d_diag);
        for (void *q : d_sched) hipFree(q);
        d_sched.clear();
        fused_ok = false;
        hipFree(d_margs); hipFree(d_pos_inc); hipFree(d_inc_pos);
        hipFree(d_nzrow); hipFree(d_imap_ptr); hipFree(d_imap); hipFree(d_zero);
        d_nzrow = nullptr; d_imap_ptr = nullptr; d_imap = nullptr; d_zero = nullptr;
        for (int t = 0; t < 2; ++t) {
            hipFree(d_plan_blocks[t]);
            d_plan_blocks[t] = nullptr;
            plan_ok[t] = false;
        }
        d_margs = nullptr; d_pos_inc = nullptr; d_inc_pos = nullptr;
        last_slot = -1;
        d_diag = nullptr;
        d_conn = nullptr; d_x = d_y = nullptr; d_tab = nullptr; d_rowptr_n = nullptr;
        d_inc_ptr = d_inc = nullptr; d_pos = nullptr; d_b = d_RE = d_EB = nullptr;
        d_fields = d_R = d_T = d_F = d_K = nullptr;
        has_mesh = false;
    }
    int timer_slot(const char *name) {
        for (size_t k = 0; k < timers.size(); ++k)
            if (std::strcmp(timers[k].name, name) == 0) return int(k);
        timers.push_back(KernelTimer{name});
        return int(timers.size() - 1);
    }
    // Brackets one launch with events when profiling is on.
    template <class F>
    hipError_t timed(const char *name, hipStream_t s, F &&launch) {
        if (!profiling) return launch();
        PendingEvent pe{timer_slot(name), nullptr, nullptr};
        hipEventCreate(&pe.a);
        hipEventCreate(&pe.b);
        hipEventRecord(pe.a, s);
        hipError_t e = launch();
        hipEventRecord(pe.b, s);
        pending.push_back(pe);
        return e;
    }
    void resolve_timers() {
        for (auto &pe : pending) {
            hipEventSynchronize(pe.b);
            float ms = 0.f;
            hipEventElapsedTime(&ms, pe.a, pe.b);
            timers[pe.slot].total_ms += ms;
            timers[pe.slot].launches += 1;
            hipEventDestroy(pe.a);
            hipEventDestroy(pe.b);
        }
        pending.clear();
    }
};

namespace {

uvh::Params make_params(const uvh_ctrl &c, uint32_t flags) {
    uvh::Params p{};
    p.dt = c.dt; p.theta = c.theta; p.beta0 = c.supg_beta0; p.tau_mult = c.supg_tau_multiplier;
    p.uz = c.speed_zero; p.eps0 = c.eps_zero; p.eta0 = c.eta_zero; p.czero = c.czero;
    p.kH = c.kH; p.Hh0 = c.Hh0; p.g = c.g; p.rhow = c.rhow;
    p.ca = std::cos(c.alpha); p.sa = std::sin(c.alpha);
    p.thick_min = c.thick_min;
    p.tau_opt = c.supg_tau;
    p.zero_fields = (flags & UVH_ZEROFIELDS) ? 1 : 0;
    p.reset_thickness = c.reset_thickness;
    p.enforce_b_above = c.enforce_b_above_B;
    p.cmin = 0.0; p.aglenmin = 0.0; p.dvisk = 1.0; p.uv = 0;
    p.nzero = c.nzero; p.law = c.sliding_law;
    return p;
}

// UVH_OK, or why the sliding law cannot run: an unknown law, or a field it reads is missing
int check_law(const uvh_ctrl &c, const double *q, const double *muk, const double *V0, std::string &why) {
    switch (c.sliding_law) {
        case UVH_SLIDING_WEERTMAN: return UVH_OK;
        case UVH_SLIDING_BUDD:
            if (q) return UVH_OK;
            why = "sliding law Budd needs q";
            return UVH_E_ARG;
        case UVH_SLIDING_COULOMB:
        case UVH_SLIDING_TSAI:
        case UVH_SLIDING_CORNFORD:
        case UVH_SLIDING_UMBI:
            if (muk) return UVH_OK;
            why = "sliding laws Coulomb, Tsai, Cornford and Umbi need muk";
            return UVH_E_ARG;
        case UVH_SLIDING_JOUGHIN:
            if (V0) return UVH_OK;
            why = "sliding law Joughin needs V0";
            return UVH_E_ARG;
        default: why = "sliding law " + std::to_string(c.sliding_law) + " is not supported"; return UVH_E_UNSUPPORTED;
    }
}

constexpr int kHostFields = 21;   // uvh_fields arrays the host entry points upload

bool fields_complete(const uvh_fields &f) {
    const double *req[] = {f.h, f.ub, f.vb, f.S, f.B, f.rho, f.C, f.m, f.n, f.AGlen, f.as, f.ab,
                           f.h0, f.ub0, f.vb0, f.as0, f.ab0};
    for (auto p : req)
        if (!p) return false;
    return true;
}

uvh::FieldPtrs to_ptrs(const uvh_fields &f) {
    return uvh::FieldPtrs{f.h, f.ub, f.vb, f.S, f.B, f.rho, f.C, f.m, f.n, f.AGlen, f.as, f.ab,
                          f.h0, f.ub0, f.vb0, f.as0, f.ab0, f.dadh, f.q, f.muk, f.V0};
}

// Chunk plan of the pipelined assembly for element tiles of `epb` elements: tile boundaries
// of nchunks equal chunks, and for each chunk the node blocks (row_nodes_per_block consecutive
// nodes) whose last incident element lies in it. The incidence lists are in element order, so
// a node's last incident element is the last entry of its list.
int build_plan(uvh_ctx *ctx, int t) {
    const uvh::MeshHost &m = ctx->mesh;
    const int epb = ctx->epb[t];
    const int ntiles = (m.nele + epb - 1) / epb;
    const int C = std::max(1, std::min(ctx->nchunks, ntiles));
    std::vector<int> tiles(size_t(C) + 1);
    for (int c = 0; c <= C; ++c) tiles[c] = int((int64_t(ntiles) * c) / C);
    const int nbn = uvh::row_nodes_per_block(m.nod);
    const int nblk = (m.nn + nbn - 1) / nbn;
    std::vector<std::vector<int32_t>> per(static_cast<size_t>(C));
    for (int b = 0; b < nblk; ++b) {
        int last = 0;
        for (int k = b * nbn; k < std::min(m.nn, (b + 1) * nbn); ++k)
            if (m.inc_ptr[k + 1] > m.inc_ptr[k]) last = std::max(last, m.inc[m.inc_ptr[k + 1] - 1] / m.nod);
        const int tile = last / epb;
        const int c = int(std::upper_bound(tiles.begin(), tiles.end(), tile) - tiles.begin()) - 1;
        per[size_t(std::max(0, std::min(C - 1, c)))].push_back(b);
    }
    std::vector<int32_t> flat;
    std::vector<int> off(size_t(C) + 1, 0);
    for (int c = 0; c < C; ++c) {
        off[c] = int(flat.size());
        flat.insert(flat.end(), per[c].begin(), per[c].end());
    }
    off[C] = int(flat.size());
    hipFree(ctx->d_plan_blocks[t]);
    ctx->d_plan_blocks[t] = nullptr;
    if (dalloc(&ctx->d_plan_blocks[t], flat.size()) != hipSuccess ||
        hipMemcpy(ctx->d_plan_blocks[t], flat.data(), flat.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
            hipSuccess)
        return ctx->fail(UVH_E_NOMEM, "uvh_assemble: cannot upload the chunk plan");
    ctx->plan_tiles[t] = tiles;
    ctx->plan_boff[t] = off;
    ctx->plan_ok[t] = true;
    return UVH_OK;
}

}  // namespace

// ------------------------------------------------------------------ allocation-free assembly
// uvh_assemble_device_noalloc takes its work buffers (the status word, the nodal lower
// surface b, the element residual rows and, for R+K, the element rows) from `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs on holds until the
// process ends. Host code never passes `scratch` itself (on the host the name is only a
// stand-in): uvh_create resolves its device address once per device, with that device
// current, and the calls use that address. All calls on one GPU share the array, so calls
// that overlap on one GPU race; the status word is cleared by the nodal pre-pass of each call,
// and every other scratch value is written before it is read.
namespace {
constexpr size_t kScratchBytes = size_t(256) << 20;   // 256 MiB per GPU
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the device address of this device's copy of `scratch`.
int resolve_scratch(int device, std::string &err) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) {
        err = "allocation-free assembly: device index out of range";
        return UVH_E_ARG;
    }
    if (g_scratch_ptr[device]) return UVH_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        err = std::string("allocation-free assembly: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return UVH_E_HIP;
    }
    g_scratch_ptr[device] = ptr;
    return UVH_OK;
}

struct NoallocLayout {
    size_t status, b, RE, EB, total;   // byte offsets into the scratch array, and its use
};
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
NoallocLayout noalloc_layout(const uvh::MeshHost &m, uint32_t flags) {
    NoallocLayout L{};
    L.status = 0;
    L.b = 256;
    L.RE = L.b + align256(size_t(m.nn) * sizeof(double));
    L.EB = L.RE + align256(size_t(m.nele) * m.nod * 6 * sizeof(double));
    const bool ronly = (flags & UVH_RONLY) != 0;
    L.total = L.EB + (ronly ? 0 : align256(size_t(m.nele) * m.nod * uvh::element_row_stride(m.nod, false) * sizeof(double)));
    return L;
}
}  // namespace

extern "C" {

int uvh_create(uvh_ctx **out, int device) {
    if (!out) return UVH_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return UVH_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return UVH_E_HIP;
    auto *ctx = new uvh_ctx();
    ctx->device = device;
    // UVH_ASSEMBLY=fused selects the single-kernel LDS-window assembly for FP64; the default is
    // the two-pass path (element rows through HBM, then one lane group per node), which is
    // faster on MI355X (measured: 0.59 ms against 1.27-1.9 ms at 199,712 6-node elements).
    const char *am = std::getenv("UVH_ASSEMBLY");
    ctx->fused_wanted = am && std::strcmp(am, "fused") == 0;
    const char *rows = std::getenv("UVH_ROWS");
    ctx->rows_mode = !rows ? 0 : std::strcmp(rows, "full") == 0 ? 1 : std::strcmp(rows, "split") == 0 ? 2 : 0;
    const char *sc = std::getenv("UVH_SCATTER");
    ctx->gather_rows = sc && std::strcmp(sc, "gather") == 0;
    const char *et = std::getenv("UVH_ELEMENT_THREADS");
    const int nt = et ? std::atoi(et) : 0;
    if (nt == 64 || nt == 128 || nt == 192 || nt == 256) ctx->el_threads[0] = ctx->el_threads[1] = nt;
    const char *eo = std::getenv("UVH_EPB");
    ctx->epb_override = eo ? std::atoi(eo) : 0;
    const char *ch = std::getenv("UVH_CHUNKS");
    ctx->nchunks = ch ? std::max(1, std::min(64, std::atoi(ch))) : kDefaultChunks;
    hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&ctx->d_status), sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&ctx->h_status), sizeof(int), 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->done_ev, hipEventDisableTiming);
    if (e == hipSuccess) ctx->scratch_rc = resolve_scratch(device, ctx->err);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&ctx->h_cargs), sizeof(uvh::CallArgs) * uvh_ctx::kRing, 0);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&ctx->d_cargs), sizeof(uvh::CallArgs) * uvh_ctx::kRing);
    for (int k = 0; e == hipSuccess && k < uvh_ctx::kRing; ++k) e = hipEventCreateWithFlags(&ctx->ring_ev[k], hipEventDisableTiming);
    if (e == hipSuccess && ctx->nchunks > 1) {
        e = hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking);  // default priority
        ctx->chunk_ev.assign(size_t(ctx->nchunks), nullptr);
        for (int k = 0; e == hipSuccess && k < ctx->nchunks; ++k)
            e = hipEventCreateWithFlags(&ctx->chunk_ev[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        delete ctx;
        return UVH_E_HIP;
    }
    *out = ctx;
    return UVH_OK;
}

void uvh_destroy(uvh_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    ctx->resolve_timers();
    ctx->free_mesh();
    hipFree(ctx->d_status);
    if (ctx->h_status) hipHostFree(ctx->h_status);
    if (ctx->done_ev) hipEventDestroy(ctx->done_ev);
    for (int k = 0; k < uvh_ctx::kRing; ++k)
        if (ctx->ring_ev[k]) hipEventDestroy(ctx->ring_ev[k]);
    hipHostFree(ctx->h_cargs);
    hipFree(ctx->d_cargs);
    for (auto ev : ctx->chunk_ev)
        if (ev) hipEventDestroy(ev);
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) hipEventDestroy(ctx->join_ev);
    if (ctx->aux_stream) hipStreamDestroy(ctx->aux_stream);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char *uvh_last_error(const uvh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int uvh_set_mesh(uvh_ctx *ctx, int32_t nele, int32_t nod, int32_t nnodes, const int32_t *conn, const double *x,
                 const double *y, int32_t nip, const double *pts, const double *wts) {
    if (!ctx) return UVH_E_ARG;
    if (nele <= 0 || nnodes <= 0 || !conn || !x || !y || !pts || !wts)
        return ctx->fail(UVH_E_ARG, "uvh_set_mesh: empty mesh or null pointer");
    if (nod != 3 && nod != 6 && nod != 10) return ctx->fail(UVH_E_ARG, "uvh_set_mesh: nod must be 3, 6 or 10");
    if (nip < 1 || nip > uvh::kMaxNip) return ctx->fail(UVH_E_ARG, "uvh_set_mesh: nip must be in 1..32");
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();  // buffers may be in use on any stream the caller passed
    ctx->free_mesh();
    uvh::MeshHost &m = ctx->mesh;
    m = uvh::MeshHost{};
    m.nele = nele; m.nod = nod; m.nn = nnodes; m.nip = nip;
    m.conn.assign(conn, conn + size_t(nele) * nod);
    m.x.assign(x, x + nnodes);
    m.y.assign(y, y + nnodes);
    std::string msg = uvh::fill_shape_tables(m.tab, nod, nip, pts, wts);
    if (msg.empty()) msg = uvh::build_mesh_host(m);
    if (!msg.empty()) return ctx->fail(UVH_E_ARG, "uvh_set_mesh: " + msg);
    ctx->layout = uvh::fused_layout(nod, nip);
    ctx->sched = uvh::Schedule{};
    ctx->fused_why.clear();
    bool fused = ctx->fused_wanted;
    if (fused && m.has_isolated) {
        fused = false;
        ctx->fused_why = "a node belongs to no element";
    }
    if (fused) {  // the fused kernel's patch schedule
        int pelems = 0, band = 0;
        uvh::fused_patch_shape(nod, pelems, band);
        std::string why = uvh::build_schedule(m, ctx->layout.epb, ctx->layout.window_doubles, pelems, band, ctx->sched);
        if (why.empty() && ctx->sched.max_stage > ctx->layout.stage_items) {
            // more ordered items in a batch than the staging holds: grow it, shrink the window
            const int grow = ctx->sched.max_stage - ctx->layout.stage_items;
            ctx->layout.stage_items = ctx->sched.max_stage;
            ctx->layout.window_doubles -= ((int64_t(grow) * uvh::kStageValues + 1) / 2) * 2;
            why = uvh::build_schedule(m, ctx->layout.epb, ctx->layout.window_doubles, pelems, band, ctx->sched);
            if (why.empty() && ctx->sched.max_stage > ctx->layout.stage_items) why = "ordered staging does not fit";
        }
        if (!why.empty()) {
            fused = false;
            ctx->fused_why = why;
            ctx->sched = uvh::Schedule{};
        }
    }
    hipError_t e = hipSuccess;
    const size_t ne = size_t(nele), nn = size_t(nnodes);
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    chk(dalloc(&ctx->d_conn, ne * nod));
    chk(dalloc(&ctx->d_x, nn));
    chk(dalloc(&ctx->d_y, nn));
    chk(dalloc(&ctx->d_tab, 1));
    chk(dalloc(&ctx->d_tab32, 1));
    chk(dalloc(&ctx->d_rowptr_n, nn + 1));
    chk(dalloc(&ctx->d_inc_ptr, nn + 1));
    chk(dalloc(&ctx->d_inc, ne * nod));
    chk(dalloc(&ctx->d_pos, ne * nod * nod));
    chk(dalloc(&ctx->d_b, nn));
    chk(dalloc(&ctx->d_diag, nn));
    // the packed schedule: patch ranges, batch headers, gather nodes, item/fixup/zero/write-out records
    uvh::PackedSchedule &P = ctx->psched;
    P = uvh::PackedSchedule{};
    struct Arr { const void *src; size_t bytes; };
    std::vector<Arr> arrs;
    if (fused) {
        uvh::pack_schedule(m, ctx->sched, P);
        const uvh::Schedule &S = ctx->sched;
        arrs = {{S.patch_batch_ptr.data(), S.patch_batch_ptr.size() * sizeof(int32_t)},
                {P.hdr.data(), P.hdr.size() * sizeof(uvh::BatchHdr)},
                {P.batch_node.data(), P.batch_node.size() * sizeof(int32_t)},
                {P.items.data(), P.items.size() * sizeof(uvh::PatchItem)},
                {P.fixes.data(), P.fixes.size() * sizeof(uvh::PatchFix)},
                {P.zeros.data(), P.zeros.size() * sizeof(uvh::PatchZero)},
                {P.flushes.data(), P.flushes.size() * sizeof(uvh::PatchFlush)}};
        ctx->d_sched.assign(arrs.size(), nullptr);
        for (size_t k = 0; k < arrs.size(); ++k)
            chk(hipMalloc(&ctx->d_sched[k], std::max<size_t>(16, arrs[k].bytes)));
    }
    {
        chk(dalloc(&ctx->d_RE, ne * nod * 6));
        chk(dalloc(&ctx->d_inc_pos, ne * nod));
        chk(dalloc(&ctx->d_pos_inc, ne * nod * uvh::pos_inc_stride(nod)));
        if (ctx->gather_rows) {
            chk(dalloc(&ctx->d_nzrow, m.nzrow.size()));
            chk(dalloc(&ctx->d_imap_ptr, m.imap_ptr.size()));
            chk(dalloc(&ctx->d_imap, m.imap.size()));
        }
    }
    if (e != hipSuccess) {
        ctx->free_mesh();
        return ctx->fail(UVH_E_NOMEM, std::string("uvh_set_mesh: device allocation failed: ") + hipGetErrorString(e));
    }
    chk(hipMemcpy(ctx->d_conn, m.conn.data(), ne * nod * sizeof(int32_t), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_x, m.x.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_y, m.y.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_tab, &m.tab, sizeof(uvh::ShapeTables), hipMemcpyHostToDevice));
    {
        uvh::ShapeTables32 t32{};
        for (int q = 0; q < uvh::kMaxNip; ++q)
            for (int a = 0; a < uvh::kMaxNod; ++a) {
                t32.N[q][a] = float(m.tab.N[q][a]);
                t32.dxi[q][a] = float(m.tab.dxi[q][a]);
                t32.deta[q][a] = float(m.tab.deta[q][a]);
            }
        chk(hipMemcpy(ctx->d_tab32, &t32, sizeof(t32), hipMemcpyHostToDevice));
    }
    chk(hipMemcpy(ctx->d_rowptr_n, m.rowptr_n.data(), (nn + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_inc_ptr, m.inc_ptr.data(), (nn + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_inc, m.inc.data(), ne * nod * sizeof(int32_t), hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_pos, m.pos.data(), ne * nod * nod, hipMemcpyHostToDevice));
    chk(hipMemcpy(ctx->d_diag, m.diag.data(), nn, hipMemcpyHostToDevice));
    {
        chk(hipMemcpy(ctx->d_inc_pos, m.inc_pos.data(), ne * nod * sizeof(int32_t), hipMemcpyHostToDevice));
        chk(hipMemcpy(ctx->d_pos_inc, m.pos_inc.data(), m.pos_inc.size(), hipMemcpyHostToDevice));
        if (ctx->gather_rows) {
            chk(hipMemcpy(ctx->d_nzrow, m.nzrow.data(), m.nzrow.size() * sizeof(int32_t), hipMemcpyHostToDevice));
            chk(hipMemcpy(ctx->d_imap_ptr, m.imap_ptr.data(), m.imap_ptr.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
            chk(hipMemcpy(ctx->d_imap, m.imap.data(), m.imap.size(), hipMemcpyHostToDevice));
        }
    }
    if (fused && e == hipSuccess)
        for (size_t k = 0; k < arrs.size(); ++k)
            if (arrs[k].bytes) chk(hipMemcpy(ctx->d_sched[k], arrs[k].src, arrs[k].bytes, hipMemcpyHostToDevice));
    chk(dalloc(&ctx->d_margs, 1));
    if (e == hipSuccess) {
        uvh::MeshArgs ma{};
        if (fused) {
            ma.patch_batch_ptr = static_cast<const int32_t *>(ctx->d_sched[0]);
            ma.hdr = static_cast<const uvh::BatchHdr *>(ctx->d_sched[1]);
            ma.batch_node = static_cast<const int32_t *>(ctx->d_sched[2]);
            ma.items = static_cast<const uvh::PatchItem *>(ctx->d_sched[3]);
            ma.fixes = static_cast<const uvh::PatchFix *>(ctx->d_sched[4]);
            ma.zeros = static_cast<const uvh::PatchZero *>(ctx->d_sched[5]);
            ma.flushes = static_cast<const uvh::PatchFlush *>(ctx->d_sched[6]);
        }
        ma.x = ctx->d_x; ma.y = ctx->d_y;
        ma.tab = ctx->d_tab; ma.b = ctx->d_b; ma.status = ctx->d_status;
        chk(hipMemcpy(ctx->d_margs, &ma, sizeof(ma), hipMemcpyHostToDevice));
    }
    if (e != hipSuccess) {
        ctx->free_mesh();
        return ctx->hip_fail(e, "uvh_set_mesh: upload");
    }
    for (int t = 0; t < 2; ++t) {
        ctx->epb[t] = uvh::element_epb_uvh(nod, nip, t == 1, ctx->el_threads[t]);
        ctx->epb_wide[t] = uvh::element_epb_default(nod, nip, t == 1, ctx->el_threads[t]);
        const int o = ctx->epb_override;
        if (o > 0 && o * nip <= ctx->el_threads[t] && o * nod <= ctx->el_threads[t] &&
            uvh::element_lds_bytes(nod, nip, o, t == 1) <= 156 * 1024)
            ctx->epb[t] = ctx->epb_wide[t] = o;
        ctx->bepb[t] = uvh::element_epb_default(nod, nip, t == 1, 256);   // batches: autotuned (uvhb_set_tile)
    }
    ctx->fused_ok = fused;
    ctx->has_mesh = true;
    return UVH_OK;
}

int uvh_pattern_info(const uvh_ctx *ctx, int64_t *neq, int64_t *nnz) {
    if (!ctx || !ctx->has_mesh) return UVH_E_STATE;
    if (neq) *neq = 3 * int64_t(ctx->mesh.nn);
    if (nnz) *nnz = 9 * ctx->mesh.nnzN();
    return UVH_OK;
}

int uvh_get_pattern(const uvh_ctx *ctx, int64_t *rowptr, int32_t *colind) {
    if (!ctx || !ctx->has_mesh) return UVH_E_STATE;
    if (!rowptr || !colind) return UVH_E_ARG;
    const uvh::MeshHost &m = ctx->mesh;
    const int64_t nn = m.nn, nnzN = m.nnzN();
    for (int bi = 0; bi < 3; ++bi)
        for (int64_t k = 0; k < nn; ++k) {
            const int64_t r = bi * nn + k;
            const int64_t start = int64_t(bi) * 3 * nnzN + 3 * m.rowptr_n[k];
            const int64_t deg = m.rowptr_n[k + 1] - m.rowptr_n[k];
            rowptr[r] = start;
            for (int bj = 0; bj < 3; ++bj)
                for (int64_t t = 0; t < deg; ++t)
                    colind[start + bj * deg + t] = int32_t(m.colind_n[m.rowptr_n[k] + t] + bj * nn);
        }
    rowptr[3 * nn] = 9 * nnzN;
    return UVH_OK;
}

int uvh_uv_pattern_info(const uvh_ctx *ctx, int64_t *neq, int64_t *nnz) {
    if (!ctx || !ctx->has_mesh) return UVH_E_STATE;
    if (neq) *neq = 2 * int64_t(ctx->mesh.nn);
    if (nnz) *nnz = 4 * ctx->mesh.nnzN();
    return UVH_OK;
}

int uvh_uv_get_pattern(const uvh_ctx *ctx, int64_t *rowptr, int32_t *colind) {
    if (!ctx || !ctx->has_mesh) return UVH_E_STATE;
    if (!rowptr || !colind) return UVH_E_ARG;
    const uvh::MeshHost &m = ctx->mesh;
    const int64_t nn = m.nn, nnzN = m.nnzN();
    for (int bi = 0; bi < 2; ++bi)
        for (int64_t k = 0; k < nn; ++k) {
            const int64_t start = int64_t(bi) * 2 * nnzN + 2 * m.rowptr_n[k];
            const int64_t deg = m.rowptr_n[k + 1] - m.rowptr_n[k];
            rowptr[bi * nn + k] = start;
            for (int bj = 0; bj < 2; ++bj)
                for (int64_t t = 0; t < deg; ++t)
                    colind[start + bj * deg + t] = int32_t(m.colind_n[m.rowptr_n[k] + t] + bj * nn);
        }
    rowptr[2 * nn] = 4 * nnzN;
    return UVH_OK;
}

int uvh_get_csc_map(const uvh_ctx *ctx, int64_t *map) {
    if (!ctx || !ctx->has_mesh) return UVH_E_STATE;
    if (!map) return UVH_E_ARG;
    uvh::build_csc_map(ctx->mesh, map);
    return UVH_OK;
}

extern "C++" {
namespace {
bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Orders a call on `s` after the context's previous call (see uvh_ctx::done_ev).
hipError_t begin_call(uvh_ctx *ctx, hipStream_t s) {
    if (ctx->inflight && s != ctx->last_stream && !stream_capturing(s)) return hipStreamWaitEvent(s, ctx->done_ev, 0);
    return hipSuccess;
}

hipError_t end_call(uvh_ctx *ctx, hipStream_t s) {
    ctx->last_stream = s;
    if (stream_capturing(s)) {   // the graph's replays are ordered by whoever launches them
        ctx->inflight = false;
        return hipSuccess;
    }
    ctx->inflight = true;
    return hipEventRecord(ctx->done_ev, s);
}

int assemble_device_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                         double *dK, double *dT, double *dF, void *stream);
int assemble_noalloc_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                          double *dK, double *dT, double *dF, void *stream);
int uv_assemble_device_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc, const uvh_uv_fields *df,
                            uint32_t flags, double *dR, double *dK, double *dT, double *dF, void *stream);

template <class F>
int ordered_call(uvh_ctx *ctx, void *stream, const char *what, F &&body) {
    if (!ctx) return UVH_E_ARG;
    hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = begin_call(ctx, s);
    if (e != hipSuccess) return ctx->hip_fail(e, what);
    const int st = body();
    if (st != UVH_OK) return st;
    e = end_call(ctx, s);
    if (e != hipSuccess) return ctx->hip_fail(e, what);
    return UVH_OK;
}
}  // namespace
}  // extern "C++"

int uvh_assemble_device(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                        double *dK, double *dT, double *dF, void *stream) {
    return ordered_call(ctx, stream, "uvh_assemble_device", [&] {
        ctx->last_status = ctx->d_status;
        return assemble_device_impl(ctx, ctrl, df, flags, dR, dK, dT, dF, stream);
    });
}

int uvh_assemble_device_noalloc(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                                double *dK, double *dT, double *dF, void *stream) {
    return ordered_call(ctx, stream, "uvh_assemble_device_noalloc", [&] {
        return assemble_noalloc_impl(ctx, ctrl, df, flags, dR, dK, dT, dF, stream);
    });
}

int uvh_uv_assemble_device(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc, const uvh_uv_fields *df,
                           uint32_t flags, double *dR, double *dK, double *dT, double *dF, void *stream) {
    return ordered_call(ctx, stream, "uvh_uv_assemble_device", [&] {
        ctx->last_status = ctx->d_status;
        return uv_assemble_device_impl(ctx, ctrl, uvc, df, flags, dR, dK, dT, dF, stream);
    });
}

int uvh_status_async(uvh_ctx *ctx, int32_t *dst, void *stream) {
    if (!ctx) return UVH_E_ARG;
    if (!dst) return ctx->fail(UVH_E_ARG, "uvh_status_async: null destination");
    hipSetDevice(ctx->device);
    const hipError_t e = hipMemcpyAsync(dst, ctx->last_status ? ctx->last_status : ctx->d_status, sizeof(int32_t),
                                        hipMemcpyDefault, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_status_async");
    return UVH_OK;
}

extern "C++" {
namespace {
int assemble_device_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                         double *dK, double *dT, double *dF, void *stream) {
    if (!ctx->has_mesh) return ctx->fail(UVH_E_STATE, "uvh_assemble: no mesh (call uvh_set_mesh first)");
    if (!ctrl || !df || !dR) return ctx->fail(UVH_E_ARG, "uvh_assemble: null ctrl, fields or R");
    if (!fields_complete(*df)) return ctx->fail(UVH_E_ARG, "uvh_assemble: a required field pointer is null");
    {
        std::string why;
        const int st = check_law(*ctrl, df->q, df->muk, df->V0, why);
        if (st != UVH_OK) return ctx->fail(st, "uvh_assemble: " + why);
    }
    const bool fp32 = (flags & UVH_FP32) != 0;
    if (ctrl->supg_tau < 0 || ctrl->supg_tau > 4) return ctx->fail(UVH_E_ARG, "uvh_assemble: supg_tau out of range");
    const bool ronly = (flags & UVH_RONLY) != 0;
    if (!ronly && !dK) return ctx->fail(UVH_E_ARG, "uvh_assemble: K output required unless UVH_RONLY");
    hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL is the HIP null stream
    const uvh::MeshHost &m = ctx->mesh;
    // FP64 R+K and R on the fused kernel where the mesh has its schedule (R alone then sums in
    // the same order as R of R+K); the precision form on the two-pass kernels
    const bool fused = ctx->fused_ok && !fp32;
    if (!fused && !ronly && !ctx->d_EB) {
        // FP64 rows; the FP32 rows of the precision form fit in the same buffer
        const size_t stride = uvh::element_row_stride(m.nod, false);
        if (dalloc(&ctx->d_EB, size_t(m.nele) * m.nod * stride) != hipSuccess)
            return ctx->fail(UVH_E_NOMEM, "uvh_assemble: cannot allocate element rows");
    }
    const uvh::Params p = make_params(*ctrl, flags);
    const uvh::FieldPtrs f = to_ptrs(*df);
    const int check = (flags & UVH_NO_NAN_CHECK) ? 0 : 1;
    // the fused kernel reads b from a nodal pre-pass; the two-pass element kernel computes b in
    // its gathers and clears the NaN flags itself
    hipError_t e = hipSuccess;
    if (fused)
        e = ctx->timed("nodal_b", s, [&] { return uvh::launch_nodal_b(m.nn, f, p, ctx->d_b, ctx->d_status, s); });
    if (e == hipSuccess && fused) {
        uvh::CallArgs ca{};
        ca.f = f; ca.p = p; ca.R = dR; ca.Tint = dT; ca.Fext = dF; ca.Kval = ronly ? nullptr : dK;
        int slot = ctx->last_slot;
        if (slot < 0 || std::memcmp(&ca, &ctx->last_args, sizeof(ca)) != 0) {
            slot = ctx->ring_pos;
            ctx->ring_pos = (ctx->ring_pos + 1) % uvh_ctx::kRing;
            if (ctx->ring_used[slot]) hipEventSynchronize(ctx->ring_ev[slot]);
            ctx->h_cargs[slot] = ca;
            e = hipMemcpyAsync(ctx->d_cargs + slot, ctx->h_cargs + slot, sizeof(ca), hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipEventRecord(ctx->ring_ev[slot], s);
            ctx->ring_used[slot] = true;
            ctx->last_args = ca;
            ctx->last_slot = slot;
        }
        if (e == hipSuccess)
            e = ctx->timed(ronly ? "fused_r" : "fused_rk", s, [&] {
                return uvh::launch_fused(m.nod, m.nip, ctx->sched.npatch, ctx->layout, ctx->d_margs,
                                         ctx->d_cargs + slot, m.nnzN(), m.nn, p.law != 0, ronly, check, s);
            });
        if (e != hipSuccess) return ctx->hip_fail(e, "uvh_assemble: kernel launch");
        return UVH_OK;
    }
    const bool full = ctx->rows_mode == 1 || (ctx->rows_mode == 0 && fp32);
    const int t = fp32 ? 1 : 0;
    const char *el_name = fp32 ? (ronly ? "element32_r" : "element32_rk") : (ronly ? "element_r" : "element_rk");
    const char *row_name = fp32 ? (ronly ? "rows32_r" : "rows32_rk") : (ronly ? "rows_r" : "rows_rk");
    auto elements = [&](int lb0, int lb1, hipStream_t st) {
        return ctx->timed(el_name, st, [&] {
            return uvh::launch_element(m.nod, m.nele, m.nip, ronly ? ctx->epb_wide[t] : ctx->epb[t],
                                       ctx->el_threads[t], fp32, full, ctx->d_conn, ctx->d_x, ctx->d_y,
                                       f, nullptr, ctx->d_tab, ctx->d_tab32, p, ctx->d_inc_pos, ronly, ctx->d_EB,
                                       ctx->d_RE, lb0, lb1, ctx->d_status, st);
        });
    };
    auto rows = [&](const int32_t *blocks, int nblk, hipStream_t st) {
        return ctx->timed(row_name, st, [&] {
            // the gather form has one thread per node-graph entry, so a node in no element
            // (no entries) would keep a stale R: such meshes use the scatter form
            if (ctx->gather_rows && !ronly && !m.has_isolated)
                return uvh::launch_rows_gather(m.nod, m.nn, m.nnzN(), fp32, ctx->d_nzrow, ctx->d_rowptr_n,
                                               ctx->d_inc_ptr, ctx->d_imap_ptr, ctx->d_imap, ctx->d_diag, ctx->d_EB,
                                               ctx->d_RE, dR, dT, dF, dK, check, ctx->d_status, st);
            return uvh::launch_rows(m.nod, m.nn, m.max_block_deg, fp32, ctx->d_inc_ptr, ctx->d_pos_inc, ctx->d_rowptr_n,
                                    ctx->d_diag, m.nnzN(), ctx->d_EB, ctx->d_RE, dR, dT, dF, ronly ? nullptr : dK,
                                    check, ctx->d_status, blocks, nblk, st);
        });
    };
    if (ronly || ctx->nchunks <= 1 || ctx->gather_rows) {
        if (e == hipSuccess) e = elements(0, -1, s);
        if (e == hipSuccess) e = rows(nullptr, 0, s);
    } else {
        if (!ctx->plan_ok[t]) {
            const int st = build_plan(ctx, t);
            if (st != UVH_OK) return st;
        }
        const std::vector<int> &tiles = ctx->plan_tiles[t], &off = ctx->plan_boff[t];
        const int C = int(tiles.size()) - 1;
        hipStream_t a = ctx->aux_stream;
        if (e == hipSuccess) e = hipEventRecord(ctx->fork_ev, s);   // status cleared and b computed
        if (e == hipSuccess) e = hipStreamWaitEvent(a, ctx->fork_ev, 0);
        for (int c = 0; c < C && e == hipSuccess; ++c) {
            e = elements(tiles[c], tiles[c + 1], s);
            if (e == hipSuccess) e = hipEventRecord(ctx->chunk_ev[c], s);
            if (e == hipSuccess) e = hipStreamWaitEvent(a, ctx->chunk_ev[c], 0);
            if (e == hipSuccess && off[c + 1] > off[c]) e = rows(ctx->d_plan_blocks[t] + off[c], off[c + 1] - off[c], a);
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->join_ev, a);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->join_ev, 0);   // the caller's stream sees the end
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_assemble: kernel launch");
    return UVH_OK;
}

// The two-pass assembly with its work buffers in the scratch array: no allocation, no
// profiling events; the same kernels, launch shapes and summation order as the default path.
int assemble_noalloc_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *df, uint32_t flags, double *dR,
                          double *dK, double *dT, double *dF, void *stream) {
    if (!ctx->has_mesh) return ctx->fail(UVH_E_STATE, "uvh_assemble_device_noalloc: no mesh (call uvh_set_mesh first)");
    if (!ctrl || !df || !dR) return ctx->fail(UVH_E_ARG, "uvh_assemble_device_noalloc: null ctrl, fields or R");
    if (!fields_complete(*df)) return ctx->fail(UVH_E_ARG, "uvh_assemble_device_noalloc: a required field pointer is null");
    {
        std::string why;
        const int st = check_law(*ctrl, df->q, df->muk, df->V0, why);
        if (st != UVH_OK) return ctx->fail(st, "uvh_assemble_device_noalloc: " + why);
    }
    if (ctrl->supg_tau < 0 || ctrl->supg_tau > 4)
        return ctx->fail(UVH_E_ARG, "uvh_assemble_device_noalloc: supg_tau out of range");
    const bool fp32 = (flags & UVH_FP32) != 0, ronly = (flags & UVH_RONLY) != 0;
    if (!ronly && !dK) return ctx->fail(UVH_E_ARG, "uvh_assemble_device_noalloc: K output required unless UVH_RONLY");
    if (ctx->scratch_rc != UVH_OK || !g_scratch_ptr[ctx->device])
        return ctx->fail(UVH_E_HIP, "uvh_assemble_device_noalloc: the scratch array's address was not resolved");
    const uvh::MeshHost &m = ctx->mesh;
    const NoallocLayout L = noalloc_layout(m, flags);
    if (L.total > kScratchBytes)
        return ctx->fail(UVH_E_ARG, "uvh_assemble_device_noalloc: this mesh needs " + std::to_string(L.total) +
                                        " bytes of work buffers; the allocation-free entry point holds at most " +
                                        std::to_string(kScratchBytes) + " (use uvh_assemble_device)");
    unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[ctx->device]);
    int *status = reinterpret_cast<int *>(base + L.status);
    double *b = reinterpret_cast<double *>(base + L.b);
    double *RE = reinterpret_cast<double *>(base + L.RE);
    double *EB = ronly ? nullptr : reinterpret_cast<double *>(base + L.EB);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uvh::Params p = make_params(*ctrl, flags);
    const uvh::FieldPtrs f = to_ptrs(*df);
    const int check = (flags & UVH_NO_NAN_CHECK) ? 0 : 1;
    const bool full = ctx->rows_mode == 1 || (ctx->rows_mode == 0 && fp32);
    const int t = fp32 ? 1 : 0;
    ctx->last_status = status;
    (void)b;   // b is computed in the element kernel's gathers
    hipError_t e = uvh::launch_element(m.nod, m.nele, m.nip, ronly ? ctx->epb_wide[t] : ctx->epb[t],
                                       ctx->el_threads[t], fp32, full, ctx->d_conn,
                                       ctx->d_x, ctx->d_y, f, nullptr,
                                       ctx->d_tab, ctx->d_tab32, p, ctx->d_inc_pos, ronly, EB, RE, 0, -1, status, s);
    if (e == hipSuccess)
        e = uvh::launch_rows(m.nod, m.nn, m.max_block_deg, fp32, ctx->d_inc_ptr, ctx->d_pos_inc, ctx->d_rowptr_n,
                             ctx->d_diag, m.nnzN(), EB, RE, dR, dT, dF, ronly ? nullptr : dK, check, status, nullptr, 0, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_assemble_device_noalloc: kernel launch");
    return UVH_OK;
}

}  // namespace
}  // extern "C++"

int uvh_sync_status(uvh_ctx *ctx) {
    if (!ctx) return UVH_E_ARG;
    hipSetDevice(ctx->device);
    // waits for the context's last assembly only (not for the device)
    hipError_t e = ctx->inflight ? hipEventSynchronize(ctx->done_ev) : hipSuccess;
    const int *src = ctx->last_status ? ctx->last_status : ctx->d_status;
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->h_status, src, sizeof(int), hipMemcpyDeviceToHost, ctx->own_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->own_stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_sync_status");
    const int st = *ctx->h_status;
    if (st & 2) return ctx->fail(UVH_E_NAN_K, " NaN in K ");
    if (st & 1) return ctx->fail(UVH_E_NAN_R, "uvhMatrixAssembly:NaNinR  NaN in R ");
    return UVH_OK;
}

int uvh_assemble(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *hf, uint32_t flags, double *R, double *Kval,
                 double *Tint, double *Fext) {
    if (!ctx) return UVH_E_ARG;
    if (!ctx->has_mesh) return ctx->fail(UVH_E_STATE, "uvh_assemble: no mesh (call uvh_set_mesh first)");
    if (!ctrl || !hf || !R) return ctx->fail(UVH_E_ARG, "uvh_assemble: null ctrl, fields or R");
    if (!fields_complete(*hf)) return ctx->fail(UVH_E_ARG, "uvh_assemble: a required field pointer is null");
    const bool ronly = (flags & UVH_RONLY) != 0;
    hipSetDevice(ctx->device);
    const uvh::MeshHost &m = ctx->mesh;
    const size_t nn = size_t(m.nn);
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    if (!ctx->d_fields) chk(dalloc(&ctx->d_fields, kHostFields * nn));
    if (!ctx->d_R) chk(dalloc(&ctx->d_R, 3 * nn));
    if (!ctx->d_T) chk(dalloc(&ctx->d_T, 3 * nn));
    if (!ctx->d_F) chk(dalloc(&ctx->d_F, 3 * nn));
    if (!ronly && Kval && !ctx->d_K) chk(dalloc(&ctx->d_K, size_t(9) * m.nnzN()));
    if (e != hipSuccess) return ctx->fail(UVH_E_NOMEM, "uvh_assemble: device allocation failed");
    const double *src[kHostFields] = {hf->h, hf->ub, hf->vb, hf->S, hf->B, hf->rho, hf->C, hf->m, hf->n, hf->AGlen,
                                      hf->as, hf->ab, hf->h0, hf->ub0, hf->vb0, hf->as0, hf->ab0, hf->dadh,
                                      hf->q, hf->muk, hf->V0};
    const double *dst[kHostFields];
    hipStream_t s = ctx->own_stream;
    for (int k = 0; k < kHostFields; ++k) {
        double *d = ctx->d_fields + k * nn;
        dst[k] = src[k] ? d : nullptr;
        if (src[k]) chk(hipMemcpyAsync(d, src[k], nn * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_assemble: field upload");
    uvh_fields df{dst[0],  dst[1],  dst[2],  dst[3],  dst[4],  dst[5],  dst[6],  dst[7],  dst[8],  dst[9],
                  dst[10], dst[11], dst[12], dst[13], dst[14], dst[15], dst[16], dst[17], dst[18], dst[19],
                  dst[20]};
    double *dK = (!ronly && Kval) ? ctx->d_K : nullptr;
    if (!ronly && !Kval) flags |= UVH_RONLY;
    int st = uvh_assemble_device(ctx, ctrl, &df, flags, ctx->d_R, dK, ctx->d_T, ctx->d_F, s);
    if (st != UVH_OK) return st;
    chk(hipMemcpyAsync(R, ctx->d_R, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (Tint) chk(hipMemcpyAsync(Tint, ctx->d_T, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (Fext) chk(hipMemcpyAsync(Fext, ctx->d_F, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (dK) chk(hipMemcpyAsync(Kval, dK, size_t(9) * m.nnzN() * sizeof(double), hipMemcpyDeviceToHost, s));
    chk(hipStreamSynchronize(s));
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_assemble: download");
    if (flags & UVH_NO_NAN_CHECK) return UVH_OK;
    return uvh_sync_status(ctx);
}

extern "C++" {
namespace {
int uv_assemble_device_impl(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc, const uvh_uv_fields *df,
                            uint32_t flags, double *dR, double *dK, double *dT, double *dF, void *stream) {
    if (!ctx->has_mesh) return ctx->fail(UVH_E_STATE, "uvh_uv_assemble: no mesh (call uvh_set_mesh first)");
    if (!ctrl || !uvc || !df || !dR) return ctx->fail(UVH_E_ARG, "uvh_uv_assemble: null ctrl, fields or R");
    for (const double *q : {df->h, df->s, df->ub, df->vb, df->S, df->B, df->rho, df->C, df->m, df->n, df->AGlen})
        if (!q) return ctx->fail(UVH_E_ARG, "uvh_uv_assemble: a required field pointer is null");
    if (flags & UVH_ZEROFIELDS)
        return ctx->fail(UVH_E_UNSUPPORTED, "uvh_uv_assemble: ZeroFields (pass the zeroed velocities instead)");
    {
        std::string why;
        const int st = check_law(*ctrl, df->q, df->muk, df->V0, why);
        if (st != UVH_OK) return ctx->fail(st, "uvh_uv_assemble: " + why);
    }
    const bool fp32 = (flags & UVH_FP32) != 0, ronly = (flags & UVH_RONLY) != 0;
    if (!ronly && !dK) return ctx->fail(UVH_E_ARG, "uvh_uv_assemble: K output required unless UVH_RONLY");
    hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uvh::MeshHost &m = ctx->mesh;
    if (!ronly && !ctx->d_EB &&
        dalloc(&ctx->d_EB, size_t(m.nele) * m.nod * uvh::element_row_stride(m.nod, false)) != hipSuccess)
        return ctx->fail(UVH_E_NOMEM, "uvh_uv_assemble: cannot allocate element rows");
    if (!ctx->d_zero) {   // stands in for the thickness-equation fields the element gather reads
        if (dalloc(&ctx->d_zero, size_t(m.nn)) != hipSuccess ||
            hipMemset(ctx->d_zero, 0, size_t(m.nn) * sizeof(double)) != hipSuccess)
            return ctx->fail(UVH_E_NOMEM, "uvh_uv_assemble: cannot allocate the zero field");
    }
    uvh::Params p = make_params(*ctrl, flags & ~uint32_t(UVH_ZEROFIELDS));
    p.uv = 1;
    p.reset_thickness = 0;
    p.enforce_b_above = 0;
    p.cmin = uvc->cmin;
    p.aglenmin = uvc->aglenmin;
    p.dvisk = uvc->nrviscosity;
    const double *z = ctx->d_zero;
    const uvh::FieldPtrs f{df->h, df->ub, df->vb, df->S, df->B, df->rho, df->C, df->m, df->n, df->AGlen, z, z,
                           df->h, df->ub, df->vb, z, z, nullptr, df->q, df->muk, df->V0};
    const int check = (flags & UVH_NO_NAN_CHECK) ? 0 : 1;
    const bool full = ctx->rows_mode == 1 || (ctx->rows_mode == 0 && fp32);
    const int t = fp32 ? 1 : 0;
    hipError_t e = ctx->timed("nodal_s", s, [&] { return uvh::launch_nodal_s(m.nn, df->s, df->h, ctx->d_b, ctx->d_status, s); });
    if (e == hipSuccess)
        e = ctx->timed(fp32 ? (ronly ? "uv_element32_r" : "uv_element32_rk") : (ronly ? "uv_element_r" : "uv_element_rk"),
                       s, [&] {
                           return uvh::launch_element(m.nod, m.nele, m.nip, ctx->epb_wide[t], ctx->el_threads[t], fp32,
                                                      full, ctx->d_conn,
                                                      ctx->d_x, ctx->d_y, f, ctx->d_b, ctx->d_tab, ctx->d_tab32, p,
                                                      ctx->d_inc_pos, ronly, ctx->d_EB, ctx->d_RE, 0, -1, nullptr, s);
                       });
    if (e == hipSuccess)
        e = ctx->timed(fp32 ? (ronly ? "uv_rows32_r" : "uv_rows32_rk") : (ronly ? "uv_rows_r" : "uv_rows_rk"), s, [&] {
            return uvh::launch_rows(m.nod, m.nn, m.max_block_deg, fp32, ctx->d_inc_ptr, ctx->d_pos_inc,
                                    ctx->d_rowptr_n, ctx->d_diag, m.nnzN(), ctx->d_EB, ctx->d_RE, dR, dT, dF,
                                    ronly ? nullptr : dK, check, ctx->d_status, nullptr, 0, s, 2);
        });
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_uv_assemble: kernel launch");
    return UVH_OK;
}
}  // namespace
}  // extern "C++"

int uvh_uv_assemble(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc, const uvh_uv_fields *hf, uint32_t flags,
                    double *R, double *Kval, double *Tint, double *Fext) {
    if (!ctx) return UVH_E_ARG;
    if (!ctx->has_mesh) return ctx->fail(UVH_E_STATE, "uvh_uv_assemble: no mesh (call uvh_set_mesh first)");
    if (!ctrl || !uvc || !hf || !R) return ctx->fail(UVH_E_ARG, "uvh_uv_assemble: null ctrl, fields or R");
    const double *src[11] = {hf->h, hf->s, hf->ub, hf->vb, hf->S, hf->B, hf->rho, hf->C, hf->m, hf->n, hf->AGlen};
    for (auto q : src)
        if (!q) return ctx->fail(UVH_E_ARG, "uvh_uv_assemble: a required field pointer is null");
    const bool ronly = (flags & UVH_RONLY) != 0;
    hipSetDevice(ctx->device);
    const uvh::MeshHost &m = ctx->mesh;
    const size_t nn = size_t(m.nn);
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    if (!ctx->d_fields) chk(dalloc(&ctx->d_fields, kHostFields * nn));
    if (!ctx->d_R) chk(dalloc(&ctx->d_R, 3 * nn));
    if (!ctx->d_T) chk(dalloc(&ctx->d_T, 3 * nn));
    if (!ctx->d_F) chk(dalloc(&ctx->d_F, 3 * nn));
    if (!ronly && Kval && !ctx->d_K) chk(dalloc(&ctx->d_K, size_t(9) * m.nnzN()));
    if (e != hipSuccess) return ctx->fail(UVH_E_NOMEM, "uvh_uv_assemble: device allocation failed");
    hipStream_t s = ctx->own_stream;
    const double *all[14] = {src[0], src[1], src[2], src[3], src[4], src[5], src[6], src[7], src[8], src[9], src[10],
                             hf->q, hf->muk, hf->V0};
    const double *dst[14];
    for (int k = 0; k < 14; ++k) {
        double *d = ctx->d_fields + k * nn;
        dst[k] = all[k] ? d : nullptr;
        if (all[k]) chk(hipMemcpyAsync(d, all[k], nn * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_uv_assemble: field upload");
    const uvh_uv_fields df{dst[0], dst[1], dst[2],  dst[3],  dst[4],  dst[5],  dst[6],
                           dst[7], dst[8], dst[9], dst[10], dst[11], dst[12], dst[13]};
    double *dK = (!ronly && Kval) ? ctx->d_K : nullptr;
    if (!ronly && !Kval) flags |= UVH_RONLY;
    int st = uvh_uv_assemble_device(ctx, ctrl, uvc, &df, flags, ctx->d_R, dK, ctx->d_T, ctx->d_F, s);
    if (st != UVH_OK) return st;
    chk(hipMemcpyAsync(R, ctx->d_R, 2 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (Tint) chk(hipMemcpyAsync(Tint, ctx->d_T, 2 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (Fext) chk(hipMemcpyAsync(Fext, ctx->d_F, 2 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (dK) chk(hipMemcpyAsync(Kval, dK, size_t(4) * m.nnzN() * sizeof(double), hipMemcpyDeviceToHost, s));
    chk(hipStreamSynchronize(s));
    if (e != hipSuccess) return ctx->hip_fail(e, "uvh_uv_assemble: download");
    if (flags & UVH_NO_NAN_CHECK) return UVH_OK;
    return uvh_sync_status(ctx);
}

int uvh_profile(uvh_ctx *ctx, int enable) {
    if (!ctx) return UVH_E_ARG;
    hipSetDevice(ctx->device);
    ctx->resolve_timers();
    ctx->profiling = enable != 0;
    for (auto &t : ctx->timers) { t.total_ms = 0; t.launches = 0; }
    return UVH_OK;
}

int uvh_kernel_times(uvh_ctx *ctx, const char **names, double *total_ms, int64_t *launches, int cap) {
    if (!ctx) return UVH_E_ARG;
    hipSetDevice(ctx->device);
    ctx->resolve_timers();
    int n = 0;
    for (auto &t : ctx->timers) {
        if (n >= cap) break;
        if (names) names[n] = t.name;
        if (total_ms) total_ms[n] = t.total_ms;
        if (launches) launches[n] = t.launches;
        ++n;
    }
    return n;
}

}  // extern "C"

// ------------------------------------------------------------------ batched assembly
struct uvh_batch {
    int device = 0;
    std::string err;
    std::vector<uvh_ctx *> members;
    // work lists for the FP64 [0] and FP32 [1] element tiles
    bool built[2] = {false, false};
    int2 *d_items[2][3] = {};   // nodal, element, row items (element and row items grouped by type)
    int n_items[2][3] = {};
    int ebeg[2][4] = {}, rbeg[2][4] = {};  // per element type (3, 6, 10 nodes): item ranges
    size_t elem_lds[2][3] = {}, row_lds[2][3] = {};
    uvh::BatchMember *d_mem = nullptr, *h_mem = nullptr;  // device array and its pinned host copy
    std::vector<uvh::BatchMember> last;
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false;
    int *d_ctr = nullptr, *d_status = nullptr;
    // one workgroup per work item (default), or persistent workgroups taking items from a
    // counter (UVH_BATCH=persistent): the persistent loop keeps more registers live (227 vs
    // 146 VGPRs for the FP64 element kernel) and runs 1.3-1.5x slower
    bool flat = true;
    // batched residual cost (uvhb_cost_device): member table, (member, chunk) items, partials
    uvh::CostMember *d_cost = nullptr, *h_cost = nullptr;
    std::vector<uvh::CostMember> cost_last;
    int2 *d_cost_items = nullptr;
    int n_cost_items = 0;
    double2 *d_cost_partial = nullptr;
    hipEvent_t cost_upload = nullptr;
    bool cost_pending = false;
    int fail(int code, const std::string &m) { err = m; return code; }
    void free_lists() {
        for (int t = 0; t < 2; ++t) {
            for (int k = 0; k < 3; ++k) { hipFree(d_items[t][k]); d_items[t][k] = nullptr; n_items[t][k] = 0; }
            built[t] = false;
        }
        // the cost tables depend on the members too
        if (cost_pending) hipEventSynchronize(cost_upload);
        hipFree(d_cost); hipFree(d_cost_items); hipFree(d_cost_partial); hipHostFree(h_cost);
        d_cost = nullptr; d_cost_items = nullptr; d_cost_partial = nullptr; h_cost = nullptr;
        if (cost_upload) hipEventDestroy(cost_upload);
        cost_upload = nullptr;
        cost_pending = false;
        cost_last.clear();
        n_cost_items = 0;
    }
};

namespace {

// Work lists of one element tile choice: (member, block) pairs; element and row items grouped
// by element type (3, 6, 10 nodes), members in order within a type.
int batch_build_lists(uvh_batch *b, int t) {
    const bool fp32 = t == 1;
    std::vector<int2> items[3];
    const int nods[3] = {3, 6, 10};
    for (int k = 0; k < 3; ++k) { b->elem_lds[t][k] = 0; b->row_lds[t][k] = 0; }
    for (int mi = 0; mi < int(b->members.size()); ++mi) {
        const uvh::MeshHost &m = b->members[mi]->mesh;
        for (int k = 0; k < (m.nn + 255) / 256; ++k) items[0].push_back(make_int2(mi, k));
    }
    for (int ty = 0; ty < 3; ++ty) {
        b->ebeg[t][ty] = int(items[1].size());
        b->rbeg[t][ty] = int(items[2].size());
        for (int mi = 0; mi < int(b->members.size()); ++mi) {
            const uvh_ctx *c = b->members[mi];
            const uvh::MeshHost &m = c->mesh;
            if (m.nod != nods[ty]) continue;
            const int epb = c->bepb[t];
            for (int k = 0; k < (m.nele + epb - 1) / epb; ++k) items[1].push_back(make_int2(mi, k));
            for (int k = 0; k < (m.nn + uvh::row_nodes_per_block(m.nod) - 1) / uvh::row_nodes_per_block(m.nod); ++k)
                items[2].push_back(make_int2(mi, k));
            b->elem_lds[t][ty] = std::max(b->elem_lds[t][ty], uvh::element_lds_bytes(m.nod, m.nip, epb, fp32));
            b->row_lds[t][ty] = std::max(b->row_lds[t][ty], size_t(9) * size_t(m.max_block_deg) * sizeof(double));
        }
    }
    b->ebeg[t][3] = int(items[1].size());
    b->rbeg[t][3] = int(items[2].size());
    for (int k = 0; k < 3; ++k) {
        hipFree(b->d_items[t][k]);
        b->d_items[t][k] = nullptr;
        b->n_items[t][k] = int(items[k].size());
        if (items[k].empty()) continue;
        if (dalloc(&b->d_items[t][k], items[k].size()) != hipSuccess ||
            hipMemcpy(b->d_items[t][k], items[k].data(), items[k].size() * sizeof(int2), hipMemcpyHostToDevice) !=
                hipSuccess)
            return b->fail(UVH_E_NOMEM, "uvhb: cannot upload work lists");
    }
    b->built[t] = true;
    return UVH_OK;
}

}  // namespace

extern "C" {

int uvhb_create(uvh_batch **out, int device) {
    if (!out) return UVH_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return UVH_E_HIP;
    hipSetDevice(device);
    auto *b = new uvh_batch();
    b->device = device;
    const char *bm = std::getenv("UVH_BATCH");
    b->flat = !(bm && std::strcmp(bm, "persistent") == 0);
    if (hipMalloc(reinterpret_cast<void **>(&b->d_ctr), 6 * sizeof(int)) != hipSuccess ||
        hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return UVH_E_HIP;
    }
    *out = b;
    return UVH_OK;
}

void uvhb_destroy(uvh_batch *b) {
    if (!b) return;
    hipSetDevice(b->device);
    hipDeviceSynchronize();
    b->free_lists();
    for (auto *c : b->members) uvh_destroy(c);
    hipFree(b->d_mem);
    hipHostFree(b->h_mem);
    hipFree(b->d_ctr);
    hipFree(b->d_status);
    if (b->upload_done) hipEventDestroy(b->upload_done);
    delete b;
}

const char *uvhb_last_error(const uvh_batch *b) { return b ? b->err.c_str() : "null batch"; }

int uvhb_add_mesh(uvh_batch *b, int32_t nele, int32_t nod, int32_t nnodes, const int32_t *conn, const double *x,
                  const double *y, int32_t nip, const double *pts, const double *wts, int32_t *member) {
    if (!b) return UVH_E_ARG;
    uvh_ctx *c = nullptr;
    int st = uvh_create(&c, b->device);
    if (st != UVH_OK) return b->fail(st, "uvhb_add_mesh: cannot create a member context");
    c->fused_wanted = false;   // the batched kernels are the two-pass ones
    st = uvh_set_mesh(c, nele, nod, nnodes, conn, x, y, nip, pts, wts);
    if (st != UVH_OK) {
        b->err = std::string("uvhb_add_mesh: ") + c->err;
        uvh_destroy(c);
        return st;
    }
    const uvh::MeshHost &m = c->mesh;
    if (dalloc(&c->d_EB, size_t(m.nele) * m.nod * uvh::element_row_stride(m.nod, false)) != hipSuccess) {
        uvh_destroy(c);
        return b->fail(UVH_E_NOMEM, "uvhb_add_mesh: cannot allocate element rows");
    }
    if (member) *member = int32_t(b->members.size());
    b->members.push_back(c);
    b->free_lists();
    hipFree(b->d_mem); b->d_mem = nullptr;
    hipHostFree(b->h_mem); b->h_mem = nullptr;
    hipFree(b->d_status); b->d_status = nullptr;
    b->last.clear();
    return UVH_OK;
}

int uvhb_size(const uvh_batch *b) { return b ? int(b->members.size()) : 0; }

uvh_ctx *uvhb_member(uvh_batch *b, int32_t member) {
    if (!b || member < 0 || member >= int(b->members.size())) return nullptr;
    return b->members[member];
}

int uvhb_set_tile(uvh_batch *b, int32_t nod, int32_t fp32, int32_t epb) {
    if (!b) return UVH_E_ARG;
    for (auto *c : b->members) {
        if (c->mesh.nod != nod) continue;
        if (epb < 1 || epb * c->mesh.nip > 256 || epb * nod > 256 ||
            uvh::element_lds_bytes(nod, c->mesh.nip, epb, fp32 != 0) > 156 * 1024)
            return b->fail(UVH_E_ARG, "uvhb_set_tile: tile does not fit (epb*nip, epb*nod <= 256, LDS <= 156 KiB)");
    }
    for (auto *c : b->members)
        if (c->mesh.nod == nod) c->bepb[fp32 ? 1 : 0] = epb;
    b->built[fp32 ? 1 : 0] = false;
    return UVH_OK;
}

int uvhb_get_tile(const uvh_batch *b, int32_t nod, int32_t fp32) {
    if (!b) return 0;
    for (auto *c : b->members)
        if (c->mesh.nod == nod) return c->bepb[fp32 ? 1 : 0];
    return 0;
}

int uvhb_assemble_device(uvh_batch *b, const uvh_ctrl *ctrl, const uvh_fields *dfields, uint32_t flags,
                         double *const *dR, double *const *dK, void *stream) {
    if (!b) return UVH_E_ARG;
    const int nm = int(b->members.size());
    if (nm == 0) return b->fail(UVH_E_STATE, "uvhb_assemble: empty batch");
    if (!ctrl || !dfields || !dR) return b->fail(UVH_E_ARG, "uvhb_assemble: null ctrl, fields or R");
    const bool ronly = (flags & UVH_RONLY) != 0, fp32 = (flags & UVH_FP32) != 0;
    const int t = fp32 ? 1 : 0;
    if (!ronly && !dK) return b->fail(UVH_E_ARG, "uvhb_assemble: K outputs required unless UVH_RONLY");
    for (int i = 0; i < nm; ++i) {
        if (!fields_complete(dfields[i]) || !dR[i] || (!ronly && !dK[i]))
            return b->fail(UVH_E_ARG, "uvhb_assemble: member " + std::to_string(i) + ": null field or output");
        std::string why;
        const int st = check_law(*ctrl, dfields[i].q, dfields[i].muk, dfields[i].V0, why);
        if (st != UVH_OK) return b->fail(st, "uvhb_assemble: member " + std::to_string(i) + ": " + why);
        if (ctrl->sliding_law != UVH_SLIDING_WEERTMAN && !b->flat)
            return b->fail(UVH_E_UNSUPPORTED,
                           "uvhb_assemble: the persistent batch schedule runs the Weertman law only");
    }
    hipSetDevice(b->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->d_mem) {
        if (dalloc(&b->d_mem, size_t(nm)) != hipSuccess || dalloc(&b->d_status, size_t(nm)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&b->h_mem), sizeof(uvh::BatchMember) * nm, 0) != hipSuccess)
            return b->fail(UVH_E_NOMEM, "uvhb_assemble: cannot allocate the member table");
    }
    if (!b->built[t]) {
        int st = batch_build_lists(b, t);
        if (st != UVH_OK) return st;
    }
    // member table: uploaded only when a pointer differs from the last upload
    std::vector<uvh::BatchMember> cur(nm);
    for (int i = 0; i < nm; ++i) {
        const uvh_ctx *c = b->members[i];
        const uvh::MeshHost &m = c->mesh;
        uvh::BatchMember &M = cur[i];
        std::memset(&M, 0, sizeof(M));
        M.conn = c->d_conn; M.x = c->d_x; M.y = c->d_y; M.tab = c->d_tab; M.tab32 = c->d_tab32;
        M.inc_ptr = c->d_inc_ptr; M.inc_pos = c->d_inc_pos; M.pos_inc = c->d_pos_inc; M.diag = c->d_diag;
        M.rowptr_n = c->d_rowptr_n; M.b = c->d_b; M.EB = c->d_EB; M.RE = c->d_RE;
        M.nnzN = m.nnzN(); M.nele = m.nele; M.nn = m.nn; M.nod = m.nod; M.nip = m.nip; M.epb = c->bepb[t];
        M.f = to_ptrs(dfields[i]);
        // K pointers are kept with UVH_RONLY too (the R-only kernels do not touch them), so that
        // R+K and R-only calls on the same buffers share one member table and a HIP graph
        // captured after eager calls of both replays with no table upload
        M.R = dR[i]; M.Tint = nullptr; M.Fext = nullptr; M.Kval = dK ? dK[i] : nullptr;
    }
    hipError_t e = hipSuccess;
    if (b->last.size() != cur.size() || std::memcmp(b->last.data(), cur.data(), sizeof(uvh::BatchMember) * nm) != 0) {
        if (b->upload_pending) hipEventSynchronize(b->upload_done);  // the pinned copy may still be in flight
        std::memcpy(b->h_mem, cur.data(), sizeof(uvh::BatchMember) * nm);
        e = hipMemcpyAsync(b->d_mem, b->h_mem, sizeof(uvh::BatchMember) * nm, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(b->upload_done, s);
        b->upload_pending = true;
        b->last = cur;
    }
    if (e != hipSuccess) return b->fail(UVH_E_HIP, std::string("uvhb_assemble: upload: ") + hipGetErrorString(e));
    const uvh::Params p = make_params(*ctrl, flags);
    const int check = (flags & UVH_NO_NAN_CHECK) ? 0 : 1;
    e = uvh::launch_batch(b->d_mem, nm, b->d_items[t][0], b->n_items[t][0], b->d_items[t][1], b->ebeg[t],
                          b->d_items[t][2], b->rbeg[t], b->elem_lds[t], b->row_lds[t], fp32, ronly, check,
                          b->d_ctr, b->d_status, p, b->flat, s);
    if (e != hipSuccess) return b->fail(UVH_E_HIP, std::string("uvhb_assemble: launch: ") + hipGetErrorString(e));
    return UVH_OK;
}

int uvhb_cost_device(uvh_batch *b, const double *const *dR, const double *const *dd, const double *const *dw,
                     double *rr, double *d2, void *stream) {
    if (!b) return UVH_E_ARG;
    const int nm = int(b->members.size());
    if (nm == 0) return b->fail(UVH_E_STATE, "uvhb_cost: empty batch");
    if (!dR || !dd || !rr || !d2) return b->fail(UVH_E_ARG, "uvhb_cost: null argument");
    hipSetDevice(b->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<uvh::CostMember> cur(static_cast<size_t>(nm));
    int first = 0;
    for (int i = 0; i < nm; ++i) {
        if (!dR[i] || !dd[i]) return b->fail(UVH_E_ARG, "uvhb_cost: member " + std::to_string(i) + ": null R or d");
        const int64_t n = 3 * int64_t(b->members[i]->mesh.nn);
        const int nch = int((n + uvh::kCostChunk - 1) / uvh::kCostChunk);
        cur[size_t(i)] = uvh::CostMember{dR[i], dd[i], dw ? dw[i] : nullptr, n, first, nch};
        first += nch;
    }
    if (!b->d_cost) {   // first call: tables and partials (later calls allocate nothing)
        std::vector<int2> items;
        for (int i = 0; i < nm; ++i)
            for (int c = 0; c < cur[size_t(i)].nchunk; ++c) items.push_back(make_int2(i, c));
        if (dalloc(&b->d_cost, size_t(nm)) != hipSuccess || dalloc(&b->d_cost_items, std::max<size_t>(1, items.size())) != hipSuccess ||
            dalloc(&b->d_cost_partial, std::max<size_t>(1, items.size())) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&b->h_cost), sizeof(uvh::CostMember) * nm, 0) != hipSuccess ||
            hipEventCreateWithFlags(&b->cost_upload, hipEventDisableTiming) != hipSuccess)
            return b->fail(UVH_E_NOMEM, "uvhb_cost: cannot allocate its tables");
        if (!items.empty() &&
            hipMemcpy(b->d_cost_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
            return b->fail(UVH_E_HIP, "uvhb_cost: item upload");
        b->n_cost_items = int(items.size());
    }
    hipError_t e = hipSuccess;
    if (b->cost_last.size() != cur.size() || std::memcmp(b->cost_last.data(), cur.data(), sizeof(uvh::CostMember) * nm) != 0) {
        if (b->cost_pending) hipEventSynchronize(b->cost_upload);
        std::memcpy(b->h_cost, cur.data(), sizeof(uvh::CostMember) * nm);
        e = hipMemcpyAsync(b->d_cost, b->h_cost, sizeof(uvh::CostMember) * nm, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(b->cost_upload, s);
        b->cost_pending = true;
        b->cost_last = cur;
    }
    if (e == hipSuccess) e = uvh::launch_cost(b->d_cost, nm, b->d_cost_items, b->n_cost_items, b->d_cost_partial, rr, d2, s);
    if (e != hipSuccess) return b->fail(UVH_E_HIP, std::string("uvhb_cost: ") + hipGetErrorString(e));
    return UVH_OK;
}

int uvhb_sync_status(uvh_batch *b, int32_t *member_status) {
    if (!b) return UVH_E_ARG;
    hipSetDevice(b->device);
    const int nm = int(b->members.size());
    std::vector<int> st(nm, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && b->d_status) e = hipMemcpy(st.data(), b->d_status, nm * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return b->fail(UVH_E_HIP, std::string("uvhb_sync_status: ") + hipGetErrorString(e));
    int first = UVH_OK;
    for (int i = 0; i < nm; ++i) {
        const int code = (st[i] & 2) ? UVH_E_NAN_K : (st[i] & 1) ? UVH_E_NAN_R : UVH_OK;
        if (member_status) member_status[i] = code;
        if (code != UVH_OK && first == UVH_OK) {
            first = code;
            b->err = "member " + std::to_string(i) + (code == UVH_E_NAN_K ? ":  NaN in K " : ": uvhMatrixAssembly:NaNinR");
        }
    }
    return first;
}

}  // extern "C"

namespace uvh {
const MeshHost *ctx_mesh(const uvh_ctx *ctx) { return ctx && ctx->has_mesh ? &ctx->mesh : nullptr; }
int ctx_device(const uvh_ctx *ctx) { return ctx ? ctx->device : 0; }
}  // namespace uvh
