// Host-only entry points of the once-per-mesh set-up (node graph and fused-assembly
// schedule). They run without a GPU, so the schedule can be checked on any machine.
// Not part of the public ABI in include/uvh.h.
#include <cstring>
#include <string>

#include "uvh_internal.h"

namespace {
struct Built {
    uvh::MeshHost m;
    uvh::Schedule s;
};

template <class T>
void put(void *dst, const std::vector<T> &v) {
    if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}
}  // namespace

extern "C" {

void *uvhh_build(int32_t nele, int32_t nod, int32_t nn, const int32_t *conn, const double *x, const double *y,
                 int32_t epb, int64_t window, int32_t patch_elems, int32_t band_cells, char *err, int32_t errlen) {
    auto *b = new Built();
    b->m.nele = nele; b->m.nod = nod; b->m.nn = nn;
    b->m.conn.assign(conn, conn + size_t(nele) * nod);
    b->m.x.assign(x, x + nn);
    b->m.y.assign(y, y + nn);
    std::string msg = uvh::build_mesh_host(b->m);
    if (msg.empty() && epb > 0) {
        int pe = 0, bc = 0;
        uvh::fused_patch_shape(nod, pe, bc);
        msg = uvh::build_schedule(b->m, epb, window, patch_elems > 0 ? patch_elems : pe, band_cells > 0 ? band_cells : bc,
                                  b->s);
    }
    if (!msg.empty()) {
        if (err && errlen > 0) {
            std::strncpy(err, msg.c_str(), size_t(errlen - 1));
            err[errlen - 1] = 0;
        }
        delete b;
        return nullptr;
    }
    return b;
}

// Schedule tables in SchedPtrs order, after the mesh tables 0-4.
namespace {
const std::vector<int32_t> *sched_table(const uvh::Schedule &s, int32_t k) {
    const std::vector<int32_t> *t[] = {&s.patch_batch_ptr, &s.batch_elem, &s.item_ptr, &s.item_ea, &s.item_slot,
                                       &s.item_stage, &s.fix_ptr, &s.fix_node, &s.fix_slot, &s.fix_item_ptr,
                                       &s.fix_item, &s.zero_ptr, &s.zero_slot, &s.zero_len, &s.flush_ptr,
                                       &s.flush_node, &s.flush_slot};
    return (k >= 0 && k < int32_t(sizeof(t) / sizeof(t[0]))) ? t[k] : nullptr;
}
constexpr int32_t kSched0 = 5, kSchedN = 17;
}  // namespace

// Table sizes (element counts) and scalars; `which` as in uvhh_copy:
// 0 rowptr_n, 1 colind_n, 2 pos, 3 inc_ptr, 4 inc, 5..21 the schedule tables in SchedPtrs order,
// 30 npatch, 31 peak_window, 32 patch_work, 33 epb, 34 max_stage, 35 max_items, 36 diag (size)
int64_t uvhh_size(void *h, int32_t which) {
    auto *b = static_cast<Built *>(h);
    const uvh::MeshHost &m = b->m;
    const uvh::Schedule &s = b->s;
    if (which >= kSched0 && which < kSched0 + kSchedN) return int64_t(sched_table(s, which - kSched0)->size());
    switch (which) {
        case 0: return int64_t(m.rowptr_n.size());
        case 1: return int64_t(m.colind_n.size());
        case 2: return int64_t(m.pos.size());
        case 3: return int64_t(m.inc_ptr.size());
        case 4: return int64_t(m.inc.size());
        case 30: return s.npatch;
        case 31: return s.peak_window;
        case 32: return s.patch_work;
        case 33: return s.epb;
        case 34: return s.max_stage;
        case 35: return s.max_items;
        case 36: return int64_t(m.diag.size());
        default: return -1;
    }
}

void uvhh_copy(void *h, int32_t which, void *dst) {
    auto *b = static_cast<Built *>(h);
    const uvh::MeshHost &m = b->m;
    if (which >= kSched0 && which < kSched0 + kSchedN) {
        put(dst, *sched_table(b->s, which - kSched0));
        return;
    }
    switch (which) {
        case 0: put(dst, m.rowptr_n); break;
        case 1: put(dst, m.colind_n); break;
        case 2: put(dst, m.pos); break;
        case 3: put(dst, m.inc_ptr); break;
        case 4: put(dst, m.inc); break;
        case 36: put(dst, m.diag); break;
        default: break;
    }
}

// Default LDS layout of the fused kernel for (nod, nip): epb, stage items, window doubles.
void uvhh_fused_layout(int32_t nod, int32_t nip, int64_t *out3) {
    const uvh::FusedLayout L = uvh::fused_layout(nod, nip);
    out3[0] = L.epb;
    out3[1] = L.stage_items;
    out3[2] = L.window_doubles;
}

void uvhh_csc_map(void *h, int64_t *map) { uvh::build_csc_map(static_cast<Built *>(h)->m, map); }

void uvhh_free(void *h) { delete static_cast<Built *>(h); }

}  // extern "C"
