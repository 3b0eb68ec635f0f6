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
                 int32_t epb, int64_t window, int32_t target_patches, char *err, int32_t errlen) {
    auto *b = new Built();
    b->m.nele = nele; b->m.nod = nod; b->m.nn = nn;
    b->m.conn.assign(conn, conn + size_t(nele) * nod);
    b->m.x.assign(x, x + nn);
    b->m.y.assign(y, y + nn);
    std::string msg = uvh::build_mesh_host(b->m);
    if (msg.empty() && epb > 0) msg = uvh::build_schedule(b->m, epb, window, target_patches, b->s);
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

// Table sizes (element counts) and scalars; `which` as in uvhh_copy.
int64_t uvhh_size(void *h, int32_t which) {
    auto *b = static_cast<Built *>(h);
    const uvh::MeshHost &m = b->m;
    const uvh::Schedule &s = b->s;
    switch (which) {
        case 0: return int64_t(m.rowptr_n.size());
        case 1: return int64_t(m.colind_n.size());
        case 2: return int64_t(m.pos.size());
        case 3: return int64_t(m.inc_ptr.size());
        case 4: return int64_t(m.inc.size());
        case 5: return int64_t(s.patch_batch_ptr.size());
        case 6: return int64_t(s.batch_elem.size());
        case 7: return int64_t(s.ea_slot.size());
        case 8: return int64_t(s.ea_rank.size());
        case 9: return int64_t(s.batch_maxrank.size());
        case 10: return int64_t(s.zero_ptr.size());
        case 11: return int64_t(s.zero_slot.size());
        case 12: return int64_t(s.zero_len.size());
        case 13: return int64_t(s.flush_ptr.size());
        case 14: return int64_t(s.flush_node.size());
        case 15: return int64_t(s.flush_slot.size());
        case 16: return s.npatch;
        case 17: return s.peak_window;
        case 18: return s.patch_work;
        case 19: return s.epb;
        default: return -1;
    }
}

void uvhh_copy(void *h, int32_t which, void *dst) {
    auto *b = static_cast<Built *>(h);
    const uvh::MeshHost &m = b->m;
    const uvh::Schedule &s = b->s;
    switch (which) {
        case 0: put(dst, m.rowptr_n); break;
        case 1: put(dst, m.colind_n); break;
        case 2: put(dst, m.pos); break;
        case 3: put(dst, m.inc_ptr); break;
        case 4: put(dst, m.inc); break;
        case 5: put(dst, s.patch_batch_ptr); break;
        case 6: put(dst, s.batch_elem); break;
        case 7: put(dst, s.ea_slot); break;
        case 8: put(dst, s.ea_rank); break;
        case 9: put(dst, s.batch_maxrank); break;
        case 10: put(dst, s.zero_ptr); break;
        case 11: put(dst, s.zero_slot); break;
        case 12: put(dst, s.zero_len); break;
        case 13: put(dst, s.flush_ptr); break;
        case 14: put(dst, s.flush_node); break;
        case 15: put(dst, s.flush_slot); break;
        default: break;
    }
}

void uvhh_csc_map(void *h, int64_t *map) { uvh::build_csc_map(static_cast<Built *>(h)->m, map); }

void uvhh_free(void *h) { delete static_cast<Built *>(h); }

}  // extern "C"
