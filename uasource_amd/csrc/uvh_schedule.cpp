// Host-side, once-per-mesh schedule of the fused assembly kernel (k_patch).
//
// The mesh is cut into compact patches of elements (recursive coordinate bisection); each
// patch is one workgroup. A node's three K rows (x, y, h) and its three residual entries are
// owned by exactly one patch: the patch holding most of the node's incident elements. A patch
// processes every element that touches one of its nodes (its own elements plus a thin ring of
// neighbours' elements whose rows it needs), in a sweep order along the patch's longer side,
// in batches of `epb` elements. Owned rows live in an LDS window from the batch of their first
// element to the write-out after the batch of their last; the schedule gives every row a
// window slot (first-fit interval allocation), lists the row items of each batch ((element,
// local node) pairs of owned nodes), the rows to zero before a batch and to write out after it.
//
// Summation order. Two distinct nodes share at most two elements on a conforming mesh, so
// every K entry off the diagonal, and every entry of a node with at most two incident
// elements, is a sum of at most two terms, which floating-point addition gives the same in
// either order: these are added into the window in any order. The diagonal block and the
// residual entries of a node with three or more incident elements are summed in a fixed
// order: by batch, and within a batch, where the node has several items, by the items' order
// in the batch (their values are staged and a fixup adds their sum once per batch). Results
// are therefore bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "uvh_internal.h"

namespace uvh {

namespace {

struct Box {
    double x0, x1, y0, y1;
};

Box bbox(const std::vector<int32_t> &els, const std::vector<double> &cx, const std::vector<double> &cy) {
    Box b{1e300, -1e300, 1e300, -1e300};
    for (int32_t e : els) {
        b.x0 = std::min(b.x0, cx[e]); b.x1 = std::max(b.x1, cx[e]);
        b.y0 = std::min(b.y0, cy[e]); b.y1 = std::max(b.y1, cy[e]);
    }
    return b;
}

// Splits `els` in two equal halves across the longer side of their centroid box (the cut runs
// along the sweep direction, so each half keeps the length and has half the front).
void bisect_across(std::vector<int32_t> &els, const std::vector<double> &cx, const std::vector<double> &cy,
                   std::vector<int32_t> &lo, std::vector<int32_t> &hi) {
    const Box b = bbox(els, cx, cy);
    const bool along_x = (b.x1 - b.x0) >= (b.y1 - b.y0);
    const std::vector<double> &key = along_x ? cy : cx;
    const std::vector<double> &key2 = along_x ? cx : cy;
    const size_t half = els.size() / 2;
    std::nth_element(els.begin(), els.begin() + half, els.end(), [&](int32_t a, int32_t c) {
        if (key[a] != key[c]) return key[a] < key[c];
        if (key2[a] != key2[c]) return key2[a] < key2[c];
        return a < c;
    });
    lo.assign(els.begin(), els.begin() + half);
    hi.assign(els.begin() + half, els.end());
}

struct PatchPlan {
    std::vector<int32_t> work;     // elements in sweep order
    std::vector<int32_t> owned;    // owned nodes
};

// First-fit allocator over [0, cap) doubles.
struct Arena {
    int64_t cap;
    std::vector<std::pair<int64_t, int64_t>> free_;  // sorted (start, len)
    int64_t used = 0, peak = 0;
    explicit Arena(int64_t c) : cap(c) { free_.push_back({0, c}); }
    int64_t alloc(int64_t len) {
        for (size_t k = 0; k < free_.size(); ++k) {
            if (free_[k].second >= len) {
                const int64_t s = free_[k].first;
                free_[k].first += len;
                free_[k].second -= len;
                if (free_[k].second == 0) free_.erase(free_.begin() + k);
                used += len;
                peak = std::max(peak, used);
                return s;
            }
        }
        return -1;
    }
    void release(int64_t s, int64_t len) {
        used -= len;
        auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(s, int64_t(0)));
        it = free_.insert(it, {s, len});
        size_t k = size_t(it - free_.begin());
        if (k + 1 < free_.size() && free_[k].first + free_[k].second == free_[k + 1].first) {
            free_[k].second += free_[k + 1].second;
            free_.erase(free_.begin() + k + 1);
        }
        if (k > 0 && free_[k - 1].first + free_[k - 1].second == free_[k].first) {
            free_[k - 1].second += free_[k].second;
            free_.erase(free_.begin() + k);
        }
    }
};

}  // namespace

int64_t slot_doubles(int64_t deg) { return 9 * deg + 6; }

// Patches of the fused kernel: bands `band_cells` element sizes high (the sweep front, which
// the LDS window holds) cut into parts of `patch_elems` elements.
void fused_patch_shape(int nod, int &patch_elems, int &band_cells) {
    band_cells = nod == 3 ? 8 : nod == 6 ? 5 : 3;
    patch_elems = nod == 3 ? 960 : nod == 6 ? 480 : 160;
    if (const char *e = std::getenv("UVH_PATCH_ELEMS")) patch_elems = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("UVH_PATCH_BAND")) band_cells = std::max(1, std::atoi(e));
}

// Entries of K that receive three or more element contributions are summed in a fixed order;
// all others are added in any order. On a conforming mesh (every edge in at most two
// elements) two distinct nodes share at most two elements, so only the diagonal block and the
// residual entries of a node with three or more incident elements (a vertex) are such sums.
// Returns "" when the mesh has that property, else a description of the first exception.
std::string check_two_contributions(const MeshHost &m) {
    const int nod = m.nod;
    std::vector<int32_t> cnt;
    for (int32_t k = 0; k < m.nn; ++k) {
        const int64_t r0 = m.rowptr_n[k], deg = m.rowptr_n[k + 1] - r0;
        cnt.assign(size_t(deg), 0);
        for (int32_t t = m.inc_ptr[k]; t < m.inc_ptr[k + 1]; ++t) {
            const int32_t e = m.inc[t] / nod, a = m.inc[t] % nod;
            for (int b = 0; b < nod; ++b)
                if (b != a) ++cnt[m.pos[(size_t(e) * nod + a) * nod + b]];
        }
        for (int64_t c = 0; c < deg; ++c)
            if (c != m.diag[k] && cnt[size_t(c)] > 2)
                return "node " + std::to_string(k) + " shares more than two elements with node " +
                       std::to_string(m.colind_n[size_t(r0 + c)]);
    }
    return "";
}

std::string build_schedule(MeshHost &m, int epb, int64_t window, int patch_elems, int band_cells, Schedule &S) {
    const int32_t nele = m.nele, nod = m.nod, nn = m.nn;
    S = Schedule{};
    S.epb = epb;
    S.window = window;
    if (epb < 1) return "schedule: epb < 1";
    {
        const std::string why = check_two_contributions(m);
        if (!why.empty()) return "schedule: " + why;
    }
    std::vector<double> cx(nele), cy(nele);
    for (int32_t e = 0; e < nele; ++e) {
        double sx = 0, sy = 0;
        for (int a = 0; a < nod; ++a) {
            sx += m.x[m.conn[size_t(e) * nod + a]];
            sy += m.y[m.conn[size_t(e) * nod + a]];
        }
        cx[e] = sx / nod;
        cy[e] = sy / nod;
    }
    // Bands across the domain's shorter side, `band_cells` element sizes high, cut along the
    // longer side into parts of about `patch_elems` elements: each part is swept along the band,
    // so its window holds the rows of a front `band_cells` cells long.
    std::vector<std::vector<int32_t>> parts;
    {
        double area = 0;
        const int c1 = nod == 3 ? 1 : nod == 6 ? 2 : 3, c2 = nod == 3 ? 2 : nod == 6 ? 4 : 6;
        for (int32_t e = 0; e < nele; ++e) {
            const int32_t *ce = &m.conn[size_t(e) * nod];
            area += 0.5 * std::fabs((m.x[ce[c1]] - m.x[ce[0]]) * (m.y[ce[c2]] - m.y[ce[0]]) -
                                    (m.x[ce[c2]] - m.x[ce[0]]) * (m.y[ce[c1]] - m.y[ce[0]]));
        }
        const double cell = std::sqrt(2.0 * area / std::max(1, nele));   // a two-triangle cell's side
        std::vector<int32_t> all(nele);
        std::iota(all.begin(), all.end(), 0);
        const Box b = bbox(all, cx, cy);
        const bool along_x = (b.x1 - b.x0) >= (b.y1 - b.y0);
        const std::vector<double> &kl = along_x ? cx : cy, &ks = along_x ? cy : cx;
        const double across = along_x ? (b.y1 - b.y0) : (b.x1 - b.x0);
        const int nbands = std::max(1, int(std::lround(across / (std::max(1, band_cells) * cell))));
        std::sort(all.begin(), all.end(), [&](int32_t a, int32_t c) {
            if (ks[a] != ks[c]) return ks[a] < ks[c];
            if (kl[a] != kl[c]) return kl[a] < kl[c];
            return a < c;
        });
        for (int k = 0; k < nbands; ++k) {
            std::vector<int32_t> band(all.begin() + int64_t(nele) * k / nbands, all.begin() + int64_t(nele) * (k + 1) / nbands);
            std::sort(band.begin(), band.end(), [&](int32_t a, int32_t c) {
                if (kl[a] != kl[c]) return kl[a] < kl[c];
                if (ks[a] != ks[c]) return ks[a] < ks[c];
                return a < c;
            });
            const int64_t nseg = std::max<int64_t>(1, (int64_t(band.size()) + patch_elems / 2) / std::max(1, patch_elems));
            for (int64_t j = 0; j < nseg; ++j)
                parts.emplace_back(band.begin() + int64_t(band.size()) * j / nseg,
                                   band.begin() + int64_t(band.size()) * (j + 1) / nseg);
        }
    }
    std::vector<int32_t> deg_e(nn);
    for (int32_t k = 0; k < nn; ++k) deg_e[k] = m.inc_ptr[k + 1] - m.inc_ptr[k];
    // Owner of a node: the part with most incident elements (ties: lowest part id).
    // Parts that overflow the LDS window are bisected again, so iterate to a fixed point.
    for (int iter = 0; iter < 64; ++iter) {
        const int32_t np = int32_t(parts.size());
        std::vector<int32_t> part_of(nele);
        for (int32_t p = 0; p < np; ++p)
            for (int32_t e : parts[p]) part_of[e] = p;
        std::vector<int32_t> owner(nn, -1);
        {
            std::vector<int32_t> cand;
            for (int32_t k = 0; k < nn; ++k) {
                cand.clear();
                for (int32_t t = m.inc_ptr[k]; t < m.inc_ptr[k + 1]; ++t) cand.push_back(part_of[m.inc[t] / nod]);
                std::sort(cand.begin(), cand.end());
                int32_t best = -1, bestc = 0;
                for (size_t s = 0; s < cand.size();) {
                    size_t t = s;
                    while (t < cand.size() && cand[t] == cand[s]) ++t;
                    if (int32_t(t - s) > bestc) { bestc = int32_t(t - s); best = cand[s]; }
                    s = t;
                }
                owner[k] = best;
            }
        }
        std::vector<PatchPlan> plan(np);
        {
            std::vector<int32_t> last_seen(np, -1);
            for (int32_t e = 0; e < nele; ++e) {
                for (int a = 0; a < nod; ++a) {
                    const int32_t p = owner[m.conn[size_t(e) * nod + a]];
                    if (p >= 0 && last_seen[p] != e) {
                        last_seen[p] = e;
                        plan[p].work.push_back(e);
                    }
                }
            }
            for (int32_t k = 0; k < nn; ++k)
                if (owner[k] >= 0) plan[owner[k]].owned.push_back(k);
        }
        std::vector<int32_t> overflow;
        std::vector<int32_t> first_b(nn, -1), last_b(nn, -1);
        std::vector<int64_t> slot_of(nn, -1);
        Schedule T;
        T.epb = epb;
        T.window = window;
        T.patch_batch_ptr.push_back(0);
        T.item_ptr.push_back(0);
        T.fix_ptr.push_back(0);
        T.fix_item_ptr.push_back(0);
        T.zero_ptr.push_back(0);
        T.flush_ptr.push_back(0);
        for (int32_t p = 0; p < np; ++p) {
            PatchPlan &pl = plan[p];
            if (pl.owned.empty()) {   // a part that owns no node (possible on tiny meshes)
                T.patch_batch_ptr.push_back(T.patch_batch_ptr.back());
                continue;
            }
            // sweep order: along the longer side of the part's own centroid box
            const Box b = bbox(parts[p], cx, cy);
            const bool along_x = (b.x1 - b.x0) >= (b.y1 - b.y0);
            const std::vector<double> &k1 = along_x ? cx : cy;
            const std::vector<double> &k2 = along_x ? cy : cx;
            std::sort(pl.work.begin(), pl.work.end(), [&](int32_t a, int32_t c) {
                if (k1[a] != k1[c]) return k1[a] < k1[c];
                if (k2[a] != k2[c]) return k2[a] < k2[c];
                return a < c;
            });
            const int32_t nb = int32_t((pl.work.size() + epb - 1) / epb);
            for (int32_t k : pl.owned) { first_b[k] = -1; last_b[k] = -1; }
            for (int32_t bi = 0; bi < nb; ++bi)
                for (int32_t s = bi * epb; s < std::min<int32_t>(int32_t(pl.work.size()), (bi + 1) * epb); ++s) {
                    const int32_t e = pl.work[s];
                    for (int a = 0; a < nod; ++a) {
                        const int32_t k = m.conn[size_t(e) * nod + a];
                        if (owner[k] != p) continue;
                        if (first_b[k] < 0) first_b[k] = bi;
                        last_b[k] = bi;
                    }
                }
            std::vector<std::vector<int32_t>> starts(nb), ends(nb);
            for (int32_t k : pl.owned) {
                starts[first_b[k]].push_back(k);
                ends[last_b[k]].push_back(k);
            }
            // A row lives in the window from the zeroing in iteration first_b to its write-out
            // in iteration last_b + 1, so its slot is free again for rows starting at last_b + 2.
            Arena arena(window);
            bool ok = true;
            for (int32_t bi = 0; bi < nb && ok; ++bi) {
                if (bi >= 2)
                    for (int32_t k : ends[bi - 2])
                        arena.release(slot_of[k], (slot_doubles(m.rowptr_n[k + 1] - m.rowptr_n[k]) + 1) & ~int64_t(1));
                for (int32_t k : starts[bi]) {
                    const int64_t len2 = (slot_doubles(m.rowptr_n[k + 1] - m.rowptr_n[k]) + 1) & ~int64_t(1);
                    const int64_t s0 = arena.alloc(len2);
                    if (s0 < 0) { ok = false; break; }
                    slot_of[k] = s0;
                    T.zero_slot.push_back(int32_t(s0));
                    T.zero_len.push_back(int32_t(len2));
                }
                if (!ok) break;
                T.zero_ptr.push_back(int32_t(T.zero_slot.size()));
                // elements, row items, ordered staging and fixup groups of this batch
                std::vector<std::pair<int32_t, std::vector<int32_t>>> groups;   // node -> items (in batch order)
                int32_t nstage = 0;
                const int32_t item0 = int32_t(T.item_ea.size());
                for (int32_t s = 0; s < epb; ++s) {
                    const int32_t idx = bi * epb + s;
                    const int32_t e = idx < int32_t(pl.work.size()) ? pl.work[idx] : -1;
                    T.batch_elem.push_back(e);
                    if (e < 0) continue;
                    for (int a = 0; a < nod; ++a) {
                        const int32_t k = m.conn[size_t(e) * nod + a];
                        if (owner[k] != p) continue;
                        T.item_ea.push_back(s * nod + a);
                        T.item_slot.push_back(int32_t(slot_of[k]));
                        T.item_stage.push_back(-1);
                        if (deg_e[k] < 3) continue;
                        auto it = std::find_if(groups.begin(), groups.end(), [&](auto &g) { return g.first == k; });
                        if (it == groups.end()) groups.push_back({k, {int32_t(T.item_ea.size() - 1)}});
                        else it->second.push_back(int32_t(T.item_ea.size() - 1));
                    }
                }
                // a node with one item in the batch takes its ordered values directly (one add per
                // batch, in batch order); with more, the items stage them and a fixup sums them
                for (auto &g : groups) {
                    if (g.second.size() < 2) continue;
                    T.fix_node.push_back(g.first);
                    T.fix_slot.push_back(int32_t(slot_of[g.first]));
                    for (int32_t itx : g.second) {
                        T.item_stage[size_t(itx)] = nstage;
                        T.fix_item.push_back(nstage);
                        ++nstage;
                    }
                    T.fix_item_ptr.push_back(int32_t(T.fix_item.size()));
                }
                T.max_stage = std::max(T.max_stage, nstage);
                T.max_items = std::max<int32_t>(T.max_items, int32_t(T.item_ea.size()) - item0);
                T.item_ptr.push_back(int32_t(T.item_ea.size()));
                T.fix_ptr.push_back(int32_t(T.fix_node.size()));
                for (int32_t k : ends[bi]) {
                    T.flush_node.push_back(k);
                    T.flush_slot.push_back(int32_t(slot_of[k]));
                }
                T.flush_ptr.push_back(int32_t(T.flush_node.size()));
            }
            if (!ok) {  // the whole schedule is rebuilt after the overflowing parts are split
                overflow.push_back(p);
                continue;
            }
            T.peak_window = std::max<int64_t>(T.peak_window, arena.peak);
            T.patch_batch_ptr.push_back(T.patch_batch_ptr.back() + nb);
            T.patch_work += int64_t(pl.work.size());
        }
        if (overflow.empty()) {
            T.npatch = int32_t(T.patch_batch_ptr.size() - 1);
            S = std::move(T);
            return "";
        }
        // bisect the parts that did not fit and rebuild the whole schedule
        std::vector<std::vector<int32_t>> next;
        std::vector<char> split(np, 0);
        for (int32_t p : overflow) split[p] = 1;
        for (int32_t p = 0; p < np; ++p) {
            if (!split[p] || parts[p].size() < 2) {
                if (split[p]) return "schedule: a single element does not fit the LDS window";
                next.push_back(std::move(parts[p]));
                continue;
            }
            // halve the part across its sweep direction, which halves its front
            std::vector<int32_t> lo, hi;
            bisect_across(parts[p], cx, cy, lo, hi);
            next.push_back(std::move(lo));
            next.push_back(std::move(hi));
        }
        parts = std::move(next);
    }
    return "schedule: could not fit the LDS window";
}

void pack_schedule(const MeshHost &m, const Schedule &S, PackedSchedule &P) {
    P = PackedSchedule{};
    const int nod = m.nod;
    const int32_t nb = S.nbatch();
    P.hdr.resize(size_t(nb));
    for (int32_t b = 0; b < nb; ++b)
        P.hdr[size_t(b)] = BatchHdr{S.item_ptr[b], S.item_ptr[b + 1], S.zero_ptr[b], S.zero_ptr[b + 1],
                                    S.fix_ptr[b], S.fix_ptr[b + 1], S.flush_ptr[b], S.flush_ptr[b + 1]};
    P.batch_node.assign(size_t(nb) * S.epb * nod, -1);
    for (size_t k = 0; k < S.batch_elem.size(); ++k) {
        const int32_t e = S.batch_elem[k];
        if (e >= 0)
            for (int a = 0; a < nod; ++a) P.batch_node[k * nod + a] = m.conn[size_t(e) * nod + a];
    }
    P.items.resize(S.item_ea.size());
    for (int32_t b = 0; b < nb; ++b)
        for (int32_t k = S.item_ptr[b]; k < S.item_ptr[b + 1]; ++k) {
            const int32_t ea = S.item_ea[k], el = ea / nod, a = ea % nod;
            const int32_t e = S.batch_elem[size_t(b) * S.epb + el];
            const int32_t node = m.conn[size_t(e) * nod + a];
            PatchItem it{};
            it.slot = S.item_slot[k];
            it.stage = S.item_stage[k];
            it.ea = ea;
            it.deg = int32_t(m.rowptr_n[node + 1] - m.rowptr_n[node]);
            for (int j = 0; j < nod; ++j) it.pos[j] = m.pos[(size_t(e) * nod + a) * nod + j];
            P.items[size_t(k)] = it;
        }
    P.fixes.resize(S.fix_node.size());
    for (size_t g = 0; g < S.fix_node.size(); ++g) {
        const int32_t node = S.fix_node[g];
        const int32_t deg = int32_t(m.rowptr_n[node + 1] - m.rowptr_n[node]);
        const int32_t i0 = S.fix_item_ptr[g], i1 = S.fix_item_ptr[g + 1];
        for (int32_t i = i0 + 1; i < i1; ++i)   // a group's staging indices are consecutive
            if (S.fix_item[i] != S.fix_item[i - 1] + 1) throw std::runtime_error("pack_schedule: staging not consecutive");
        P.fixes[g] = PatchFix{S.fix_slot[g], S.fix_item[i0], i1 - i0, deg | (int32_t(m.diag[node]) << 16)};
    }
    P.zeros.resize(S.zero_slot.size());
    for (size_t z = 0; z < S.zero_slot.size(); ++z) P.zeros[z] = PatchZero{S.zero_slot[z], S.zero_len[z]};
    P.flushes.resize(S.flush_node.size());
    for (size_t f = 0; f < S.flush_node.size(); ++f) {
        const int32_t node = S.flush_node[f];
        P.flushes[f] = PatchFlush{node, S.flush_slot[f], int32_t(m.rowptr_n[node + 1] - m.rowptr_n[node]),
                                  int32_t(m.diag[node]), 3 * m.rowptr_n[node], 0};
    }
}

}  // namespace uvh
