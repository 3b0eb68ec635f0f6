// Host-side, once-per-mesh schedule of the fused assembly kernel.
//
// The mesh is cut into compact patches of elements; each patch is one workgroup. A node's
// three K rows (x, y, h) and its three residual entries are owned by exactly one patch: the
// patch holding most of the node's incident elements. A patch processes every element that
// touches one of its nodes (its own elements plus a thin ring of neighbours' elements whose
// rows it needs), in a sweep order along the patch's longer side, in batches of `epb`
// elements. Owned rows live in an LDS window only between the first and the last batch that
// touches them; the schedule gives every row a window slot (first-fit interval allocation),
// every (element, local node) a slot and a rank (its order among the batch's elements that
// touch the same node), and every batch the rows to zero before it and to write out after it.
// Summation order per K entry is therefore fixed: by batch, then by rank.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
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

// Splits `els` in two equal halves along the longer side of their centroid box.
void bisect(std::vector<int32_t> &els, const std::vector<double> &cx, const std::vector<double> &cy,
            std::vector<int32_t> &lo, std::vector<int32_t> &hi) {
    const Box b = bbox(els, cx, cy);
    const bool along_x = (b.x1 - b.x0) >= (b.y1 - b.y0);
    const std::vector<double> &key = along_x ? cx : cy;
    const std::vector<double> &key2 = along_x ? cy : cx;
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

std::string build_schedule(MeshHost &m, int epb, int64_t window, int target_patches, Schedule &S) {
    const int32_t nele = m.nele, nod = m.nod, nn = m.nn;
    S = Schedule{};
    S.epb = epb;
    S.window = window;
    if (epb < 1) return "schedule: epb < 1";
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
    // recursive bisection to ~target_patches parts
    std::vector<std::vector<int32_t>> parts;
    {
        std::vector<int32_t> all(nele);
        std::iota(all.begin(), all.end(), 0);
        const int64_t target = std::max<int64_t>(1, (int64_t(nele) + target_patches - 1) / std::max(1, target_patches));
        std::vector<std::vector<int32_t>> stack{all};
        while (!stack.empty()) {
            std::vector<int32_t> p = std::move(stack.back());
            stack.pop_back();
            if (int64_t(p.size()) <= target || p.size() < 2) {
                parts.push_back(std::move(p));
                continue;
            }
            std::vector<int32_t> lo, hi;
            bisect(p, cx, cy, lo, hi);
            stack.push_back(std::move(hi));
            stack.push_back(std::move(lo));
        }
    }
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
                    if (last_seen[p] != e) {
                        last_seen[p] = e;
                        plan[p].work.push_back(e);
                    }
                }
            }
            for (int32_t k = 0; k < nn; ++k) plan[owner[k]].owned.push_back(k);
        }
        // sweep order: along the longer side of the part's own centroid box
        std::vector<int32_t> overflow;
        std::vector<int64_t> first_b(nn), last_b(nn);
        Schedule T;
        T.epb = epb;
        T.window = window;
        T.patch_batch_ptr.push_back(0);
        for (int32_t p = 0; p < np; ++p) {
            PatchPlan &pl = plan[p];
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
            // allocate in batch order
            std::vector<std::vector<int32_t>> starts(nb), ends(nb);
            for (int32_t k : pl.owned) {
                starts[first_b[k]].push_back(k);
                ends[last_b[k]].push_back(k);
            }
            Arena arena(window);
            bool ok = true;
            // node -> slot lookup: owned nodes of this patch only
            std::vector<std::pair<int32_t, int64_t>> slot_map;
            slot_map.reserve(pl.owned.size());
            auto find_slot = [&](int32_t k) -> int64_t {
                auto it = std::lower_bound(slot_map.begin(), slot_map.end(), std::make_pair(k, int64_t(-1)));
                return (it != slot_map.end() && it->first == k) ? it->second : -1;
            };
            std::vector<std::pair<int32_t, int64_t>> pending;
            for (int32_t bi = 0; bi < nb && ok; ++bi) {
                pending.clear();
                T.zero_ptr.push_back(int32_t(T.zero_slot.size()));
                for (int32_t k : starts[bi]) {
                    const int64_t len = slot_doubles(m.rowptr_n[k + 1] - m.rowptr_n[k]);
                    const int64_t len2 = (len + 1) & ~int64_t(1);
                    const int64_t s = arena.alloc(len2);
                    if (s < 0) { ok = false; break; }
                    pending.push_back({k, s});
                    T.zero_slot.push_back(int32_t(s));
                    T.zero_len.push_back(int32_t(len2));
                }
                if (!ok) break;
                for (auto &pr : pending) {
                    auto it = std::lower_bound(slot_map.begin(), slot_map.end(), std::make_pair(pr.first, int64_t(-1)));
                    slot_map.insert(it, pr);
                }
                // per (element, local node): slot and rank
                std::vector<std::pair<int32_t, int32_t>> cnt;  // node -> count in this batch
                int maxrank = 0;
                for (int32_t s = 0; s < epb; ++s) {
                    const int32_t idx = bi * epb + s;
                    const int32_t e = idx < int32_t(pl.work.size()) ? pl.work[idx] : -1;
                    T.batch_elem.push_back(e);
                    for (int a = 0; a < nod; ++a) {
                        int32_t slot = -1, rank = 0;
                        if (e >= 0) {
                            const int32_t k = m.conn[size_t(e) * nod + a];
                            if (owner[k] == p) {
                                slot = int32_t(find_slot(k));
                                auto it = std::find_if(cnt.begin(), cnt.end(), [&](auto &c) { return c.first == k; });
                                if (it == cnt.end()) { cnt.push_back({k, 1}); rank = 0; }
                                else { rank = it->second++; }
                                maxrank = std::max(maxrank, rank);
                            }
                        }
                        T.ea_slot.push_back(slot);
                        T.ea_rank.push_back(uint8_t(rank));
                    }
                }
                T.batch_maxrank.push_back(uint8_t(maxrank));
                T.flush_ptr.push_back(int32_t(T.flush_node.size()));
                for (int32_t k : ends[bi]) {
                    const int64_t s = find_slot(k);
                    T.flush_node.push_back(k);
                    T.flush_slot.push_back(int32_t(s));
                    const int64_t len = slot_doubles(m.rowptr_n[k + 1] - m.rowptr_n[k]);
                    arena.release(s, (len + 1) & ~int64_t(1));
                    auto it = std::lower_bound(slot_map.begin(), slot_map.end(), std::make_pair(k, int64_t(-1)));
                    slot_map.erase(it);
                }
            }
            if (!ok) {  // the whole schedule is rebuilt after the overflowing parts are split
                overflow.push_back(p);
                continue;
            }
            T.peak_window = std::max<int64_t>(T.peak_window, arena.peak);
            T.patch_batch_ptr.push_back(int32_t(T.batch_maxrank.size()));
            T.patch_work += int64_t(pl.work.size());
        }
        if (overflow.empty()) {
            T.zero_ptr.push_back(int32_t(T.zero_slot.size()));
            T.flush_ptr.push_back(int32_t(T.flush_node.size()));
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
            std::vector<int32_t> lo, hi;
            bisect(parts[p], cx, cy, lo, hi);
            next.push_back(std::move(lo));
            next.push_back(std::move(hi));
        }
        parts = std::move(next);
    }
    return "schedule: could not fit the LDS window";
}

}  // namespace uvh
