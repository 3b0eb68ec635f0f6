// Host-side, once-per-mesh set-up: the node graph (CSR pattern of K), node incidence
// lists, the per-element placement table and the shape-function tables.
//
// Pattern: every element couples all its nodes in all nine blocks
// (UaMain/uvhMatrixAssembly.m:470-509), so K's pattern is fixed by the node graph
// "shares an element". Row (blk, n) of K lists adj(n), adj(n)+Nn, adj(n)+2Nn.
#include "uvh_internal.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace uvh {

namespace {

// Úa's local numbering (UaMain/shape_fun.m:52-82): c1 = xi, c2 = eta, c3 = 1-xi-eta;
// 3-node N = [c1, c3, c2]; 6-node corners 1,3,5 at c1,c3,c2 with edge nodes between;
// 10-node corners 1,4,7, edge nodes at thirds, node 10 at the centroid.
void shape(int nod, double xi, double eta, double *N, double *Nxi, double *Neta) {
    const double c1 = xi, c2 = eta, c3 = 1.0 - xi - eta;
    if (nod == 3) {
        N[0] = c1; N[1] = c3; N[2] = c2;
        Nxi[0] = 1; Nxi[1] = -1; Nxi[2] = 0;
        Neta[0] = 0; Neta[1] = -1; Neta[2] = 1;
        return;
    }
    if (nod == 6) {
        N[0] = (2 * c1 - 1) * c1;  N[1] = 4 * c3 * c1;  N[2] = (2 * c3 - 1) * c3;
        N[3] = 4 * c2 * c3;        N[4] = (2 * c2 - 1) * c2;  N[5] = 4 * c1 * c2;
        Nxi[0] = 4 * c1 - 1;  Nxi[1] = 4 * (c3 - c1);  Nxi[2] = 1 - 4 * c3;
        Nxi[3] = -4 * c2;     Nxi[4] = 0;              Nxi[5] = 4 * c2;
        Neta[0] = 0;          Neta[1] = -4 * c1;       Neta[2] = 1 - 4 * c3;
        Neta[3] = 4 * (c3 - c2);  Neta[4] = 4 * c2 - 1;  Neta[5] = 4 * c1;
        return;
    }
    // 10 nodes: corner k at barycentric l: l(3l-1)(3l-2)/2; edge node near corner a,
    // far corner b: 9/2 * l_a l_b (3 l_a - 1); centroid 27 c1 c2 c3.
    // d/dxi = d/dc1 - d/dc3, d/deta = d/dc2 - d/dc3.
    auto corner = [](double l, double &v, double &dl) {
        v = 0.5 * l * (3 * l - 1) * (3 * l - 2);
        dl = 0.5 * (27 * l * l - 18 * l + 2);
    };
    double v, dl;
    corner(c1, v, dl); N[0] = v; Nxi[0] = dl;  Neta[0] = 0;
    corner(c3, v, dl); N[3] = v; Nxi[3] = -dl; Neta[3] = -dl;
    corner(c2, v, dl); N[6] = v; Nxi[6] = 0;   Neta[6] = dl;
    // (node, near, far) with 0 = c1, 1 = c2, 2 = c3
    const int node[6] = {1, 2, 4, 5, 7, 8};
    const int nearc[6] = {0, 2, 2, 1, 1, 0};
    const int farc[6] = {2, 0, 1, 2, 0, 1};
    const double L[3] = {c1, c2, c3};
    const double dLdxi[3] = {1, 0, -1}, dLdeta[3] = {0, 1, -1};
    for (int k = 0; k < 6; ++k) {
        const double a = L[nearc[k]], b = L[farc[k]];
        N[node[k]] = 4.5 * a * b * (3 * a - 1);
        const double da = 4.5 * b * (6 * a - 1), db = 4.5 * a * (3 * a - 1);
        Nxi[node[k]] = da * dLdxi[nearc[k]] + db * dLdxi[farc[k]];
        Neta[node[k]] = da * dLdeta[nearc[k]] + db * dLdeta[farc[k]];
    }
    N[9] = 27 * c1 * c2 * c3;
    Nxi[9] = 27 * c2 * (c3 - c1);
    Neta[9] = 27 * c1 * (c3 - c2);
}

}  // namespace

std::string fill_shape_tables(ShapeTables &t, int nod, int nip, const double *pts, const double *wts) {
    if (nod != 3 && nod != 6 && nod != 10) return "nod must be 3, 6 or 10";
    if (nip < 1 || nip > kMaxNip) return "nip must be in 1..32";
    t = ShapeTables{};
    for (int q = 0; q < nip; ++q) {
        shape(nod, pts[2 * q], pts[2 * q + 1], t.N[q], t.dxi[q], t.deta[q]);
        t.w[q] = wts[q];
    }
    return "";
}

std::string build_mesh_host(MeshHost &m) {
    const int32_t nele = m.nele, nod = m.nod, nn = m.nn;
    for (int64_t k = 0; k < int64_t(nele) * nod; ++k)
        if (m.conn[k] < 0 || m.conn[k] >= nn) return "connectivity index out of range";
    // incidence lists: counting sort by node, element order kept
    m.inc_ptr.assign(size_t(nn) + 1, 0);
    for (int64_t k = 0; k < int64_t(nele) * nod; ++k) m.inc_ptr[m.conn[k] + 1]++;
    for (int32_t k = 0; k < nn; ++k) m.inc_ptr[k + 1] += m.inc_ptr[k];
    m.inc.assign(size_t(nele) * nod, 0);
    {
        std::vector<int32_t> fill(m.inc_ptr.begin(), m.inc_ptr.end() - 1);
        for (int32_t e = 0; e < nele; ++e)
            for (int a = 0; a < nod; ++a) m.inc[fill[m.conn[size_t(e) * nod + a]]++] = e * nod + a;
    }
    // node adjacency: union of the incident elements' nodes, sorted
    std::vector<int64_t> deg(size_t(nn) + 1, 0);
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto adjacency_of = [&](int32_t k, std::vector<int32_t> &buf) {
        buf.clear();
        for (int32_t t = m.inc_ptr[k]; t < m.inc_ptr[k + 1]; ++t) {
            const int32_t e = m.inc[t] / nod;
            for (int b = 0; b < nod; ++b) buf.push_back(m.conn[size_t(e) * nod + b]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    auto parallel = [&](auto &&body) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; ++t)
            th.emplace_back([&, t] {
                std::vector<int32_t> buf;
                for (int64_t k = t; k < nn; k += nthreads) body(int32_t(k), buf);
            });
        for (auto &x : th) x.join();
    };
    parallel([&](int32_t k, std::vector<int32_t> &buf) {
        adjacency_of(k, buf);
        deg[size_t(k) + 1] = int64_t(buf.size());
    });
    m.rowptr_n.assign(size_t(nn) + 1, 0);
    for (int32_t k = 0; k < nn; ++k) {
        if (deg[size_t(k) + 1] > 255) return "node degree above 255 (per-element placement uses 8 bits)";
        m.rowptr_n[size_t(k) + 1] = m.rowptr_n[k] + deg[size_t(k) + 1];
        m.maxdeg = std::max<int32_t>(m.maxdeg, int32_t(deg[size_t(k) + 1]));
        if (deg[size_t(k) + 1] == 0) m.has_isolated = true;
    }
    const int64_t nb = row_nodes_per_block(nod);
    for (int64_t k0 = 0; k0 < nn; k0 += nb)
        m.max_block_deg = std::max<int64_t>(m.max_block_deg, m.rowptr_n[std::min<int64_t>(k0 + nb, nn)] - m.rowptr_n[k0]);
    m.colind_n.assign(size_t(m.rowptr_n[nn]), 0);
    parallel([&](int32_t k, std::vector<int32_t> &buf) {
        adjacency_of(k, buf);
        std::copy(buf.begin(), buf.end(), m.colind_n.begin() + m.rowptr_n[k]);
    });
    m.diag.assign(size_t(nn), 0);
    for (int32_t k = 0; k < nn; ++k) {
        const int32_t *c0 = m.colind_n.data() + m.rowptr_n[k];
        const int32_t *c1 = m.colind_n.data() + m.rowptr_n[size_t(k) + 1];
        m.diag[k] = uint8_t(std::lower_bound(c0, c1, k) - c0);
    }
    // placement: column position of node b within row conn[e][a]
    m.pos.assign(size_t(nele) * nod * nod, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; ++t)
            th.emplace_back([&, t] {
                for (int64_t e = t; e < nele; e += nthreads)
                    for (int a = 0; a < nod; ++a) {
                        const int32_t r = m.conn[size_t(e) * nod + a];
                        const int32_t *c0 = m.colind_n.data() + m.rowptr_n[r];
                        const int32_t *c1 = m.colind_n.data() + m.rowptr_n[size_t(r) + 1];
                        for (int b = 0; b < nod; ++b) {
                            const int32_t *p = std::lower_bound(c0, c1, m.conn[size_t(e) * nod + b]);
                            m.pos[(size_t(e) * nod + a) * nod + b] = uint8_t(p - c0);
                        }
                    }
            });
        for (auto &x : th) x.join();
    }
    // incidence-ordered copies: the element row (e, a) is stored at inc_pos[e*nod + a], so the
    // rows of each node are contiguous
    // (pos_inc rows are padded to pos_inc_stride(nod) bytes, so that a lane reads a row's
    // columns with one or two 8-byte loads)
    const int pst = pos_inc_stride(nod);
    m.inc_pos.assign(size_t(nele) * nod, 0);
    m.pos_inc.assign(size_t(nele) * nod * pst, 0);
    for (int64_t t = 0; t < int64_t(nele) * nod; ++t) {
        const int32_t ea = m.inc[t];
        m.inc_pos[ea] = int32_t(t);
        std::copy(m.pos.begin() + size_t(ea) * nod, m.pos.begin() + size_t(ea + 1) * nod,
                  m.pos_inc.begin() + size_t(t) * pst);
    }
    // gather tables of the row scatter
    m.nzrow.assign(size_t(m.nnzN()), 0);
    m.imap_ptr.assign(size_t(nn) + 1, 0);
    for (int32_t k = 0; k < nn; ++k) {
        for (int64_t x = m.rowptr_n[k]; x < m.rowptr_n[size_t(k) + 1]; ++x) m.nzrow[x] = k;
        m.imap_ptr[size_t(k) + 1] = m.imap_ptr[k] + int64_t(m.inc_ptr[k + 1] - m.inc_ptr[k]) *
                                                        (m.rowptr_n[size_t(k) + 1] - m.rowptr_n[k]);
    }
    m.imap.assign(size_t(m.imap_ptr[nn]), 0xFF);
    {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nthreads; ++w)
            th.emplace_back([&, w] {
                for (int64_t k = w; k < nn; k += nthreads) {
                    const int64_t deg = m.rowptr_n[k + 1] - m.rowptr_n[k];
                    for (int32_t t = m.inc_ptr[k]; t < m.inc_ptr[k + 1]; ++t) {
                        const int32_t ea = m.inc[t];
                        const int32_t e = ea / nod;
                        uint8_t *row = m.imap.data() + m.imap_ptr[k] + int64_t(t - m.inc_ptr[k]) * deg;
                        // pos[(e*nod + a)*nod + b] is the column of node conn[e][b] in adj(k)
                        for (int b = 0; b < nod; ++b) row[m.pos[size_t(ea) * nod + b]] = uint8_t(b);
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    return "";
}

void build_csc_map(const MeshHost &m, int64_t *map) {
    // CSR slot idx = bi*3*nnzN + 3*rowptr_n[n] + bj*deg(n) + t holds K((bi,n),(bj,m)) with
    // m = adj(n)[t]. Read as CSC (same index arrays), the slot is column (bi,n), row (bj,m),
    // and must hold K((bj,m),(bi,n)), which CSR keeps at bj*3*nnzN + 3*rowptr_n[m] + bi*deg(m)
    // + (position of n in adj(m)).
    const int64_t nn = m.nn, nnzN = m.nnzN();
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nthreads; ++w)
        th.emplace_back([&, w] {
            for (int64_t n = w; n < nn; n += nthreads) {
                const int64_t rp = m.rowptr_n[n], deg = m.rowptr_n[n + 1] - rp;
                for (int64_t t = 0; t < deg; ++t) {
                    const int32_t mm = m.colind_n[rp + t];
                    const int32_t *c0 = m.colind_n.data() + m.rowptr_n[mm];
                    const int32_t *c1 = m.colind_n.data() + m.rowptr_n[mm + 1];
                    const int64_t tpos = std::lower_bound(c0, c1, int32_t(n)) - c0;
                    const int64_t degm = m.rowptr_n[mm + 1] - m.rowptr_n[mm];
                    for (int bi = 0; bi < 3; ++bi)
                        for (int bj = 0; bj < 3; ++bj)
                            map[bi * 3 * nnzN + 3 * rp + bj * deg + t] =
                                bj * 3 * nnzN + 3 * m.rowptr_n[mm] + bi * degm + tpos;
                }
            }
        });
    for (auto &x : th) x.join();
}

}  // namespace uvh
