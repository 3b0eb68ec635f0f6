// Internal declarations shared by the host mesh builder, the kernels and the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace uvh {

constexpr int kMaxNod = 10;
constexpr int kMaxNip = 32;

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
    ShapeTables tab{};
    int64_t nnzN() const { return rowptr_n.empty() ? 0 : rowptr_n.back(); }
};

// Builds rowptr_n/colind_n, inc_ptr/inc and pos from conn; returns "" or an error message.
std::string build_mesh_host(MeshHost &m);
// Fills the shape tables for nod nodes at the nip points pts (nip x 2) with weights wts.
std::string fill_shape_tables(ShapeTables &t, int nod, int nip, const double *pts, const double *wts);

// Scalars of the assembly, device-ready (cos/sin of alpha precomputed).
struct Params {
    double dt, theta, beta0, tau_mult, uz, eps0, eta0, czero, kH, Hh0, g, rhow, ca, sa, thick_min;
    int32_t tau_opt, zero_fields, reset_thickness, enforce_b_above;
};

// Device pointers of the nodal fields, in the order of uvh_fields (dadh may be null).
struct FieldPtrs {
    const double *h, *u, *v, *S, *B, *rho, *C, *m, *n, *A, *as1, *ab1;
    const double *h0, *u0, *v0, *as0, *ab0;
    const double *dadh;
};

}  // namespace uvh
