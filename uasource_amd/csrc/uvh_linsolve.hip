// Newton-step linear solve on the device: K d = rhs with Dirichlet-fixed degrees of freedom
// eliminated, block ILU(0) preconditioning and BiCGStab.
//
// Reference: the branch CtrlVar.AsymmSolver = 'EliminateBCsSolveSystemIterativly' of
// UaMain/solveKApe.m:117-151 (constrained rows and columns removed, `ilu` with type 'nofill',
// a preconditioned Krylov solve to a relative residual tolerance; CtrlVar.LinSolveTol = 1e-8 by
// default, UaMain/Ua2D_DefaultParameters.m:659). Here
//   * K is taken in 3 x 3 node blocks, unknowns interleaved per node, so the incomplete
//     factorization is the block ILU(0) of the node graph; the nodes are renumbered by a
//     greedy colouring of the node graph (colour by colour, original order within a colour),
//     so no two nodes of a colour are coupled and each triangular solve has as many
//     sequential levels as there are colours (9-10 on 6-node meshes) instead of the ~1,600 of
//     the x-sorted node order (whose ILU(0) needs ~20 % fewer iterations but whose level-
//     scheduled solves are ~100x slower);
//   * a fixed degree of freedom gets an identity row and column and a zero right-hand side:
//     the other unknowns are then exactly those of the eliminated system and the fixed ones
//     zero (the reference's x(iConstrainedDOF) = g with g = 0, the Newton-step form);
//   * the Krylov method is BiCGStab with right preconditioning (the reference's GMRES(10)
//     took 3-6x more iterations than BiCGStab on these systems when both were tried with the
//     same ILU(0));
//   * the factorization and the triangular solves are rocSPARSE's bsrilu0 / bsrsv; the
//     CSR-to-block gather, the block product, the vector updates and the reductions are the
//     kernels below. Reductions have a fixed shape (fixed grid, fixed tree), so a solve is
//     bit for bit reproducible.

#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/uvh.h"
#include "uvh_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 512;   // fixed reduction grid: the summation order depends only on n

// Block (n, m) of the node graph entry k: value (b, c) = K[(b, n), (c, m)], read from the
// assembly's CSR order (row (b, n) at b*3*nnzN + 3*rowptr[n], its column block c at
// c*deg(n)) and written to the block's slot in the renumbered block matrix; rows and columns
// of fixed degrees of freedom become identity.
__global__ void k_to_blocks(int64_t total, int nn, int64_t nnzN, const int64_t *__restrict__ rowptr,
                            const int32_t *__restrict__ colind, const int32_t *__restrict__ node_of,
                            const int32_t *__restrict__ slot, const uint8_t *__restrict__ fixed,
                            const double *__restrict__ K, double *__restrict__ A) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t k = t / 9;
    const int bc = int(t - k * 9), b = bc / 3, c = bc - 3 * b;
    const int n = node_of[k], m = colind[k];
    const int64_t r0 = rowptr[n], deg = rowptr[n + 1] - r0;
    double v = K[int64_t(b) * 3 * nnzN + 3 * r0 + c * deg + (k - r0)];
    if (fixed && (fixed[int64_t(b) * nn + n] | fixed[int64_t(c) * nn + m])) v = (n == m && b == c) ? 1.0 : 0.0;
    A[int64_t(slot[k]) * 9 + bc] = v;
}

// y = A x for 3 x 3 row-major blocks: one thread per (node, component).
__global__ void k_block_mv(int nn, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ A, const double *__restrict__ x, double *__restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nn) return;
    const int n = t / 3, b = t - 3 * n;
    double s = 0.0;
    for (int k = rp[n]; k < rp[n + 1]; ++k) {
        const int m = ci[k];
        const double *a = A + int64_t(k) * 9 + 3 * b;
        s += a[0] * x[3 * m] + a[1] * x[3 * m + 1] + a[2] * x[3 * m + 2];
    }
    y[t] = s;
}

// [ub; vb; h] order -> renumbered, node-interleaved; zero at fixed degrees of freedom
__global__ void k_gather_rhs(int nn, const int32_t *__restrict__ newnode, const uint8_t *__restrict__ fixed,
                             const double *__restrict__ rhs, double *__restrict__ x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nn) return;
    const int b = t / nn, n = t - b * nn;
    x[3 * int64_t(newnode[n]) + b] = (fixed && fixed[t]) ? 0.0 : rhs[t];
}

__global__ void k_scatter_sol(int nn, const int32_t *__restrict__ newnode, const double *__restrict__ x,
                              double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nn) return;
    const int b = t / nn, n = t - b * nn;
    out[t] = x[3 * int64_t(newnode[n]) + b];
}

// Block-level sum of NP values per thread into part[p * gridDim.x + blockIdx.x].
template <int NP>
__device__ __forceinline__ void block_partials(double (&v)[NP], double *__restrict__ part) {
    __shared__ double red[NP][kThreads];
#pragma unroll
    for (int p = 0; p < NP; ++p) red[p][threadIdx.x] = v[p];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
#pragma unroll
            for (int p = 0; p < NP; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int p = 0; p < NP; ++p) part[p * gridDim.x + blockIdx.x] = red[p][0];
    }
}

// partials of (a0, b0) and (a1, b1) (the second pair optional)
template <int NP>
__global__ __launch_bounds__(kThreads) void k_dots(int64_t n, const double *__restrict__ a0,
                                                   const double *__restrict__ b0, const double *__restrict__ a1,
                                                   const double *__restrict__ b1, double *__restrict__ part) {
    double v[NP] = {};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        v[0] += a0[i] * b0[i];
        if (NP > 1) v[NP - 1] += a1[i] * b1[i];
    }
    block_partials<NP>(v, part);
}

// p = r + beta (p - omega v)
__global__ void k_update_p(int64_t n, const double *__restrict__ r, const double *__restrict__ v, double beta,
                           double omega, double *__restrict__ p) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// s = r - alpha v, x += alpha y, partials of (s, s)
__global__ __launch_bounds__(kThreads) void k_update_s(int64_t n, const double *__restrict__ r,
                                                       const double *__restrict__ v, const double *__restrict__ y,
                                                       double alpha, double *__restrict__ s, double *__restrict__ x,
                                                       double *__restrict__ part) {
    double acc[1] = {0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        x[i] += alpha * y[i];
        acc[0] += si * si;
    }
    block_partials<1>(acc, part);
}

// x += omega z, r = s - omega t, partials of (rhat, r) and (r, r)
__global__ __launch_bounds__(kThreads) void k_update_xr(int64_t n, const double *__restrict__ s,
                                                        const double *__restrict__ t, const double *__restrict__ z,
                                                        const double *__restrict__ rhat, double omega,
                                                        double *__restrict__ x, double *__restrict__ r,
                                                        double *__restrict__ part) {
    double acc[2] = {0.0, 0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        x[i] += omega * z[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        acc[0] += rhat[i] * ri;
        acc[1] += ri * ri;
    }
    block_partials<2>(acc, part);
}

// res = b - w (in place in w), partials of (res, res)
__global__ __launch_bounds__(kThreads) void k_residual(int64_t n, const double *__restrict__ b,
                                                       double *__restrict__ w, double *__restrict__ part) {
    double acc[1] = {0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const double d = b[i] - w[i];
        w[i] = d;
        acc[0] += d * d;
    }
    block_partials<1>(acc, part);
}

// out[p] = sum of part[p * nb .. p * nb + nb - 1] in a fixed tree order; one block
__global__ __launch_bounds__(kThreads) void k_finish(int nb, int np, const double *__restrict__ part,
                                                     double *__restrict__ out) {
    __shared__ double red[kThreads];
    for (int p = 0; p < np; ++p) {
        double v = 0.0;
        for (int i = threadIdx.x; i < nb; i += kThreads) v += part[p * nb + i];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = kThreads / 2; w > 0; w >>= 1) {
            if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[p] = red[0];
        __syncthreads();
    }
}

__global__ void k_fill(int64_t n, double v, double *__restrict__ x) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        x[i] = v;
}

__global__ void k_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        b[i] = a[i];
}

const char *sp_name(rocsparse_status s) {
    switch (s) {
        case rocsparse_status_success: return "success";
        case rocsparse_status_invalid_handle: return "invalid handle";
        case rocsparse_status_not_implemented: return "not implemented";
        case rocsparse_status_invalid_pointer: return "invalid pointer";
        case rocsparse_status_invalid_size: return "invalid size";
        case rocsparse_status_memory_error: return "memory error";
        case rocsparse_status_internal_error: return "internal error";
        case rocsparse_status_invalid_value: return "invalid value";
        case rocsparse_status_arch_mismatch: return "arch mismatch";
        case rocsparse_status_zero_pivot: return "zero pivot";
        default: return "error";
    }
}

}  // namespace

struct uvh_lin {
    std::string err;
    int device = 0;
    int nn = 0;
    int64_t nnzN = 0, n3 = 0;
    // node graph on the device
    int64_t *d_rowptr64 = nullptr;
    int32_t *d_colind = nullptr, *d_node_of = nullptr;   // original node graph
    int32_t *d_newnode = nullptr, *d_slot = nullptr;     // renumbering: node -> new node, entry -> block slot
    int32_t *d_rp = nullptr, *d_ci = nullptr;           // renumbered block pattern
    int ncolors = 0;
    uint8_t *d_fixed = nullptr;
    double *d_A = nullptr, *d_LU = nullptr;
    // BiCGStab vectors (node-interleaved) and reduction scratch
    double *d_b = nullptr, *d_x = nullptr, *d_r = nullptr, *d_rhat = nullptr, *d_p = nullptr, *d_v = nullptr;
    double *d_s = nullptr, *d_t = nullptr, *d_y = nullptr, *d_z = nullptr, *d_w = nullptr;
    double *d_part = nullptr, *d_out = nullptr;
    double *h_out = nullptr;   // pinned, 2 values
    rocsparse_handle sp = nullptr;
    rocsparse_mat_descr dM = nullptr, dL = nullptr, dU = nullptr;
    rocsparse_mat_info info = nullptr;
    void *d_buf = nullptr;
    bool factored = false;
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

namespace {

template <typename T>
hipError_t lalloc(T **p, size_t n) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void **>(p), (n ? n : 1) * sizeof(T));
}

int grid_of(int64_t n) { return int((n + kThreads - 1) / kThreads); }

// reduce NP partial sums on the stream and return them on the host
int finish(uvh_lin *L, int np, hipStream_t s, double *out) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, s, kRedBlocks, np, L->d_part, L->d_out);
    hipError_t e = hipMemcpyAsync(L->h_out, L->d_out, np * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return L->fail(UVH_E_HIP, std::string("uvh_lin_solve: ") + hipGetErrorString(e));
    for (int p = 0; p < np; ++p) out[p] = L->h_out[p];
    return UVH_OK;
}

// y = U^-1 L^-1 x
int precondition(uvh_lin *L, const double *x, double *y) {
    const double one = 1.0;
    rocsparse_status st = rocsparse_dbsrsv_solve(L->sp, rocsparse_direction_row, rocsparse_operation_none, L->nn,
                                                 rocsparse_int(L->nnzN), &one, L->dL, L->d_LU, L->d_rp, L->d_ci, 3,
                                                 L->info, x, L->d_w, rocsparse_solve_policy_auto, L->d_buf);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrsv_solve(L->sp, rocsparse_direction_row, rocsparse_operation_none, L->nn,
                                    rocsparse_int(L->nnzN), &one, L->dU, L->d_LU, L->d_rp, L->d_ci, 3, L->info, L->d_w,
                                    y, rocsparse_solve_policy_auto, L->d_buf);
    if (st != rocsparse_status_success)
        return L->fail(UVH_E_HIP, std::string("uvh_lin_solve: triangular solve: ") + sp_name(st));
    return UVH_OK;
}

void block_mv(uvh_lin *L, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_block_mv, dim3(grid_of(L->n3)), dim3(kThreads), 0, s, L->nn, L->d_rp, L->d_ci, L->d_A, x, y);
}

void destroy(uvh_lin *L) {
    if (!L) return;
    hipSetDevice(L->device);
    if (L->info) rocsparse_destroy_mat_info(L->info);
    if (L->dM) rocsparse_destroy_mat_descr(L->dM);
    if (L->dL) rocsparse_destroy_mat_descr(L->dL);
    if (L->dU) rocsparse_destroy_mat_descr(L->dU);
    if (L->sp) rocsparse_destroy_handle(L->sp);
    for (void *p : {(void *)L->d_rowptr64, (void *)L->d_colind, (void *)L->d_newnode, (void *)L->d_slot,
                    (void *)L->d_rp, (void *)L->d_ci, (void *)L->d_node_of, (void *)L->d_fixed,
                    (void *)L->d_A, (void *)L->d_LU, (void *)L->d_b, (void *)L->d_x, (void *)L->d_r, (void *)L->d_rhat,
                    (void *)L->d_p, (void *)L->d_v, (void *)L->d_s, (void *)L->d_t, (void *)L->d_y, (void *)L->d_z,
                    (void *)L->d_w, (void *)L->d_part, (void *)L->d_out, L->d_buf})
        if (p) hipFree(p);
    if (L->h_out) hipHostFree(L->h_out);
    delete L;
}

}  // namespace

extern "C" {

int uvh_lin_create(uvh_ctx *ctx, const uint8_t *fixed, uvh_lin **out) {
    if (!out) return UVH_E_ARG;
    *out = nullptr;
    const uvh::MeshHost *M = uvh::ctx_mesh(ctx);
    if (!M) return UVH_E_STATE;
    if (M->nnzN() >= (int64_t(1) << 31) / 9) return UVH_E_ARG;   // rocSPARSE indices are 32-bit
    auto *L = new uvh_lin();
    L->device = uvh::ctx_device(ctx);
    hipSetDevice(L->device);
    L->nn = M->nn;
    L->nnzN = M->nnzN();
    L->n3 = 3 * int64_t(M->nn);
    // greedy colouring of the node graph, then new numbers colour by colour
    const int nn = M->nn;
    const std::vector<int64_t> &rowptr = M->rowptr_n;
    const std::vector<int32_t> &colind = M->colind_n;
    std::vector<int32_t> color(nn, -1), mark(M->maxdeg + 2, -1);
    int ncol = 0;
    for (int n = 0; n < nn; ++n) {
        for (int64_t k = rowptr[n]; k < rowptr[n + 1]; ++k) {
            const int c = color[colind[k]];
            if (c >= 0) mark[c] = n;
        }
        int c = 0;
        while (mark[c] == n) ++c;
        color[n] = c;
        ncol = std::max(ncol, c + 1);
    }
    L->ncolors = ncol;
    std::vector<int32_t> cstart(ncol + 1, 0), newnode(nn), oldnode(nn);
    for (int n = 0; n < nn; ++n) ++cstart[color[n] + 1];
    for (int c = 0; c < ncol; ++c) cstart[c + 1] += cstart[c];
    for (int n = 0; n < nn; ++n) {
        newnode[n] = cstart[color[n]]++;
        oldnode[newnode[n]] = n;
    }
    // renumbered block pattern (columns sorted) and the slot of every original entry
    std::vector<int32_t> rp32(nn + 1, 0), ci32(L->nnzN), slot(L->nnzN), node_of(L->nnzN);
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int r = 0; r < nn; ++r) {
        const int n = oldnode[r];
        rp32[r + 1] = rp32[r] + int32_t(rowptr[n + 1] - rowptr[n]);
        row.clear();
        for (int64_t k = rowptr[n]; k < rowptr[n + 1]; ++k) {
            row.emplace_back(newnode[colind[k]], int32_t(k));
            node_of[k] = n;
        }
        std::sort(row.begin(), row.end());
        for (size_t q = 0; q < row.size(); ++q) {
            ci32[rp32[r] + q] = row[q].first;
            slot[row[q].second] = rp32[r] + int32_t(q);
        }
    }
    bool ok = lalloc(&L->d_rowptr64, nn + 1) == hipSuccess && lalloc(&L->d_rp, nn + 1) == hipSuccess &&
              lalloc(&L->d_ci, L->nnzN) == hipSuccess && lalloc(&L->d_node_of, L->nnzN) == hipSuccess &&
              lalloc(&L->d_colind, L->nnzN) == hipSuccess && lalloc(&L->d_slot, L->nnzN) == hipSuccess &&
              lalloc(&L->d_newnode, nn) == hipSuccess &&
              lalloc(&L->d_A, 9 * L->nnzN) == hipSuccess && lalloc(&L->d_LU, 9 * L->nnzN) == hipSuccess &&
              lalloc(&L->d_part, 2 * kRedBlocks) == hipSuccess && lalloc(&L->d_out, 2) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&L->h_out), 2 * sizeof(double), 0) == hipSuccess;
    for (double **v : {&L->d_b, &L->d_x, &L->d_r, &L->d_rhat, &L->d_p, &L->d_v, &L->d_s, &L->d_t, &L->d_y, &L->d_z,
                       &L->d_w})
        ok = ok && lalloc(v, L->n3) == hipSuccess;
    if (ok && fixed) ok = lalloc(&L->d_fixed, L->n3) == hipSuccess;
    if (!ok) {
        destroy(L);
        return UVH_E_NOMEM;
    }
    hipError_t e = hipMemcpy(L->d_rowptr64, rowptr.data(), (nn + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_rp, rp32.data(), rp32.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_ci, ci32.data(), L->nnzN * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_colind, colind.data(), L->nnzN * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_slot, slot.data(), L->nnzN * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_newnode, newnode.data(), nn * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(L->d_node_of, node_of.data(), L->nnzN * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && fixed) e = hipMemcpy(L->d_fixed, fixed, L->n3, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        destroy(L);
        return UVH_E_HIP;
    }
    // rocSPARSE: block ILU(0) and the two triangular solves share one analysis of the pattern
    rocsparse_status st = rocsparse_create_handle(&L->sp);
    if (st == rocsparse_status_success) st = rocsparse_create_mat_descr(&L->dM);
    if (st == rocsparse_status_success) st = rocsparse_create_mat_descr(&L->dL);
    if (st == rocsparse_status_success) st = rocsparse_create_mat_descr(&L->dU);
    if (st == rocsparse_status_success) st = rocsparse_create_mat_info(&L->info);
    if (st == rocsparse_status_success) st = rocsparse_set_mat_fill_mode(L->dL, rocsparse_fill_mode_lower);
    if (st == rocsparse_status_success) st = rocsparse_set_mat_diag_type(L->dL, rocsparse_diag_type_unit);
    if (st == rocsparse_status_success) st = rocsparse_set_mat_fill_mode(L->dU, rocsparse_fill_mode_upper);
    if (st == rocsparse_status_success) st = rocsparse_set_mat_diag_type(L->dU, rocsparse_diag_type_non_unit);
    size_t b0 = 0, b1 = 0, b2 = 0;
    const rocsparse_int mb = L->nn, nnzb = rocsparse_int(L->nnzN);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrilu0_buffer_size(L->sp, rocsparse_direction_row, mb, nnzb, L->dM, L->d_LU, L->d_rp, L->d_ci,
                                            3, L->info, &b0);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrsv_buffer_size(L->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb, L->dL,
                                          L->d_LU, L->d_rp, L->d_ci, 3, L->info, &b1);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrsv_buffer_size(L->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb, L->dU,
                                          L->d_LU, L->d_rp, L->d_ci, 3, L->info, &b2);
    if (st == rocsparse_status_success) {
        const size_t nb = std::max(b0, std::max(b1, b2));
        if (hipMalloc(&L->d_buf, nb ? nb : 1) != hipSuccess) {
            destroy(L);
            return UVH_E_NOMEM;
        }
    }
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrilu0_analysis(L->sp, rocsparse_direction_row, mb, nnzb, L->dM, L->d_LU, L->d_rp, L->d_ci, 3,
                                         L->info, rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto,
                                         L->d_buf);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrsv_analysis(L->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb, L->dL,
                                       L->d_LU, L->d_rp, L->d_ci, 3, L->info, rocsparse_analysis_policy_reuse,
                                       rocsparse_solve_policy_auto, L->d_buf);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrsv_analysis(L->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb, L->dU,
                                       L->d_LU, L->d_rp, L->d_ci, 3, L->info, rocsparse_analysis_policy_reuse,
                                       rocsparse_solve_policy_auto, L->d_buf);
    if (st == rocsparse_status_success && hipDeviceSynchronize() != hipSuccess) st = rocsparse_status_internal_error;
    if (st != rocsparse_status_success) {
        destroy(L);
        return UVH_E_HIP;
    }
    *out = L;
    return UVH_OK;
}

void uvh_lin_destroy(uvh_lin *L) { destroy(L); }

const char *uvh_lin_last_error(const uvh_lin *L) { return L ? L->err.c_str() : "null solver"; }

int uvh_lin_factor(uvh_lin *L, const double *dK, void *stream) {
    if (!L) return UVH_E_ARG;
    if (!dK) return L->fail(UVH_E_ARG, "uvh_lin_factor: null K");
    hipSetDevice(L->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t total = 9 * L->nnzN;
    hipLaunchKernelGGL(k_to_blocks, dim3(grid_of(total)), dim3(kThreads), 0, s, total, L->nn, L->nnzN, L->d_rowptr64,
                       L->d_colind, L->d_node_of, L->d_slot, L->d_fixed, dK, L->d_A);
    hipLaunchKernelGGL(k_copy, dim3(kRedBlocks * 4), dim3(kThreads), 0, s, total, L->d_A, L->d_LU);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return L->fail(UVH_E_HIP, std::string("uvh_lin_factor: ") + hipGetErrorString(e));
    rocsparse_status st = rocsparse_set_stream(L->sp, s);
    if (st == rocsparse_status_success)
        st = rocsparse_dbsrilu0(L->sp, rocsparse_direction_row, L->nn, rocsparse_int(L->nnzN), L->dM, L->d_LU, L->d_rp,
                                L->d_ci, 3, L->info, rocsparse_solve_policy_auto, L->d_buf);
    if (st != rocsparse_status_success)
        return L->fail(UVH_E_HIP, std::string("uvh_lin_factor: bsrilu0: ") + sp_name(st));
    rocsparse_int pivot = -1;
    st = rocsparse_bsrilu0_zero_pivot(L->sp, L->info, &pivot);   // waits for the factorization
    if (st == rocsparse_status_zero_pivot) {
        L->factored = false;
        return L->fail(UVH_E_NOCONV, "uvh_lin_factor: zero pivot in block row " + std::to_string(pivot));
    }
    if (st != rocsparse_status_success)
        return L->fail(UVH_E_HIP, std::string("uvh_lin_factor: zero-pivot check: ") + sp_name(st));
    L->factored = true;
    return UVH_OK;
}

int uvh_lin_solve(uvh_lin *L, const double *drhs, double *dx, double rtol, int32_t maxit, int32_t *iterations,
                  double *relres, void *stream) {
    if (!L) return UVH_E_ARG;
    if (!drhs || !dx) return L->fail(UVH_E_ARG, "uvh_lin_solve: null right-hand side or solution");
    if (!L->factored) return L->fail(UVH_E_STATE, "uvh_lin_solve: no factorization (uvh_lin_factor)");
    if (!(rtol > 0.0) || maxit < 1) return L->fail(UVH_E_ARG, "uvh_lin_solve: rtol must be > 0 and maxit >= 1");
    hipSetDevice(L->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (rocsparse_set_stream(L->sp, s) != rocsparse_status_success)
        return L->fail(UVH_E_HIP, "uvh_lin_solve: rocsparse_set_stream");
    const int64_t n = L->n3;
    const dim3 G(kRedBlocks), B(kThreads);
    double red[2];
    int st;
    hipLaunchKernelGGL(k_gather_rhs, dim3(grid_of(n)), B, 0, s, L->nn, L->d_newnode, L->d_fixed, drhs, L->d_b);
    hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_b, L->d_r);
    hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_b, L->d_rhat);
    for (double *v : {L->d_x, L->d_p, L->d_v}) hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, v);
    hipLaunchKernelGGL(k_dots<1>, G, B, 0, s, n, L->d_b, L->d_b, nullptr, nullptr, L->d_part);
    if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
    const double bnorm = std::sqrt(red[0]);
    int it = 0;
    double res = 0.0;
    bool converged = bnorm == 0.0;
    const char *breakdown = nullptr;
    if (!converged) {
        // near-breakdown (rhat nearly orthogonal to r or v, or omega = 0): restart with the
        // current residual as the new shadow vector, at most kMaxRestarts times
        constexpr double kBreak = 1e-12;
        constexpr int kMaxRestarts = 20;
        double rho = red[0], rho_old = 1.0, alpha = 1.0, omega = 1.0;   // rhat = r = b
        double rhat2 = red[0], rr = red[0];
        int restarts = 0;
        auto restart = [&]() {
            hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_r, L->d_rhat);
            hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_p);
            hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_v);
            rho = rhat2 = rr;
            rho_old = alpha = omega = 1.0;
            ++restarts;
        };
        for (it = 1; it <= maxit; ++it) {
            if (std::fabs(rho) <= kBreak * std::sqrt(rhat2 * rr)) {
                if (restarts == kMaxRestarts) { breakdown = "(rhat, r) = 0"; break; }
                restart();
            }
            const double beta = (rho / rho_old) * (alpha / omega);
            hipLaunchKernelGGL(k_update_p, G, B, 0, s, n, L->d_r, L->d_v, beta, omega, L->d_p);
            if ((st = precondition(L, L->d_p, L->d_y)) != UVH_OK) return st;
            block_mv(L, L->d_y, L->d_v, s);
            hipLaunchKernelGGL(k_dots<2>, G, B, 0, s, n, L->d_rhat, L->d_v, L->d_v, L->d_v, L->d_part);
            if ((st = finish(L, 2, s, red)) != UVH_OK) return st;
            if (std::fabs(red[0]) <= kBreak * std::sqrt(rhat2 * red[1])) {
                if (restarts == kMaxRestarts) { breakdown = "(rhat, v) = 0"; break; }
                restart();   // p and v are discarded; this iteration starts again
                continue;
            }
            alpha = rho / red[0];
            hipLaunchKernelGGL(k_update_s, G, B, 0, s, n, L->d_r, L->d_v, L->d_y, alpha, L->d_s, L->d_x, L->d_part);
            if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
            if (std::sqrt(red[0]) <= rtol * bnorm) { converged = true; break; }
            if ((st = precondition(L, L->d_s, L->d_z)) != UVH_OK) return st;
            block_mv(L, L->d_z, L->d_t, s);
            hipLaunchKernelGGL(k_dots<2>, G, B, 0, s, n, L->d_t, L->d_s, L->d_t, L->d_t, L->d_part);
            if ((st = finish(L, 2, s, red)) != UVH_OK) return st;
            omega = red[1] > 0.0 ? red[0] / red[1] : 0.0;
            hipLaunchKernelGGL(k_update_xr, G, B, 0, s, n, L->d_s, L->d_t, L->d_z, L->d_rhat, omega, L->d_x, L->d_r,
                               L->d_part);
            if ((st = finish(L, 2, s, red)) != UVH_OK) return st;
            rr = red[1];
            if (std::sqrt(rr) <= rtol * bnorm) { converged = true; break; }
            rho_old = rho;
            rho = red[0];
            if (omega == 0.0) {
                if (restarts == kMaxRestarts) { breakdown = "omega = 0"; break; }
                restart();
            }
        }
        if (it > maxit) it = maxit;
        // the true residual of the returned solution
        block_mv(L, L->d_x, L->d_w, s);
        hipLaunchKernelGGL(k_residual, G, B, 0, s, n, L->d_b, L->d_w, L->d_part);
        if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
        res = std::sqrt(red[0]) / bnorm;
    } else {
        hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_x);
    }
    hipLaunchKernelGGL(k_scatter_sol, dim3(grid_of(n)), B, 0, s, L->nn, L->d_newnode, L->d_x, dx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return L->fail(UVH_E_HIP, std::string("uvh_lin_solve: ") + hipGetErrorString(e));
    if (iterations) *iterations = it;
    if (relres) *relres = res;
    if (breakdown) return L->fail(UVH_E_NOCONV, std::string("uvh_lin_solve: BiCGStab breakdown: ") + breakdown);
    if (!converged)
        return L->fail(UVH_E_NOCONV, "uvh_lin_solve: relative residual " + std::to_string(res) + " after " +
                                         std::to_string(it) + " iterations");
    return UVH_OK;
}

}  // extern "C"
