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
//   * in that order the rows of one colour share no entries, so the factorization and both
//     triangular solves run colour by colour, one thread per block row, all rows of a colour
//     at once (k_ilu_color, k_lsolve_color, k_usolve_color); the diagonal blocks of U are kept
//     inverted. Reductions have a fixed shape (fixed grid, fixed tree): a solve is bit for bit
//     reproducible;
//   * the factorization is FP64; the triangular solves read its factors rounded to FP32 (half
//     the bytes of the preconditioner, which is most of an iteration's traffic), with FP64 sums
//     and vectors, and the products with K stay FP64;
//   * BiCGStab keeps its scalars on the device, so an iteration needs one host read of a stop
//     word, and the next iteration is queued before that read.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/uvh.h"
#include "uvh_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 512;   // fixed reduction grid: the summation order depends only on n

// index of value q = 3 b + c of block p (blocks row major, one after another)
__device__ __forceinline__ int64_t bidx(int64_t p, int q) { return p * 9 + q; }

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
    A[bidx(slot[k], bc)] = v;
}

// Row kernels below give kRowLanes lanes to one block row: lane l takes the row's blocks
// l, l + kRowLanes, ..., and the lane sums are combined by a fixed butterfly, so the result
// does not depend on timing.
constexpr int kRowLanes = 8;

__device__ __forceinline__ void row_sum(double &t0, double &t1, double &t2) {
#pragma unroll
    for (int off = kRowLanes / 2; off > 0; off >>= 1) {
        t0 += __shfl_xor(t0, off, kRowLanes);
        t1 += __shfl_xor(t1, off, kRowLanes);
        t2 += __shfl_xor(t2, off, kRowLanes);
    }
}

// t += A_p v_ci[p] over the row's blocks p0 .. p1-1 taken by this lane (blocks of type TM,
// arithmetic in FP64)
template <typename TM>
__device__ __forceinline__ void row_dot(int p0, int p1, int lane, const int32_t *__restrict__ ci,
                                        const TM *__restrict__ A, const double *__restrict__ v, double &t0,
                                        double &t1, double &t2) {
    for (int p = p0 + lane; p < p1; p += kRowLanes) {
        double a[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) a[q] = double(A[bidx(p, q)]);
        const int64_t m = 3 * int64_t(ci[p]);
        const double v0 = v[m], v1 = v[m + 1], v2 = v[m + 2];
        t0 += a[0] * v0 + a[1] * v1 + a[2] * v2;
        t1 += a[3] * v0 + a[4] * v1 + a[5] * v2;
        t2 += a[6] * v0 + a[7] * v1 + a[8] * v2;
    }
}

// y = A x for 3 x 3 row-major blocks
__global__ void k_block_mv(int nn, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ A, const double *__restrict__ x, double *__restrict__ y) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = int(g / kRowLanes), lane = int(g % kRowLanes);
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (i < nn) row_dot(rp[i], rp[i + 1], lane, ci, A, x, t0, t1, t2);
    row_sum(t0, t1, t2);
    if (i < nn && lane == 0) {
        y[3 * int64_t(i)] = t0;
        y[3 * int64_t(i) + 1] = t1;
        y[3 * int64_t(i) + 2] = t2;
    }
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

// Block-level sum of NP values per thread into part[p * stride] (part already offset to this
// block's slot).
template <int NP>
__device__ __forceinline__ void block_partials_at(double (&v)[NP], double *__restrict__ part, int stride) {
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
        for (int p = 0; p < NP; ++p) part[p * stride] = red[p][0];
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

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        b[i] = float(a[i]);
}

__global__ void k_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        b[i] = a[i];
}

// 3 x 3 row-major block helpers
__device__ __forceinline__ void mm3(const double *a, const double *b, double *c) {   // c = a b
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// Block ILU(0), the rows of one colour (new numbers r0 .. r1-1): for each lower block (i, k)
// in ascending k, L_ik = A_ik D_k^-1 and A_ij -= L_ik U_kj for every j > k in both rows;
// then D_i^-1 replaces the diagonal block. Rows k < r0 are final (earlier colours); rows of
// this colour do not read each other.
template <typename TM>
__device__ __forceinline__ void load_block(const TM *__restrict__ LU, int p, double (&a)[9]) {
#pragma unroll
    for (int q = 0; q < 9; ++q) a[q] = double(LU[bidx(p, q)]);
}

__global__ void k_ilu_color(int r0, int r1, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                            const int32_t *__restrict__ dslot, double *__restrict__ LU, int *__restrict__ status) {
    const int i = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r1) return;
    const int e1 = rp[i + 1], di = dslot[i];
    for (int p = rp[i]; p < di; ++p) {
        const int k = ci[p];
        double a[9], dk[9], l[9];
        load_block(LU, p, a);
        load_block(LU, dslot[k], dk);
        mm3(a, dk, l);
#pragma unroll
        for (int q = 0; q < 9; ++q) LU[bidx(p, q)] = l[q];
        // merge the rest of row i with the upper part of row k (both sorted by column)
        int q = p + 1, r = dslot[k] + 1;
        const int k1 = rp[k + 1];
        while (q < e1 && r < k1) {
            const int jq = ci[q], jr = ci[r];
            if (jq == jr) {
                double u[9], lu[9];
                load_block(LU, r, u);
                mm3(l, u, lu);
#pragma unroll
                for (int t = 0; t < 9; ++t) LU[bidx(q, t)] -= lu[t];
                ++q;
                ++r;
            } else if (jq < jr) {
                ++q;
            } else {
                ++r;
            }
        }
    }
    // invert the pivot block (cofactors over the determinant)
    double d[9];
    load_block(LU, di, d);
    const double c0 = d[4] * d[8] - d[5] * d[7], c1 = d[5] * d[6] - d[3] * d[8], c2 = d[3] * d[7] - d[4] * d[6];
    const double det = d[0] * c0 + d[1] * c1 + d[2] * c2;
    if (!(det != 0.0) || !isfinite(det)) {
        atomicMin(status, i);   // first zero pivot (smallest row) reported
        return;
    }
    const double r = 1.0 / det;
    double v[9];
    v[0] = c0 * r; v[1] = (d[2] * d[7] - d[1] * d[8]) * r; v[2] = (d[1] * d[5] - d[2] * d[4]) * r;
    v[3] = c1 * r; v[4] = (d[0] * d[8] - d[2] * d[6]) * r; v[5] = (d[2] * d[3] - d[0] * d[5]) * r;
    v[6] = c2 * r; v[7] = (d[1] * d[6] - d[0] * d[7]) * r; v[8] = (d[0] * d[4] - d[1] * d[3]) * r;
#pragma unroll
    for (int q = 0; q < 9; ++q) LU[bidx(di, q)] = v[q];
}

// The preconditioner M^-1 = U^-1 L^-1 of the block ILU(0), colour by colour. Rows of colour 0
// have no lower blocks (every neighbour has a larger new number), so L^-1 x equals x there
// and the lower solves read x for those columns instead of a copy; rows of the last colour
// have no upper blocks, so their upper solve x_i = D_i^-1 y_i is done in the lower-solve launch
// of that colour.

// y_i = x_i - sum over lower blocks L_ik y_k (y_k = x_k for k < c1, the rows of colour 0) for
// row i < r1 of one colour (L has identity diagonal blocks); with last, z_i = D_i^-1 y_i is
// written instead. All kRowLanes lanes of the row's group take part (the row sum shuffles).
template <typename TM>
__device__ __forceinline__ void lsolve_row(int i, int r1, int lane, int c1, int last, const int32_t *__restrict__ rp,
                                           const int32_t *__restrict__ ci, const int32_t *__restrict__ dslot,
                                           const TM *__restrict__ LU, const double *__restrict__ x,
                                           double *__restrict__ y, double *__restrict__ z) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (i < r1) {
        for (int p = rp[i] + lane, p1 = dslot[i]; p < p1; p += kRowLanes) {
            double a[9];
            load_block(LU, p, a);
            const int k = ci[p];
            const double *src = k < c1 ? x : y;
            const int64_t m = 3 * int64_t(k);
            const double v0 = src[m], v1 = src[m + 1], v2 = src[m + 2];
            t0 += a[0] * v0 + a[1] * v1 + a[2] * v2;
            t1 += a[3] * v0 + a[4] * v1 + a[5] * v2;
            t2 += a[6] * v0 + a[7] * v1 + a[8] * v2;
        }
    }
    row_sum(t0, t1, t2);
    if (i < r1 && lane == 0) {
        t0 = x[3 * int64_t(i)] - t0;
        t1 = x[3 * int64_t(i) + 1] - t1;
        t2 = x[3 * int64_t(i) + 2] - t2;
        if (last) {
            double d[9];
            load_block(LU, dslot[i], d);
            z[3 * int64_t(i)] = d[0] * t0 + d[1] * t1 + d[2] * t2;
            z[3 * int64_t(i) + 1] = d[3] * t0 + d[4] * t1 + d[5] * t2;
            z[3 * int64_t(i) + 2] = d[6] * t0 + d[7] * t1 + d[8] * t2;
        } else {
            y[3 * int64_t(i)] = t0;
            y[3 * int64_t(i) + 1] = t1;
            y[3 * int64_t(i) + 2] = t2;
        }
    }
}

// z_i = D_i^-1 (y_i - sum over upper blocks U_ij z_j) for row i < r1 of one colour; y_i is read
// from x for the rows of colour 0 (i < c1)
template <typename TM>
__device__ __forceinline__ void usolve_row(int i, int r1, int lane, int c1, const int32_t *__restrict__ rp,
                                           const int32_t *__restrict__ ci, const int32_t *__restrict__ dslot,
                                           const TM *__restrict__ LU, const double *__restrict__ x,
                                           const double *__restrict__ y, double *__restrict__ z) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (i < r1) row_dot(dslot[i] + 1, rp[i + 1], lane, ci, LU, z, t0, t1, t2);
    row_sum(t0, t1, t2);
    if (i < r1 && lane == 0) {
        const double *yi = (i < c1 ? x : y) + 3 * int64_t(i);
        t0 = yi[0] - t0;
        t1 = yi[1] - t1;
        t2 = yi[2] - t2;
        double d[9];
        load_block(LU, dslot[i], d);
        z[3 * int64_t(i)] = d[0] * t0 + d[1] * t1 + d[2] * t2;
        z[3 * int64_t(i) + 1] = d[3] * t0 + d[4] * t1 + d[5] * t2;
        z[3 * int64_t(i) + 2] = d[6] * t0 + d[7] * t1 + d[8] * t2;
    }
}

// the rows r0 .. r1-1 of one colour, kRowLanes lanes per row
template <typename TM>
__global__ void k_lsolve_color(int r0, int r1, int c1, int last, const int32_t *__restrict__ rp,
                               const int32_t *__restrict__ ci, const int32_t *__restrict__ dslot,
                               const TM *__restrict__ LU, const double *__restrict__ x,
                               double *__restrict__ y, double *__restrict__ z) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    lsolve_row(r0 + int(g / kRowLanes), r1, int(g % kRowLanes), c1, last, rp, ci, dslot, LU, x, y, z);
}

template <typename TM>
__global__ void k_usolve_color(int r0, int r1, int c1, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                               const int32_t *__restrict__ dslot, const TM *__restrict__ LU,
                               const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ z) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    usolve_row(r0 + int(g / kRowLanes), r1, int(g % kRowLanes), c1, rp, ci, dslot, LU, x, y, z);
}

// BiCGStab with its scalars on the device. Each iteration's kernels read rho, alpha, omega and
// the dot products from `sc` and decide breakdown, convergence and the best iterate there, so
// the host reads one stop word per iteration and can launch the next iteration before it has
// read the last one's word: every kernel that changes the solver state returns at once while
// the stop word is set, which makes an iteration launched after a stop a no-op. The arithmetic
// and its order are those of the host-driven form (the same IEEE operations on the same values).
enum : int { S_RHO, S_RHO_OLD, S_ALPHA, S_OMEGA, S_RHAT2, S_RR, S_BEST, S_THR, S_RV, S_VV, S_SS, S_TS, S_TT,
             S_RHATR, S_RRN, S_COUNT };
enum : int { F_STOP, F_COPY, F_COUNT };   // stop word; x is to be copied to xbest
enum : int { B_RUN, B_BREAK_V, B_CONV_S, B_CONV_R, B_BREAK_R, B_OMEGA0 };   // stop word values
constexpr double kBreak = 1e-12;   // near-breakdown threshold of the inner products

__global__ void k_bi_init(double *__restrict__ sc, int *__restrict__ flag, double bnorm, double thr) {
    sc[S_RHO] = sc[S_RHAT2] = sc[S_RR] = bnorm * bnorm;
    sc[S_RHO_OLD] = sc[S_ALPHA] = sc[S_OMEGA] = 1.0;
    sc[S_BEST] = bnorm;
    sc[S_THR] = thr;
    flag[F_STOP] = B_RUN;
    flag[F_COPY] = 0;
}

// restart with the current residual as the shadow vector (its copy and p = v = 0 are separate
// launches): rho = (rhat, rhat) = (r, r)
__global__ void k_bi_restart(double *__restrict__ sc, int *__restrict__ flag) {
    sc[S_RHO] = sc[S_RHAT2] = sc[S_RR];
    sc[S_RHO_OLD] = sc[S_ALPHA] = sc[S_OMEGA] = 1.0;
    flag[F_STOP] = B_RUN;
}

// xbest = x if the last residual check asked for it; then, unless stopped,
// p = r + beta (p - omega v), beta = (rho / rho_old) (alpha / omega)
__global__ void k_bi_p(int64_t n, const double *__restrict__ r, const double *__restrict__ v,
                       double *__restrict__ p, const double *__restrict__ x, double *__restrict__ xbest,
                       const double *__restrict__ sc, const int *__restrict__ flag) {
    const bool cp = flag[F_COPY] != 0, stop = flag[F_STOP] != B_RUN;
    const double beta = (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]), omega = sc[S_OMEGA];
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        if (cp) xbest[i] = x[i];
        if (!stop) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
}

// unless stopped or (rhat, v) near zero: alpha = rho / (rhat, v), s = r - alpha v,
// x += alpha y, partials of (s, s)
__global__ __launch_bounds__(kThreads) void k_bi_s(int64_t n, const double *__restrict__ r,
                                                   const double *__restrict__ v, const double *__restrict__ y,
                                                   double *__restrict__ s, double *__restrict__ x,
                                                   double *__restrict__ part, double *__restrict__ sc,
                                                   int *__restrict__ flag) {
    const double rv = sc[S_RV];
    // the same values in every thread, so every block takes the same branch
    const bool brk = std::fabs(rv) <= kBreak * std::sqrt(sc[S_RHAT2] * sc[S_VV]);
    if (flag[F_STOP] != B_RUN || brk) {
        if (brk && blockIdx.x == 0 && threadIdx.x == 0 && flag[F_STOP] == B_RUN) flag[F_STOP] = B_BREAK_V;
        return;
    }
    const double alpha = sc[S_RHO] / rv;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_ALPHA] = alpha;
    double acc[1] = {0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        x[i] += alpha * y[i];
        acc[0] += si * si;
    }
    block_partials<1>(acc, part);
}

// unless stopped: xbest = x if asked for, omega = (t, s) / (t, t), x += omega z,
// r = s - omega t, partials of (rhat, r) and (r, r)
__global__ __launch_bounds__(kThreads) void k_bi_xr(int64_t n, const double *__restrict__ s,
                                                    const double *__restrict__ t, const double *__restrict__ z,
                                                    const double *__restrict__ rhat, double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ xbest,
                                                    double *__restrict__ part, double *__restrict__ sc,
                                                    const int *__restrict__ flag) {
    if (flag[F_STOP] != B_RUN) return;
    const double tt = sc[S_TT];
    const double omega = tt > 0.0 ? sc[S_TS] / tt : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_OMEGA] = omega;
    const bool cp = flag[F_COPY] != 0;
    double acc[2] = {0.0, 0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        if (cp) xbest[i] = x[i];
        x[i] += omega * z[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        acc[0] += rhat[i] * ri;
        acc[1] += ri * ri;
    }
    block_partials<2>(acc, part);
}

// sc[slot + p] = sum of part[p * nb ..] in the fixed tree order of k_finish (np <= 2); then
// mode 1 (after k_bi_s): stop at convergence, else note a better iterate (copied in k_bi_xr);
// mode 2 (after k_bi_xr): rr, stop at convergence, note a better iterate (copied in the next
// k_bi_p), rho_old = rho, rho = (rhat, r), stop for a restart when omega = 0 or (rhat, r) is
// near zero
__global__ __launch_bounds__(kThreads) void k_bi_finish(int nb, int np, const double *__restrict__ part,
                                                        double *__restrict__ sc, int slot, int mode,
                                                        int *__restrict__ flag) {
    __shared__ double red[kThreads];
    double tot[2] = {0.0, 0.0};
    for (int p = 0; p < np; ++p) {
        double v = 0.0;
        for (int i = threadIdx.x; i < nb; i += kThreads) v += part[p * nb + i];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = kThreads / 2; w > 0; w >>= 1) {
            if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        tot[p] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int p = 0; p < np; ++p) sc[slot + p] = tot[p];
    if (mode == 0) return;
    flag[F_COPY] = 0;
    if (flag[F_STOP] != B_RUN) return;
    if (mode == 1) {
        const double nrm = std::sqrt(tot[0]);
        if (nrm <= sc[S_THR]) {
            flag[F_STOP] = B_CONV_S;
        } else if (nrm < sc[S_BEST]) {
            sc[S_BEST] = nrm;
            flag[F_COPY] = 1;
        }
        return;
    }
    const double rr = tot[1], nrm = std::sqrt(rr);
    sc[S_RR] = rr;
    if (nrm <= sc[S_THR]) {
        flag[F_STOP] = B_CONV_R;
        return;
    }
    if (nrm < sc[S_BEST]) {
        sc[S_BEST] = nrm;
        flag[F_COPY] = 1;
    }
    sc[S_RHO_OLD] = sc[S_RHO];
    sc[S_RHO] = tot[0];
    if (sc[S_OMEGA] == 0.0)
        flag[F_STOP] = B_OMEGA0;
    else if (std::fabs(tot[0]) <= kBreak * std::sqrt(sc[S_RHAT2] * rr))
        flag[F_STOP] = B_BREAK_R;
}

// xbest = x if the last residual check asked for it
__global__ void k_copy_if(int64_t n, const double *__restrict__ x, double *__restrict__ xbest,
                          const int *__restrict__ flag) {
    if (!flag[F_COPY]) return;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        xbest[i] = x[i];
}

// Batched BiCGStab (uvhb_lin_*): the systems of many members advance together, one launch per
// step of the iteration for all of them, over flat work lists of (member, vector chunk),
// (member, 32 block rows) and, per colour index, (member, 32 rows of that colour). Each member
// keeps its own scalars and stop word; restarts, the iteration count and the iteration limit
// are handled on the device too, so the host only waits until every member has stopped.
enum : int { F_ITS = F_COUNT, F_RESTARTS, F_MAXIT, F_BWORDS = 8 };   // the batched form's flag words
enum : int { B_FAIL = 6, B_MAXIT = 7, B_IDLE = 8 };                  // its further stop word values
constexpr int kMaxRestarts = 20;                                      // near-breakdown restarts per solve
constexpr int kChunk = 4 * kThreads;                                  // vector values per work item
constexpr int kRowsPerItem = kThreads / kRowLanes;                    // block rows per work item

struct LinMember {
    const int32_t *rp, *ci, *dslot;
    const double *A;
    const void *LU;   // float when lu32, else double
    double *r, *rhat, *p, *v, *s, *t, *y, *z, *w, *x, *xbest, *part, *sc;
    int *flag;
    int64_t n3;
    int nchunk, lu32;
};

__device__ __forceinline__ bool bb_terminal(int w) { return w == B_CONV_S || w == B_CONV_R || w >= B_FAIL; }
__device__ __forceinline__ bool bb_restart(int w) { return w == B_BREAK_V || w == B_BREAK_R || w == B_OMEGA0; }

// xbest = x if asked for; then p = r + beta (p - omega v), or, when the last iteration asked
// for a restart, rhat = p = r and v = 0 (the restarted scalars are set in k_bb_finish)
__global__ __launch_bounds__(kThreads) void k_bb_p(const LinMember *__restrict__ M, const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    const int w = L.flag[F_STOP];
    const bool cp = L.flag[F_COPY] != 0;
    const int64_t i0 = int64_t(it.y) * kChunk, i1 = min(L.n3, i0 + kChunk);
    if (bb_terminal(w)) {
        if (cp)
            for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) L.xbest[i] = L.x[i];
        return;
    }
    if (bb_restart(w)) {
        for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
            if (cp) L.xbest[i] = L.x[i];
            L.rhat[i] = L.r[i];
            L.p[i] = L.r[i];
            L.v[i] = 0.0;
        }
        return;
    }
    const double *sc = L.sc;
    const double beta = (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]), omega = sc[S_OMEGA];
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        if (cp) L.xbest[i] = L.x[i];
        L.p[i] = L.r[i] + beta * (L.p[i] - omega * L.v[i]);
    }
}

// z = M^-1 x, lower solves: one colour index per launch (which = 0: x = p, z = y; 1: x = s,
// z = z); item (member, first row, end of the colour's rows, last + 2 c1)
__global__ __launch_bounds__(kThreads) void k_bb_lsolve(const LinMember *__restrict__ M, const int4 *__restrict__ items,
                                                        int which) {
    const int4 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    if (bb_terminal(L.flag[F_STOP])) return;
    const int i = it.y + int(threadIdx.x) / kRowLanes, lane = int(threadIdx.x) % kRowLanes;
    const double *x = which ? L.s : L.p;
    double *z = which ? L.z : L.y;
    if (L.lu32)
        lsolve_row(i, it.z, lane, it.w >> 1, it.w & 1, L.rp, L.ci, L.dslot, static_cast<const float *>(L.LU), x, L.w,
                   z);
    else
        lsolve_row(i, it.z, lane, it.w >> 1, it.w & 1, L.rp, L.ci, L.dslot, static_cast<const double *>(L.LU), x,
                   L.w, z);
}

// upper solves, one colour index per launch; item (member, first row, end of the colour's rows, c1)
__global__ __launch_bounds__(kThreads) void k_bb_usolve(const LinMember *__restrict__ M, const int4 *__restrict__ items,
                                                        int which) {
    const int4 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    if (bb_terminal(L.flag[F_STOP])) return;
    const int i = it.y + int(threadIdx.x) / kRowLanes, lane = int(threadIdx.x) % kRowLanes;
    const double *x = which ? L.s : L.p;
    double *z = which ? L.z : L.y;
    if (L.lu32)
        usolve_row(i, it.z, lane, it.w, L.rp, L.ci, L.dslot, static_cast<const float *>(L.LU), x, L.w, z);
    else
        usolve_row(i, it.z, lane, it.w, L.rp, L.ci, L.dslot, static_cast<const double *>(L.LU), x, L.w, z);
}

// which = 0: v = A y; 1: t = A z; item (member, first block row)
__global__ __launch_bounds__(kThreads) void k_bb_mv(const LinMember *__restrict__ M, const int2 *__restrict__ items,
                                                    const int *__restrict__ nrows, int which) {
    const int2 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    if (bb_terminal(L.flag[F_STOP])) return;
    const int i = it.y + int(threadIdx.x) / kRowLanes, lane = int(threadIdx.x) % kRowLanes;
    const int nn = nrows[it.x];
    const double *x = which ? L.z : L.y;
    double *y = which ? L.t : L.v;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (i < nn) row_dot(L.rp[i], L.rp[i + 1], lane, L.ci, L.A, x, t0, t1, t2);
    row_sum(t0, t1, t2);
    if (i < nn && lane == 0) {
        y[3 * int64_t(i)] = t0;
        y[3 * int64_t(i) + 1] = t1;
        y[3 * int64_t(i) + 2] = t2;
    }
}

// partials of (rhat, v) and (v, v) (which = 0) or (t, s) and (t, t) (which = 1) into
// part[p * nchunk + chunk]
__global__ __launch_bounds__(kThreads) void k_bb_dots(const LinMember *__restrict__ M, const int2 *__restrict__ items,
                                                      int which) {
    const int2 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    if (bb_terminal(L.flag[F_STOP])) return;
    const double *a0 = which ? L.t : L.rhat, *b0 = which ? L.s : L.v, *a1 = which ? L.t : L.v;
    const int64_t i0 = int64_t(it.y) * kChunk, i1 = min(L.n3, i0 + kChunk);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        acc[0] += a0[i] * b0[i];
        acc[1] += a1[i] * a1[i];
    }
    block_partials_at<2>(acc, L.part + it.y, L.nchunk);
}

// as k_bi_s, per member chunk
__global__ __launch_bounds__(kThreads) void k_bb_s(const LinMember *__restrict__ M, const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    double *sc = L.sc;
    int *flag = L.flag;
    const double rv = sc[S_RV];
    const bool brk = std::fabs(rv) <= kBreak * std::sqrt(sc[S_RHAT2] * sc[S_VV]);
    if (flag[F_STOP] != B_RUN || brk) {
        if (brk && it.y == 0 && threadIdx.x == 0 && flag[F_STOP] == B_RUN) flag[F_STOP] = B_BREAK_V;
        return;
    }
    const double alpha = sc[S_RHO] / rv;
    if (it.y == 0 && threadIdx.x == 0) sc[S_ALPHA] = alpha;
    const int64_t i0 = int64_t(it.y) * kChunk, i1 = min(L.n3, i0 + kChunk);
    double acc[1] = {0.0};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        const double si = L.r[i] - alpha * L.v[i];
        L.s[i] = si;
        L.x[i] += alpha * L.y[i];
        acc[0] += si * si;
    }
    block_partials_at<1>(acc, L.part + it.y, L.nchunk);
}

// as k_bi_xr, per member chunk
__global__ __launch_bounds__(kThreads) void k_bb_xr(const LinMember *__restrict__ M, const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const LinMember &L = M[it.x];
    if (L.flag[F_STOP] != B_RUN) return;
    double *sc = L.sc;
    const double tt = sc[S_TT];
    const double omega = tt > 0.0 ? sc[S_TS] / tt : 0.0;
    if (it.y == 0 && threadIdx.x == 0) sc[S_OMEGA] = omega;
    const bool cp = L.flag[F_COPY] != 0;
    const int64_t i0 = int64_t(it.y) * kChunk, i1 = min(L.n3, i0 + kChunk);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        if (cp) L.xbest[i] = L.x[i];
        L.x[i] += omega * L.z[i];
        const double ri = L.s[i] - omega * L.t[i];
        L.r[i] = ri;
        acc[0] += L.rhat[i] * ri;
        acc[1] += ri * ri;
    }
    block_partials_at<2>(acc, L.part + it.y, L.nchunk);
}

// one workgroup per member: sc[slot + p] = the fixed-order sum of its chunk partials, then
// mode 0 (first dot products of an iteration): a requested restart (rho = (rhat, rhat) =
//   (r, r), the other scalars 1; past kMaxRestarts the member fails) and the iteration count;
// mode 1 (after k_bb_s) and mode 2 (after k_bb_xr): the decisions of k_bi_finish, and at the
//   end of an iteration the iteration limit; mode 3: the sums only
__global__ __launch_bounds__(kThreads) void k_bb_finish(const LinMember *__restrict__ M, int np, int slot, int mode) {
    const LinMember &L = M[blockIdx.x];
    int *flag = L.flag;
    double *sc = L.sc;
    const int w0 = flag[F_STOP];
    if (bb_terminal(w0)) {
        if (threadIdx.x == 0 && mode != 0 && mode != 3) flag[F_COPY] = 0;
        return;
    }
    __shared__ double red[kThreads];
    double tot[2] = {0.0, 0.0};
    for (int p = 0; p < np; ++p) {
        double v = 0.0;
        for (int i = threadIdx.x; i < L.nchunk; i += kThreads) v += L.part[p * L.nchunk + i];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = kThreads / 2; w > 0; w >>= 1) {
            if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        tot[p] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        if (bb_restart(w0)) {
            if (flag[F_RESTARTS] == kMaxRestarts) {
                flag[F_STOP] = B_FAIL;
                return;
            }
            ++flag[F_RESTARTS];
            sc[S_RHO] = sc[S_RHAT2] = sc[S_RR];
            sc[S_RHO_OLD] = sc[S_ALPHA] = sc[S_OMEGA] = 1.0;
            flag[F_STOP] = B_RUN;
        }
        ++flag[F_ITS];
    }
    for (int p = 0; p < np; ++p) sc[slot + p] = tot[p];
    if (mode == 0 || mode == 3) return;
    flag[F_COPY] = 0;
    if (mode == 1) {
        if (flag[F_STOP] != B_RUN) return;
        const double nrm = std::sqrt(tot[0]);
        if (nrm <= sc[S_THR]) {
            flag[F_STOP] = B_CONV_S;
        } else if (nrm < sc[S_BEST]) {
            sc[S_BEST] = nrm;
            flag[F_COPY] = 1;
        }
        return;
    }
    if (flag[F_STOP] == B_RUN) {
        const double rr = tot[1], nrm = std::sqrt(rr);
        sc[S_RR] = rr;
        if (nrm <= sc[S_THR]) {
            flag[F_STOP] = B_CONV_R;
            return;
        }
        if (nrm < sc[S_BEST]) {
            sc[S_BEST] = nrm;
            flag[F_COPY] = 1;
        }
        sc[S_RHO_OLD] = sc[S_RHO];
        sc[S_RHO] = tot[0];
        if (sc[S_OMEGA] == 0.0)
            flag[F_STOP] = B_OMEGA0;
        else if (std::fabs(tot[0]) <= kBreak * std::sqrt(sc[S_RHAT2] * rr))
            flag[F_STOP] = B_BREAK_R;
    }
    if (flag[F_ITS] >= flag[F_MAXIT] && !bb_terminal(flag[F_STOP])) flag[F_STOP] = B_MAXIT;
}

// words[m] = stop word of member m
__global__ void k_bb_words(const LinMember *__restrict__ M, int n, int *__restrict__ words) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) words[m] = M[m].flag[F_STOP];
}

// a member's counters at the start of a solve; word = B_RUN for an active member (with
// k_bi_init), B_IDLE for one left out, B_CONV_R for a zero right-hand side
__global__ void k_bb_start(int *__restrict__ flag, int word, int maxit) {
    flag[F_STOP] = word;
    flag[F_COPY] = 0;
    flag[F_ITS] = 0;
    flag[F_RESTARTS] = 0;
    flag[F_MAXIT] = maxit;
}

// GMRES helpers over a basis V of k vectors of length n, stored one after another
constexpr int kMaxKrylov = 64;   // longest GMRES cycle

// part[c * gridDim.x + blockIdx.x] = partial of (V_c, w), c = blockIdx.y
__global__ __launch_bounds__(kThreads) void k_multi_dot(int64_t n, const double *__restrict__ V,
                                                        const double *__restrict__ w, double *__restrict__ part) {
    const double *v = V + int64_t(blockIdx.y) * n;
    double acc[1] = {0.0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        acc[0] += v[i] * w[i];
    block_partials<1>(acc, part + int64_t(blockIdx.y) * gridDim.x);
}

// w += sign * sum_c coef[c] V_c (coefficients read on the device)
__global__ void k_multi_axpy(int64_t n, int k, const double *__restrict__ V, const double *__restrict__ coef,
                             double sign, double *__restrict__ w) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        double t = 0.0;
        for (int c = 0; c < k; ++c) t += coef[c] * V[int64_t(c) * n + i];
        w[i] += sign * t;
    }
}

__global__ void k_scale(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        y[i] = a * x[i];
}

__global__ void k_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
        y[i] += a * x[i];
}

__global__ void k_set_int(int *p, int v) { *p = v; }

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
    float *d_LUf = nullptr;                           // the factors rounded to FP32 (lu32)
    bool lu32 = true;                                 // triangular solves read d_LUf instead of d_LU
    // BiCGStab vectors (node-interleaved) and reduction scratch
    double *d_b = nullptr, *d_x = nullptr, *d_r = nullptr, *d_rhat = nullptr, *d_p = nullptr, *d_v = nullptr;
    double *d_s = nullptr, *d_t = nullptr, *d_y = nullptr, *d_z = nullptr, *d_w = nullptr;
    double *d_part = nullptr, *d_out = nullptr;       // reduction partials and results (kMaxKrylov + 2 values)
    double *h_out = nullptr;                          // pinned: 3 (kMaxKrylov + 2) values
    double *d_xbest = nullptr;                        // BiCGStab: iterate with the smallest residual so far
    double *d_sc = nullptr;                           // BiCGStab scalars (S_COUNT values)
    int *d_flag = nullptr;                            // BiCGStab stop word and copy flag (F_COUNT values)
    int *h_flag = nullptr;                            // pinned: the stop words of two consecutive iterations
    hipEvent_t ev[2] = {nullptr, nullptr};            // recorded after each of those copies
    // Krylov method (uvh_lin_set_krylov) and the GMRES basis (allocated on first use)
    int method = UVH_LIN_AUTO, restart = 30;
    double *d_V = nullptr, *d_coef = nullptr;
    int32_t *d_dslot = nullptr;                         // slot of each row's diagonal block
    std::vector<int32_t> cstart;                        // rows of colour c: cstart[c] .. cstart[c+1]-1
    int *d_status = nullptr;                            // smallest row with a singular pivot block
    int *h_status = nullptr;                            // pinned
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

// y = U^-1 L^-1 x: colours ascending for L, descending for U; w holds L^-1 x on the rows of
// colours 1 .. nc-1 (x and y must not be w). Colour 0 needs no lower solve and the last colour
// no separate upper solve, so nc colours take 2 nc - 2 launches (one when nc = 1). The factors
// are read as TM (float: the FP64 factors rounded once after the factorization; every sum and
// every vector stays FP64, so the preconditioner is a fixed linear operator and the Krylov
// method reaches the same tolerances).
template <typename TM>
void precondition_t(uvh_lin *L, const TM *LU, const double *x, double *y, hipStream_t s) {
    const int nc = int(L->cstart.size()) - 1;
    const int c1 = L->cstart[1];
    for (int c = nc == 1 ? 0 : 1; c < nc; ++c) {
        const int r0 = L->cstart[c], r1 = L->cstart[c + 1];
        hipLaunchKernelGGL(k_lsolve_color<TM>, dim3(grid_of(int64_t(r1 - r0) * kRowLanes)), dim3(kThreads), 0, s, r0, r1,
                           c == 0 ? 0 : c1, int(c == nc - 1), L->d_rp, L->d_ci, L->d_dslot, LU, x,
                           L->d_w, y);
    }
    for (int c = nc - 2; c >= 0; --c) {
        const int r0 = L->cstart[c], r1 = L->cstart[c + 1];
        hipLaunchKernelGGL(k_usolve_color<TM>, dim3(grid_of(int64_t(r1 - r0) * kRowLanes)), dim3(kThreads), 0, s, r0, r1,
                           c1, L->d_rp, L->d_ci, L->d_dslot, LU, x, L->d_w, y);
    }
}

void precondition(uvh_lin *L, const double *x, double *y, hipStream_t s) {
    if (L->lu32)
        precondition_t(L, L->d_LUf, x, y, s);
    else
        precondition_t(L, L->d_LU, x, y, s);
}

void block_mv(uvh_lin *L, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_block_mv, dim3(grid_of(int64_t(L->nn) * kRowLanes)), dim3(kThreads), 0, s, L->nn, L->d_rp,
                       L->d_ci, L->d_A, x, y);
}

// BiCGStab from x = 0 (d_b the right-hand side). Near-breakdown (rhat nearly orthogonal to
// r or v, or omega = 0) restarts with the current residual as the new shadow vector, at most
// kMaxRestarts times. The iterate with the smallest residual is kept and returned when the
// tolerance is not reached, so a diverging run cannot hand back a worse vector than it saw.
// The scalars live on the device (k_bi_*); the host keeps one iteration in flight ahead of the
// one whose stop word it reads, and relaunches that iteration after a restart (it ran as a
// no-op behind the stop word).
int bicgstab(uvh_lin *L, double rtol, int maxit, double bnorm, hipStream_t s, int &it, bool &converged,
             std::string &why) {
    const int64_t n = L->n3;
    const dim3 G(kRedBlocks), B(kThreads);
    hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_b, L->d_r);
    hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_b, L->d_rhat);
    hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_p);
    hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_v);
    hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_xbest);
    hipLaunchKernelGGL(k_bi_init, dim3(1), dim3(1), 0, s, L->d_sc, L->d_flag, bnorm, rtol * bnorm);
    double *sc = L->d_sc;
    int *flag = L->d_flag;
    int launched = 0, done = 0, restarts = 0;
    auto launch = [&]() {
        hipLaunchKernelGGL(k_bi_p, G, B, 0, s, n, L->d_r, L->d_v, L->d_p, L->d_x, L->d_xbest, sc, flag);
        precondition(L, L->d_p, L->d_y, s);
        block_mv(L, L->d_y, L->d_v, s);
        hipLaunchKernelGGL(k_dots<2>, G, B, 0, s, n, L->d_rhat, L->d_v, L->d_v, L->d_v, L->d_part);
        hipLaunchKernelGGL(k_bi_finish, dim3(1), B, 0, s, kRedBlocks, 2, L->d_part, sc, int(S_RV), 0, flag);
        hipLaunchKernelGGL(k_bi_s, G, B, 0, s, n, L->d_r, L->d_v, L->d_y, L->d_s, L->d_x, L->d_part, sc, flag);
        hipLaunchKernelGGL(k_bi_finish, dim3(1), B, 0, s, kRedBlocks, 1, L->d_part, sc, int(S_SS), 1, flag);
        precondition(L, L->d_s, L->d_z, s);
        block_mv(L, L->d_z, L->d_t, s);
        hipLaunchKernelGGL(k_dots<2>, G, B, 0, s, n, L->d_t, L->d_s, L->d_t, L->d_t, L->d_part);
        hipLaunchKernelGGL(k_bi_finish, dim3(1), B, 0, s, kRedBlocks, 2, L->d_part, sc, int(S_TS), 0, flag);
        hipLaunchKernelGGL(k_bi_xr, G, B, 0, s, n, L->d_s, L->d_t, L->d_z, L->d_rhat, L->d_x, L->d_r, L->d_xbest,
                           L->d_part, sc, flag);
        hipLaunchKernelGGL(k_bi_finish, dim3(1), B, 0, s, kRedBlocks, 2, L->d_part, sc, int(S_RHATR), 2, flag);
        ++launched;
        const int slot = launched & 1;
        hipError_t e = hipMemcpyAsync(L->h_flag + slot, flag + F_STOP, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(L->ev[slot], s);
        return e;
    };
    auto fail = [&](hipError_t e) {
        return L->fail(UVH_E_HIP, std::string("uvh_lin_solve: BiCGStab: ") + hipGetErrorString(e));
    };
    converged = false;
    hipError_t e = launch();
    if (e != hipSuccess) return fail(e);
    for (;;) {
        const int cur = done + 1;   // launched >= cur
        if (launched == cur && launched < maxit && (e = launch()) != hipSuccess) return fail(e);
        if ((e = hipEventSynchronize(L->ev[cur & 1])) != hipSuccess) return fail(e);
        const int word = L->h_flag[cur & 1];
        done = cur;
        if (word == B_RUN) {
            if (done == maxit) break;
            continue;
        }
        if (word == B_CONV_S || word == B_CONV_R) {
            converged = true;
            break;
        }
        // a restart: after a breakdown inside iteration `done` (it counts as done), or before
        // the next one; not past the last iteration unless omega = 0 ended this one
        if (word == B_BREAK_R && done == maxit) break;
        if (restarts == kMaxRestarts) {
            why = word == B_BREAK_V ? "BiCGStab breakdown: (rhat, v) = 0"
                  : word == B_BREAK_R ? "BiCGStab breakdown: (rhat, r) = 0"
                                      : "BiCGStab breakdown: omega = 0";
            break;
        }
        hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_r, L->d_rhat);
        hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_p);
        hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_v);
        hipLaunchKernelGGL(k_bi_restart, dim3(1), dim3(1), 0, s, sc, flag);
        ++restarts;
        launched = done;   // an iteration launched ahead ran as a no-op: launch it again
        if (done == maxit) break;
        if ((e = launch()) != hipSuccess) return fail(e);
    }
    it = done;
    if (!converged) {
        if (why.empty()) why = "BiCGStab iteration limit";
        hipLaunchKernelGGL(k_copy_if, G, B, 0, s, n, L->d_x, L->d_xbest, flag);   // a better last iterate
        hipLaunchKernelGGL(k_copy, G, B, 0, s, n, L->d_xbest, L->d_x);
    }
    return UVH_OK;
}

// GMRES(m) with right preconditioning from the current d_x, m = L->restart: Arnoldi with
// classical Gram-Schmidt and one re-orthogonalisation (one host synchronisation per step),
// Givens rotations on the host. `it` counts Arnoldi steps (one preconditioner application
// and one product each).
int gmres(uvh_lin *L, double rtol, int maxit, double bnorm, hipStream_t s, int &it, bool &converged,
          std::string &why) {
    const int64_t n = L->n3;
    const int m = L->restart;
    const dim3 G(kRedBlocks), B(kThreads);
    double red[2];
    int st;
    std::vector<double> H(size_t(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    double *h1 = L->h_out, *h2 = L->h_out + (kMaxKrylov + 2), *hw = L->h_out + 2 * (kMaxKrylov + 2);
    double prev = HUGE_VAL;
    it = 0;
    converged = false;
    while (it < maxit) {
        // r = b - A x into V_0
        block_mv(L, L->d_x, L->d_w, s);
        hipLaunchKernelGGL(k_residual, G, B, 0, s, n, L->d_b, L->d_w, L->d_part);
        if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
        const double beta = std::sqrt(red[0]);
        if (beta <= rtol * bnorm) { converged = true; break; }
        if (!(beta < prev * (1.0 - 1e-12))) { why = "GMRES stagnated"; break; }
        prev = beta;
        hipLaunchKernelGGL(k_scale, G, B, 0, s, n, 1.0 / beta, L->d_w, L->d_V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;   // basis vectors used for the update
        for (int j = 0; j < m && it < maxit; ++j) {
            ++it;
            double *vj = L->d_V + int64_t(j) * n, *vn = L->d_V + int64_t(j + 1) * n;
            precondition(L, vj, L->d_z, s);
            block_mv(L, L->d_z, vn, s);
            // two passes of classical Gram-Schmidt against V_0 .. V_j
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL(k_multi_dot, dim3(kRedBlocks, j + 1), B, 0, s, n, L->d_V, vn, L->d_part);
                hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, s, kRedBlocks, j + 1, L->d_part, L->d_out);
                hipLaunchKernelGGL(k_multi_axpy, G, B, 0, s, n, j + 1, L->d_V, L->d_out, -1.0, vn);
                hipMemcpyAsync(pass ? h2 : h1, L->d_out, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s);
            }
            hipLaunchKernelGGL(k_dots<1>, G, B, 0, s, n, vn, vn, nullptr, nullptr, L->d_part);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, s, kRedBlocks, 1, L->d_part, L->d_out);
            hipMemcpyAsync(hw, L->d_out, sizeof(double), hipMemcpyDeviceToHost, s);
            if (hipStreamSynchronize(s) != hipSuccess) return L->fail(UVH_E_HIP, "uvh_lin_solve: GMRES step");
            double *col = &H[size_t(j) * (m + 1)];   // column j, m + 1 rows
            for (int c = 0; c <= j; ++c) col[c] = h1[c] + h2[c];
            col[j + 1] = std::sqrt(hw[0]);
            const double hn = col[j + 1];
            if (hn > 0.0) hipLaunchKernelGGL(k_scale, G, B, 0, s, n, 1.0 / hn, vn, vn);
            for (int c = 0; c < j; ++c) {   // earlier rotations
                const double a = col[c], b = col[c + 1];
                col[c] = cs[c] * a + sn[c] * b;
                col[c + 1] = -sn[c] * a + cs[c] * b;
            }
            const double r = std::hypot(col[j], col[j + 1]);
            cs[j] = r > 0.0 ? col[j] / r : 1.0;
            sn[j] = r > 0.0 ? col[j + 1] / r : 0.0;
            col[j] = r;
            col[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            k = j + 1;
            if (std::fabs(g[j + 1]) <= rtol * bnorm || hn == 0.0) break;
        }
        // y = H(0:k, 0:k)^-1 g, x += M^-1 V y
        for (int r = k - 1; r >= 0; --r) {
            double t = g[r];
            for (int c = r + 1; c < k; ++c) t -= H[size_t(c) * (m + 1) + r] * y[c];
            const double d = H[size_t(r) * (m + 1) + r];
            y[r] = d != 0.0 ? t / d : 0.0;
        }
        std::copy(y.begin(), y.begin() + k, h1);
        hipMemcpyAsync(L->d_coef, h1, k * sizeof(double), hipMemcpyHostToDevice, s);
        hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_y);
        hipLaunchKernelGGL(k_multi_axpy, G, B, 0, s, n, k, L->d_V, L->d_coef, 1.0, L->d_y);
        precondition(L, L->d_y, L->d_z, s);
        hipLaunchKernelGGL(k_axpy, G, B, 0, s, n, 1.0, L->d_z, L->d_x);
    }
    if (!converged && why.empty()) why = "GMRES iteration limit";
    return UVH_OK;
}

void destroy(uvh_lin *L) {
    if (!L) return;
    hipSetDevice(L->device);
    for (void *p : {(void *)L->d_rowptr64, (void *)L->d_colind, (void *)L->d_newnode, (void *)L->d_slot,
                    (void *)L->d_rp, (void *)L->d_ci, (void *)L->d_node_of, (void *)L->d_fixed,
                    (void *)L->d_A, (void *)L->d_LU, (void *)L->d_LUf, (void *)L->d_b, (void *)L->d_x, (void *)L->d_r, (void *)L->d_rhat,
                    (void *)L->d_p, (void *)L->d_v, (void *)L->d_s, (void *)L->d_t, (void *)L->d_y, (void *)L->d_z,
                    (void *)L->d_w, (void *)L->d_part, (void *)L->d_out, (void *)L->d_dslot, (void *)L->d_status,
                    (void *)L->d_xbest, (void *)L->d_V, (void *)L->d_coef, (void *)L->d_sc, (void *)L->d_flag})
        if (p) hipFree(p);
    for (hipEvent_t e : L->ev)
        if (e) hipEventDestroy(e);
    if (L->h_flag) hipHostFree(L->h_flag);
    if (L->h_out) hipHostFree(L->h_out);
    if (L->h_status) hipHostFree(L->h_status);
    delete L;
}

}  // namespace

extern "C" {

int uvh_lin_create(uvh_ctx *ctx, const uint8_t *fixed, uvh_lin **out) {
    if (!out) return UVH_E_ARG;
    *out = nullptr;
    const uvh::MeshHost *M = uvh::ctx_mesh(ctx);
    if (!M) return UVH_E_STATE;
    if (M->nnzN() >= (int64_t(1) << 31) - 1) return UVH_E_ARG;   // 32-bit block indices
    auto *L = new uvh_lin();
    L->device = uvh::ctx_device(ctx);
    hipSetDevice(L->device);
    // UVH_LIN_FACTOR=fp64 keeps the triangular solves on the FP64 factors
    const char *fac = std::getenv("UVH_LIN_FACTOR");
    L->lu32 = !(fac && std::string(fac) == "fp64");
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
    L->cstart = cstart;
    for (int n = 0; n < nn; ++n) {
        newnode[n] = cstart[color[n]]++;
        oldnode[newnode[n]] = n;
    }
    // renumbered block pattern (columns sorted) and the slot of every original entry
    std::vector<int32_t> rp32(nn + 1, 0), ci32(L->nnzN), slot(L->nnzN), node_of(L->nnzN), dslot(nn, -1);
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
            if (row[q].first == r) dslot[r] = rp32[r] + int32_t(q);
        }
        if (dslot[r] < 0) {   // the node graph holds every node's own entry
            delete L;
            return UVH_E_ARG;
        }
    }
    bool ok = lalloc(&L->d_rowptr64, nn + 1) == hipSuccess && lalloc(&L->d_rp, nn + 1) == hipSuccess &&
              lalloc(&L->d_ci, L->nnzN) == hipSuccess && lalloc(&L->d_node_of, L->nnzN) == hipSuccess &&
              lalloc(&L->d_colind, L->nnzN) == hipSuccess && lalloc(&L->d_slot, L->nnzN) == hipSuccess &&
              lalloc(&L->d_newnode, nn) == hipSuccess && lalloc(&L->d_dslot, nn) == hipSuccess &&
              lalloc(&L->d_status, 1) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&L->h_status), sizeof(int), 0) == hipSuccess &&
              lalloc(&L->d_A, 9 * L->nnzN) == hipSuccess && lalloc(&L->d_LU, 9 * L->nnzN) == hipSuccess &&
              (!L->lu32 || lalloc(&L->d_LUf, 9 * L->nnzN) == hipSuccess) &&
              lalloc(&L->d_part, (kMaxKrylov + 2) * kRedBlocks) == hipSuccess &&
              lalloc(&L->d_out, kMaxKrylov + 2) == hipSuccess && lalloc(&L->d_coef, kMaxKrylov + 2) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&L->h_out), 3 * (kMaxKrylov + 2) * sizeof(double), 0) ==
                  hipSuccess &&
              lalloc(&L->d_sc, S_COUNT) == hipSuccess && lalloc(&L->d_flag, F_COUNT) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&L->h_flag), 2 * sizeof(int), 0) == hipSuccess &&
              hipEventCreateWithFlags(&L->ev[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&L->ev[1], hipEventDisableTiming) == hipSuccess;
    for (double **v : {&L->d_b, &L->d_x, &L->d_r, &L->d_rhat, &L->d_p, &L->d_v, &L->d_s, &L->d_t, &L->d_y, &L->d_z,
                       &L->d_w, &L->d_xbest})
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
    if (e == hipSuccess) e = hipMemcpy(L->d_dslot, dslot.data(), nn * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(L->d_node_of, node_of.data(), L->nnzN * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && fixed) e = hipMemcpy(L->d_fixed, fixed, L->n3, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
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
    // block ILU(0), colour by colour
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, L->d_status, L->nn);
    const int nc = int(L->cstart.size()) - 1;
    for (int c = 0; c < nc; ++c) {
        const int r0 = L->cstart[c], r1 = L->cstart[c + 1];
        hipLaunchKernelGGL(k_ilu_color, dim3(grid_of(r1 - r0)), dim3(kThreads), 0, s, r0, r1, L->d_rp, L->d_ci,
                           L->d_dslot, L->d_LU, L->d_status);
    }
    hipError_t e = hipMemcpyAsync(L->h_status, L->d_status, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return L->fail(UVH_E_HIP, std::string("uvh_lin_factor: ") + hipGetErrorString(e));
    if (*L->h_status < L->nn) {
        L->factored = false;
        return L->fail(UVH_E_NOCONV, "uvh_lin_factor: singular pivot block in (renumbered) block row " +
                                         std::to_string(*L->h_status));
    }
    if (L->lu32) {
        hipLaunchKernelGGL(k_to_f32, dim3(kRedBlocks * 4), dim3(kThreads), 0, s, total, L->d_LU, L->d_LUf);
        if ((e = hipGetLastError()) != hipSuccess)
            return L->fail(UVH_E_HIP, std::string("uvh_lin_factor: ") + hipGetErrorString(e));
    }
    L->factored = true;
    return UVH_OK;
}

int uvh_lin_set_krylov(uvh_lin *L, int32_t method, int32_t restart) {
    if (!L) return UVH_E_ARG;
    if (method < UVH_LIN_BICGSTAB || method > UVH_LIN_AUTO)
        return L->fail(UVH_E_ARG, "uvh_lin_set_krylov: unknown method " + std::to_string(method));
    if (restart < 2 || restart > kMaxKrylov)
        return L->fail(UVH_E_ARG, "uvh_lin_set_krylov: restart must be in [2, " + std::to_string(kMaxKrylov) + "]");
    if (L->d_V && restart > L->restart) {   // the basis grows
        hipFree(L->d_V);
        L->d_V = nullptr;
    }
    L->method = method;
    L->restart = restart;
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
    const int64_t n = L->n3;
    const dim3 G(kRedBlocks), B(kThreads);
    if (L->method != UVH_LIN_BICGSTAB && !L->d_V && lalloc(&L->d_V, size_t(L->restart + 1) * n) != hipSuccess)
        return L->fail(UVH_E_NOMEM, "uvh_lin_solve: cannot allocate the GMRES basis");
    double red[2];
    int st;
    hipLaunchKernelGGL(k_gather_rhs, dim3(grid_of(n)), B, 0, s, L->nn, L->d_newnode, L->d_fixed, drhs, L->d_b);
    hipLaunchKernelGGL(k_fill, G, B, 0, s, n, 0.0, L->d_x);
    hipLaunchKernelGGL(k_dots<1>, G, B, 0, s, n, L->d_b, L->d_b, nullptr, nullptr, L->d_part);
    if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
    const double bnorm = std::sqrt(red[0]);
    int it = 0;
    double res = 0.0;
    bool converged = bnorm == 0.0;
    std::string why;
    if (!converged) {
        if (L->method != UVH_LIN_GMRES) {
            st = bicgstab(L, rtol, maxit, bnorm, s, it, converged, why);
            if (st != UVH_OK) return st;
        }
        if (!converged && L->method != UVH_LIN_BICGSTAB && it < maxit) {
            int git = 0;
            st = gmres(L, rtol, maxit - it, bnorm, s, git, converged, why);
            if (st != UVH_OK) return st;
            it += git;
        }
        // the true residual of the returned solution
        block_mv(L, L->d_x, L->d_w, s);
        hipLaunchKernelGGL(k_residual, G, B, 0, s, n, L->d_b, L->d_w, L->d_part);
        if ((st = finish(L, 1, s, red)) != UVH_OK) return st;
        res = std::sqrt(red[0]) / bnorm;
    }
    hipLaunchKernelGGL(k_scatter_sol, dim3(grid_of(n)), B, 0, s, L->nn, L->d_newnode, L->d_x, dx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return L->fail(UVH_E_HIP, std::string("uvh_lin_solve: ") + hipGetErrorString(e));
    if (iterations) *iterations = it;
    if (relres) *relres = res;
    if (!converged)
        return L->fail(UVH_E_NOCONV, "uvh_lin_solve: relative residual " + std::to_string(res) + " after " +
                                         std::to_string(it) + " iterations" + (why.empty() ? "" : " (" + why + ")"));
    return UVH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Batched solves: the BiCGStab iterations of many members' systems (each a uvh_lin, factored
// on its own) in one launch per step, over the flat work lists built at creation.
struct uvhb_lin {
    std::string err;
    int device = 0, n = 0, maxcol = 0;
    std::vector<uvh_lin *> mem;
    LinMember *d_members = nullptr;
    int2 *d_vitems = nullptr, *d_ritems = nullptr;
    int4 *d_litems = nullptr, *d_uitems = nullptr;
    int *d_nrows = nullptr, *d_flags = nullptr, *d_words = nullptr;
    int nv = 0, nr = 0;
    std::vector<int> loff, uoff;   // colour c's lower / upper items: [off[c], off[c+1])
    int *h_words = nullptr;        // pinned: 2 n stop words (two iterations in flight)
    double *h_red = nullptr;       // pinned: n values
    int *h_its = nullptr;          // pinned: n iteration counts
    int *h_flags = nullptr;        // pinned: the members' flag words at the end of a solve
    hipEvent_t ev[2] = {nullptr, nullptr};
    // the solve in progress (uvhb_lin_begin .. uvhb_lin_end)
    std::vector<double> bnorm;
    std::vector<uint8_t> act;
    bool begun = false;
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

namespace {

void bdestroy(uvhb_lin *B) {
    if (!B) return;
    hipSetDevice(B->device);
    for (void *p : {(void *)B->d_members, (void *)B->d_vitems, (void *)B->d_ritems, (void *)B->d_litems,
                    (void *)B->d_uitems, (void *)B->d_nrows, (void *)B->d_flags, (void *)B->d_words})
        if (p) hipFree(p);
    for (hipEvent_t e : B->ev)
        if (e) hipEventDestroy(e);
    for (void *p : {(void *)B->h_words, (void *)B->h_red, (void *)B->h_its, (void *)B->h_flags})
        if (p) hipHostFree(p);
    delete B;
}

// one BiCGStab iteration of every member (launches only)
void biter_launch(uvhb_lin *B, hipStream_t s) {
    const dim3 T(kThreads);
    const int n = B->n;
    hipLaunchKernelGGL(k_bb_p, dim3(B->nv), T, 0, s, B->d_members, B->d_vitems);
    for (int which = 0; which < 2; ++which) {
        if (which == 1) {
            hipLaunchKernelGGL(k_bb_s, dim3(B->nv), T, 0, s, B->d_members, B->d_vitems);
            hipLaunchKernelGGL(k_bb_finish, dim3(n), T, 0, s, B->d_members, 1, int(S_SS), 1);
        }
        for (int c = 0; c < B->maxcol; ++c)
            if (B->loff[c + 1] > B->loff[c])
                hipLaunchKernelGGL(k_bb_lsolve, dim3(B->loff[c + 1] - B->loff[c]), T, 0, s, B->d_members,
                                   B->d_litems + B->loff[c], which);
        for (int c = B->maxcol - 1; c >= 0; --c)
            if (B->uoff[c + 1] > B->uoff[c])
                hipLaunchKernelGGL(k_bb_usolve, dim3(B->uoff[c + 1] - B->uoff[c]), T, 0, s, B->d_members,
                                   B->d_uitems + B->uoff[c], which);
        hipLaunchKernelGGL(k_bb_mv, dim3(B->nr), T, 0, s, B->d_members, B->d_ritems, B->d_nrows, which);
        hipLaunchKernelGGL(k_bb_dots, dim3(B->nv), T, 0, s, B->d_members, B->d_vitems, which);
        hipLaunchKernelGGL(k_bb_finish, dim3(n), T, 0, s, B->d_members, 2, which ? int(S_TS) : int(S_RV),
                           which ? 3 : 0);
    }
    hipLaunchKernelGGL(k_bb_xr, dim3(B->nv), T, 0, s, B->d_members, B->d_vitems);
    hipLaunchKernelGGL(k_bb_finish, dim3(n), T, 0, s, B->d_members, 2, int(S_RHATR), 2);
}

// the members' stop words into pinned slot `slot`
hipError_t bwords(uvhb_lin *B, int slot, hipStream_t s) {
    const int n = B->n;
    hipLaunchKernelGGL(k_bb_words, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, B->d_members, n,
                       B->d_words);
    return hipMemcpyAsync(B->h_words + slot * n, B->d_words, n * sizeof(int), hipMemcpyDeviceToHost, s);
}

// one iteration of every member, then its stop words into slot `slot` and that slot's event
hipError_t biter(uvhb_lin *B, int slot, hipStream_t s) {
    biter_launch(B, s);
    hipError_t e = bwords(B, slot, s);
    if (e == hipSuccess) e = hipEventRecord(B->ev[slot], s);
    return e;
}

// where a member missed the tolerance: x = xbest, after a copy of x itself if its last residual
// check asked for one; nothing for a converged or idle member
__global__ void k_bb_best(int64_t n, double *__restrict__ x, double *__restrict__ xbest, const int *__restrict__ flag) {
    const int w = flag[F_STOP];
    if (w == B_CONV_S || w == B_CONV_R || w == B_IDLE) return;
    const bool cp = flag[F_COPY] != 0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        if (cp)
            xbest[i] = x[i];
        else
            x[i] = xbest[i];
    }
}

bool word_terminal(int w) { return w == B_CONV_S || w == B_CONV_R || w >= B_FAIL; }

}  // namespace

extern "C" {

int uvhb_lin_create(uvh_lin *const *members, int32_t n, uvhb_lin **out) {
    if (!out || !members || n < 1) return UVH_E_ARG;
    *out = nullptr;
    auto *B = new uvhb_lin();
    B->n = n;
    B->mem.assign(members, members + n);
    for (uvh_lin *L : B->mem)
        if (!L || L->device != B->mem[0]->device) {
            delete B;
            return UVH_E_ARG;
        }
    B->device = B->mem[0]->device;
    hipSetDevice(B->device);
    std::vector<LinMember> tab(n);
    std::vector<int2> vit, rit;
    std::vector<int> nrows(n);
    for (int m = 0; m < n; ++m) B->maxcol = std::max(B->maxcol, int(B->mem[m]->cstart.size()) - 1);
    std::vector<std::vector<int4>> lit(B->maxcol), uit(B->maxcol);
    bool ok = lalloc(&B->d_flags, int64_t(n) * F_BWORDS) == hipSuccess;
    for (int m = 0; m < n && ok; ++m) {
        uvh_lin *L = B->mem[m];
        LinMember &T = tab[m];
        T.rp = L->d_rp;
        T.ci = L->d_ci;
        T.dslot = L->d_dslot;
        T.A = L->d_A;
        T.LU = L->lu32 ? static_cast<const void *>(L->d_LUf) : static_cast<const void *>(L->d_LU);
        T.r = L->d_r; T.rhat = L->d_rhat; T.p = L->d_p; T.v = L->d_v; T.s = L->d_s; T.t = L->d_t;
        T.y = L->d_y; T.z = L->d_z; T.w = L->d_w; T.x = L->d_x; T.xbest = L->d_xbest;
        T.part = L->d_part;
        T.sc = L->d_sc;
        T.flag = B->d_flags + int64_t(m) * F_BWORDS;
        T.n3 = L->n3;
        T.nchunk = int((L->n3 + kChunk - 1) / kChunk);
        T.lu32 = L->lu32 ? 1 : 0;
        if (2 * int64_t(T.nchunk) > int64_t(kMaxKrylov + 2) * kRedBlocks) {   // the member's partial slots
            bdestroy(B);
            return UVH_E_ARG;
        }
        nrows[m] = L->nn;
        for (int c = 0; c < T.nchunk; ++c) vit.push_back(make_int2(m, c));
        for (int r0 = 0; r0 < L->nn; r0 += kRowsPerItem) rit.push_back(make_int2(m, r0));
        const std::vector<int32_t> &cs = L->cstart;
        const int nc = int(cs.size()) - 1, c1 = cs[1];
        for (int c = nc == 1 ? 0 : 1; c < nc; ++c)
            for (int r0 = cs[c]; r0 < cs[c + 1]; r0 += kRowsPerItem)
                lit[c].push_back(make_int4(m, r0, cs[c + 1], int(c == nc - 1) + 2 * (c == 0 ? 0 : c1)));
        for (int c = nc - 2; c >= 0; --c)
            for (int r0 = cs[c]; r0 < cs[c + 1]; r0 += kRowsPerItem) uit[c].push_back(make_int4(m, r0, cs[c + 1], c1));
    }
    std::vector<int4> lflat, uflat;
    B->loff.assign(1, 0);
    B->uoff.assign(1, 0);
    for (int c = 0; c < B->maxcol; ++c) {
        lflat.insert(lflat.end(), lit[c].begin(), lit[c].end());
        uflat.insert(uflat.end(), uit[c].begin(), uit[c].end());
        B->loff.push_back(int(lflat.size()));
        B->uoff.push_back(int(uflat.size()));
    }
    B->nv = int(vit.size());
    B->nr = int(rit.size());
    ok = ok && lalloc(&B->d_members, n) == hipSuccess && lalloc(&B->d_vitems, vit.size()) == hipSuccess &&
         lalloc(&B->d_ritems, rit.size()) == hipSuccess && lalloc(&B->d_litems, lflat.size()) == hipSuccess &&
         lalloc(&B->d_uitems, uflat.size()) == hipSuccess && lalloc(&B->d_nrows, n) == hipSuccess &&
         lalloc(&B->d_words, n) == hipSuccess &&
         hipHostMalloc(reinterpret_cast<void **>(&B->h_words), 2 * n * sizeof(int), 0) == hipSuccess &&
         hipHostMalloc(reinterpret_cast<void **>(&B->h_red), n * sizeof(double), 0) == hipSuccess &&
         hipHostMalloc(reinterpret_cast<void **>(&B->h_its), n * sizeof(int), 0) == hipSuccess &&
         hipHostMalloc(reinterpret_cast<void **>(&B->h_flags), size_t(n) * F_BWORDS * sizeof(int), 0) == hipSuccess &&
         hipEventCreateWithFlags(&B->ev[0], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&B->ev[1], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        bdestroy(B);
        return UVH_E_NOMEM;
    }
    hipError_t e = hipMemcpy(B->d_members, tab.data(), n * sizeof(LinMember), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_vitems, vit.data(), vit.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_ritems, rit.data(), rit.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e == hipSuccess && !lflat.empty())
        e = hipMemcpy(B->d_litems, lflat.data(), lflat.size() * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess && !uflat.empty())
        e = hipMemcpy(B->d_uitems, uflat.data(), uflat.size() * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_nrows, nrows.data(), n * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        bdestroy(B);
        return UVH_E_HIP;
    }
    *out = B;
    return UVH_OK;
}

void uvhb_lin_destroy(uvhb_lin *B) { bdestroy(B); }

const char *uvhb_lin_last_error(const uvhb_lin *B) { return B ? B->err.c_str() : "null batch solver"; }

int uvhb_lin_begin(uvhb_lin *B, const double *const *drhs, const uint8_t *active, double rtol, int32_t maxit,
                   void *stream) {
    if (!B) return UVH_E_ARG;
    if (!drhs) return B->fail(UVH_E_ARG, "uvhb_lin_begin: null right-hand sides");
    if (!(rtol > 0.0) || maxit < 1) return B->fail(UVH_E_ARG, "uvhb_lin_begin: rtol must be > 0 and maxit >= 1");
    const int n = B->n;
    B->act.assign(n, 1);
    for (int m = 0; m < n; ++m) {
        if (active && !active[m]) {
            B->act[m] = 0;
            continue;
        }
        if (!drhs[m]) return B->fail(UVH_E_ARG, "uvhb_lin_begin: null right-hand side of member " + std::to_string(m));
        if (!B->mem[m]->factored)
            return B->fail(UVH_E_STATE, "uvhb_lin_begin: member " + std::to_string(m) + " has no factorization");
    }
    hipSetDevice(B->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 G(kRedBlocks), T(kThreads);
    // right-hand sides in the solvers' order and their norms
    for (int m = 0; m < n; ++m) {
        if (!B->act[m]) continue;
        uvh_lin *L = B->mem[m];
        hipLaunchKernelGGL(k_gather_rhs, dim3(grid_of(L->n3)), T, 0, s, L->nn, L->d_newnode, L->d_fixed, drhs[m], L->d_b);
        hipLaunchKernelGGL(k_fill, G, T, 0, s, L->n3, 0.0, L->d_x);
        hipLaunchKernelGGL(k_dots<1>, G, T, 0, s, L->n3, L->d_b, L->d_b, nullptr, nullptr, L->d_part);
        hipLaunchKernelGGL(k_finish, dim3(1), T, 0, s, kRedBlocks, 1, L->d_part, L->d_out);
        hipMemcpyAsync(B->h_red + m, L->d_out, sizeof(double), hipMemcpyDeviceToHost, s);
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return B->fail(UVH_E_HIP, std::string("uvhb_lin_begin: ") + hipGetErrorString(e));
    B->bnorm.assign(n, 0.0);
    for (int m = 0; m < n; ++m) {
        uvh_lin *L = B->mem[m];
        int *flag = B->d_flags + int64_t(m) * F_BWORDS;
        B->bnorm[m] = B->act[m] ? std::sqrt(B->h_red[m]) : 0.0;
        if (!B->act[m]) {
            hipLaunchKernelGGL(k_bb_start, dim3(1), dim3(1), 0, s, flag, int(B_IDLE), maxit);
        } else if (B->bnorm[m] == 0.0) {
            hipLaunchKernelGGL(k_bb_start, dim3(1), dim3(1), 0, s, flag, int(B_CONV_R), maxit);
        } else {
            hipLaunchKernelGGL(k_copy, G, T, 0, s, L->n3, L->d_b, L->d_r);
            hipLaunchKernelGGL(k_copy, G, T, 0, s, L->n3, L->d_b, L->d_rhat);
            hipLaunchKernelGGL(k_fill, G, T, 0, s, L->n3, 0.0, L->d_p);
            hipLaunchKernelGGL(k_fill, G, T, 0, s, L->n3, 0.0, L->d_v);
            hipLaunchKernelGGL(k_fill, G, T, 0, s, L->n3, 0.0, L->d_xbest);
            hipLaunchKernelGGL(k_bb_start, dim3(1), dim3(1), 0, s, flag, int(B_RUN), maxit);
            hipLaunchKernelGGL(k_bi_init, dim3(1), dim3(1), 0, s, L->d_sc, flag, B->bnorm[m], rtol * B->bnorm[m]);
        }
    }
    if ((e = hipGetLastError()) != hipSuccess)
        return B->fail(UVH_E_HIP, std::string("uvhb_lin_begin: ") + hipGetErrorString(e));
    B->begun = true;
    return UVH_OK;
}

int uvhb_lin_iterate(uvhb_lin *B, int32_t iterations, void *stream) {
    if (!B) return UVH_E_ARG;
    if (!B->begun) return B->fail(UVH_E_STATE, "uvhb_lin_iterate: no solve begun (uvhb_lin_begin)");
    if (iterations < 1) return B->fail(UVH_E_ARG, "uvhb_lin_iterate: iterations must be >= 1");
    hipSetDevice(B->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int k = 0; k < iterations; ++k) biter_launch(B, s);
    hipError_t e = bwords(B, 0, s);
    if (e != hipSuccess) return B->fail(UVH_E_HIP, std::string("uvhb_lin_iterate: ") + hipGetErrorString(e));
    return UVH_OK;
}

int uvhb_lin_stopped(const uvhb_lin *B) {
    if (!B) return UVH_E_ARG;
    for (int m = 0; m < B->n; ++m)
        if (!word_terminal(B->h_words[m])) return 0;
    return 1;
}

int uvhb_lin_end(uvhb_lin *B, double *const *dx, int32_t *iterations, double *relres, int32_t *converged,
                 void *stream) {
    if (!B) return UVH_E_ARG;
    if (!B->begun) return B->fail(UVH_E_STATE, "uvhb_lin_end: no solve begun (uvhb_lin_begin)");
    if (!dx) return B->fail(UVH_E_ARG, "uvhb_lin_end: null solutions");
    const int n = B->n;
    for (int m = 0; m < n; ++m)
        if (B->act[m] && !dx[m]) return B->fail(UVH_E_ARG, "uvhb_lin_end: null solution of member " + std::to_string(m));
    hipSetDevice(B->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 G(kRedBlocks), T(kThreads);
    // best iterate where the tolerance was missed, the true residual, the solution in [ub; vb; h]
    for (int m = 0; m < n; ++m) {
        if (!B->act[m]) continue;
        uvh_lin *L = B->mem[m];
        int *flag = B->d_flags + int64_t(m) * F_BWORDS;
        if (B->bnorm[m] > 0.0) {
            hipLaunchKernelGGL(k_bb_best, G, T, 0, s, L->n3, L->d_x, L->d_xbest, flag);
            block_mv(L, L->d_x, L->d_w, s);
            hipLaunchKernelGGL(k_residual, G, T, 0, s, L->n3, L->d_b, L->d_w, L->d_part);
            hipLaunchKernelGGL(k_finish, dim3(1), T, 0, s, kRedBlocks, 1, L->d_part, L->d_out);
            hipMemcpyAsync(B->h_red + m, L->d_out, sizeof(double), hipMemcpyDeviceToHost, s);
        }
        hipLaunchKernelGGL(k_scatter_sol, dim3(grid_of(L->n3)), T, 0, s, L->nn, L->d_newnode, L->d_x, dx[m]);
    }
    hipMemcpyAsync(B->h_flags, B->d_flags, size_t(n) * F_BWORDS * sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    B->begun = false;
    if (e != hipSuccess) return B->fail(UVH_E_HIP, std::string("uvhb_lin_end: ") + hipGetErrorString(e));
    int nfail = 0;
    for (int m = 0; m < n; ++m) {
        const int *fl = B->h_flags + int64_t(m) * F_BWORDS;
        const bool live = B->act[m] && B->bnorm[m] > 0.0;
        const bool conv = B->act[m] && (B->bnorm[m] == 0.0 || fl[F_STOP] == B_CONV_S || fl[F_STOP] == B_CONV_R);
        if (iterations) iterations[m] = live ? fl[F_ITS] : 0;
        if (relres) relres[m] = live ? std::sqrt(B->h_red[m]) / B->bnorm[m] : 0.0;
        if (converged) converged[m] = conv ? 1 : 0;
        if (B->act[m] && !conv) ++nfail;
    }
    if (nfail)
        return B->fail(UVH_E_NOCONV, "uvhb_lin_end: " + std::to_string(nfail) + " member(s) missed the tolerance");
    return UVH_OK;
}

int uvhb_lin_solve(uvhb_lin *B, const double *const *drhs, double *const *dx, const uint8_t *active, double rtol,
                   int32_t maxit, int32_t *iterations, double *relres, int32_t *converged, void *stream) {
    if (!B) return UVH_E_ARG;
    if (!dx) return B->fail(UVH_E_ARG, "uvhb_lin_solve: null solutions");
    int st = uvhb_lin_begin(B, drhs, active, rtol, maxit, stream);
    if (st != UVH_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // iterations, one in flight ahead of the stop words being read; every member stops on the
    // device (tolerance, breakdown past the restarts, or the iteration limit), so at most
    // maxit + kMaxRestarts + 1 iterations are needed
    const int n = B->n, cap = maxit + kMaxRestarts + 2;
    int launched = 1, done = 0;
    hipError_t e = biter(B, 1, s);
    while (e == hipSuccess) {
        const int cur = done + 1;
        if (launched == cur && launched < cap) {
            if ((e = biter(B, (launched + 1) & 1, s)) != hipSuccess) break;
            ++launched;
        }
        if ((e = hipEventSynchronize(B->ev[cur & 1])) != hipSuccess) break;
        done = cur;
        bool all = true;
        for (int m = 0; m < n && all; ++m) all = word_terminal(B->h_words[(cur & 1) * n + m]);
        if (all || done >= cap) break;
    }
    if (e != hipSuccess) {
        B->begun = false;
        return B->fail(UVH_E_HIP, std::string("uvhb_lin_solve: ") + hipGetErrorString(e));
    }
    return uvhb_lin_end(B, dx, iterations, relres, converged, stream);
}

}  // extern "C"
