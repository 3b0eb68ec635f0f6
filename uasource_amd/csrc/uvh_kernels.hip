// HIP kernels of the uvh assembly for gfx950 (CDNA4, wave64).
//
// Three kernels per assembly:
//   k_nodal_b    per node: lower surface b from floatation (UaMain/Calc_bs_From_hBS.m:71-79,109);
//                also clears the assembly's NaN flags
//   k_element    per tile of elements: gather of nodal fields through the connectivity into LDS,
//                Gauss-point physics (UaMain/uvhAssemblyIntPointImplicitSUPG.m:62-395) and the
//                element residual rows and Jacobian rows (UaMain/uvhNodalLoopSSTREAM.m:39-228),
//                stored in the node-incidence order
//   k_rows       per node: deterministic sum of the element rows into R (Tint - Fext) and into
//                the CSR rows of K (UaMain/uvhMatrixAssembly.m:416-509); incident elements are
//                summed in ascending element order, so results are bitwise reproducible.
// Alternatives of the same contract: k_rows_gather (one thread per K entry), the batched
// persistent kernels (many meshes per launch) and k_fused (owner-computes patches with an LDS
// window of node rows, no element rows in HBM).
//
// Thread mapping of k_element: a 256-thread block owns `epb` elements. Threads first gather
// (element, node) pairs into LDS, then compute the physics of (element, integration point)
// pairs and publish a record of per-point coefficients in LDS, then accumulate element rows
// ((block row, element, row) items, or (element, row) items for all nine blocks at once)
// over all points in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "uvh_internal.h"

namespace uvh {

// nodal values staged per element in LDS
enum : int {
    NV_H, NV_U, NV_V, NV_S, NV_B, NV_RHO, NV_C, NV_M, NV_N, NV_A, NV_AS1, NV_AB1,
    NV_H0, NV_U0, NV_V0, NV_AS0, NV_AB0, NV_DADH, NV_BL, NV_X, NV_Y, NV_Q, NV_MUK, NV_JV0, NV_COUNT
};

// per-(element, integration point) coefficient record; W = detJ * weight folded in where noted
enum : int {
    C_I11, C_I12, C_I21, C_I22,     // inverse Jacobian: Nx = I11 dN/dxi + I12 dN/deta, Ny = I21 .. + I22 ..
    C_TU, C_TV,                     // SUPG: psi_i = N_i + TU Nx_i + TV Ny_i  (TU = tau0 u0 / 2)
    C_WHETA, C_WHE2,                // W h eta, 2 W h E
    C_S1, C_S2, C_EXY,              // 2exx+eyy, 2eyy+exx, exy
    C_WDTXDU, C_WDTYDV, C_WDTXDV,   // W dtaux/du, W dtauy/dv, W dtaux/dv (= W dtauy/du)
    C_PX, C_PXY, C_PCX, C_PY, C_PCY,  // Kxh row factor: PX Nx + PXY Ny + PCX N;  Kyh: PY Ny + PXY Nx + PCY N
    C_W,                            // W
    C_AX, C_AY, C_RH, C_GAM, C_RU, C_RV,  // Khu/Khv/Khh column factors (dt theta folded in)
    C_WTX, C_WTY, C_WF1X, C_WF1Y, C_WF2, C_TH, C_FH,  // residual factors
    C_COUNT
};
constexpr int kNC = C_COUNT + (C_COUNT & 1);  // even, for 16-byte LDS reads

// Element rows in HBM (two-pass path): 9 x NOD values each, back to back. The rows of one
// node are contiguous (incidence order), so the row scatter reads one contiguous range per node.
template <int NOD, typename T>
constexpr int eb_stride() { return 9 * NOD; }

// Lower surface b at a node from floatation (Calc_bs_From_hBS.m:71-79,109), h being the
// node's thickness (h0 under ZeroFields). Every operation is rounded on its own (no fused
// multiply-adds), so that the nodal kernels and the element kernels' gathers, which both use
// it, give the same b bit for bit.
__device__ __forceinline__ double lower_surface(double h, double S, double B, double rho, const Params &p) {
#pragma clang fp contract(off)
    if (p.reset_thickness && h < p.thick_min) h = p.thick_min;
    const double hf = p.rhow * (S - B) / rho;
    const double GF = 0.5 + 0.5 * tanh(p.kH * ((h - hf) - p.Hh0));
    const double bfloat = S - rho * h / p.rhow;
    double bb = GF * B + (1.0 - GF) * bfloat;
    if (p.enforce_b_above && bb < B) bb = B;
    return bb;
}

__global__ void k_nodal_b(int nn, FieldPtrs f, Params p, double *__restrict__ b, int *__restrict__ status) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 && status) *status = 0;   // the assembly's NaN flags start clear
    if (k >= nn) return;
    b[k] = lower_surface(p.zero_fields ? f.h0[k] : f.h[k], f.S[k], f.B[k], f.rho[k], p);
}

// uv assembly: the lower surface at the nodes from the given upper surface, b = s - h (the
// points interpolate s and h, uvMatrixAssemblySSTREAM.m:219); clears the NaN flags like
// k_nodal_b.
__global__ void k_nodal_s(int nn, const double *__restrict__ s, const double *__restrict__ h,
                          double *__restrict__ b, int *__restrict__ status) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 && status) *status = 0;
    if (k >= nn) return;
    b[k] = s[k] - h[k];
}

template <int NOD>
__device__ __forceinline__ void corner_ids(int &c0, int &c1, int &c2) {
    if (NOD == 3) { c0 = 0; c1 = 1; c2 = 2; }
    else if (NOD == 6) { c0 = 0; c1 = 2; c2 = 4; }
    else { c0 = 0; c1 = 3; c2 = 6; }
}

// LDS add without returning the old value (ds_add_f64). Callers guarantee that the lanes of
// one call never target the same address and that successive calls to one address come
// from the same wave in program order or are separated by a barrier, so the order of the
// additions, and the rounded result, is fixed.
__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each
// XCD has its own L2. Neighbouring element ranges share nodes, so give every XCD one
// contiguous eighth of the ranges: logical block = (b % 8) * ceil(G / 8) + b / 8.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int grid) {
    const int per = (grid + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

// Transcendentals in the element arithmetic type: double for the plain form, float for the
// precision form.
template <typename T>
struct Fn;
template <>
struct Fn<double> {
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double expm1(double x) { return ::expm1(x); }
    static __device__ __forceinline__ double tanh(double x) { return ::tanh(x); }
    static __device__ __forceinline__ double abs(double x) { return ::fabs(x); }
};
template <>
struct Fn<float> {
    static __device__ __forceinline__ float exp(float x) { return ::expf(x); }
    static __device__ __forceinline__ float log(float x) { return ::logf(x); }
    static __device__ __forceinline__ float sqrt(float x) { return ::sqrtf(x); }
    static __device__ __forceinline__ float expm1(float x) { return ::expm1f(x); }
    static __device__ __forceinline__ float tanh(float x) { return ::tanhf(x); }
    static __device__ __forceinline__ float abs(float x) { return ::fabsf(x); }
};

// Reciprocal for divisors that are finite and away from zero (1 + t in [1, 2], a density, a
// Jacobian determinant, Glen's and Weertman's exponents, squared speeds and strain rates with
// their regularisation): the hardware estimate refined by two Newton steps in FP64 (one
// rounding error more than a correctly rounded division when multiplied), the hardware
// estimate alone in FP32 (about one ulp). A zero or infinite divisor gives NaN, not Inf.
template <typename T>
__device__ __forceinline__ T rcp_nr(T x);
template <>
__device__ __forceinline__ double rcp_nr<double>(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
template <>
__device__ __forceinline__ float rcp_nr<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}

// Smooth step and its derivative at x = k (z - z0) (HeavisideApprox.m:101, DiracDelta.m:29):
//   H = 1/2 + 1/2 tanh(x),   D = k/2 sech^2(x).
// One expm1 gives both: with t = exp(-2|x|), tanh|x| = (1 - t)/(1 + t) and
// sech^2 x = 4t/(1 + t)^2, each to a few ulp for every x.
template <typename T>
__device__ __forceinline__ void soft_step(T x, T k, T &H, T &D) {
    const T em = Fn<T>::expm1(T(-2) * Fn<T>::abs(x));  // t - 1, in (-1, 0]
    const T rden = rcp_nr(T(2) + em);                   // 1 / (1 + t)
    const T th = copysign(-em * rden, x);
    H = T(0.5) + T(0.5) * th;
    D = T(2) * k * ((T(1) + em) * rden * rden);
}

// Floatation terms at one point (uvhAssemblyIntPointImplicitSUPG.m:209-245):
// He = H(h - hf), HE = H(hf - h), delta and Delta their derivatives, H+ = H(H) H with
// H = S - B, the water depth d = HE rho h / rhow + He H+ and dd = d(d)/dh.
template <typename T>
__device__ __forceinline__ void floatation(T h, T hf, T H, T rho, const Params &p, T &He, T &delta, T &d, T &dd) {
    const T kH = T(p.kH), Hh0 = T(p.Hh0), rrhow = rcp_nr(T(p.rhow));
    T HE, Delta, HH, unused;
    soft_step<T>(kH * ((h - hf) - Hh0), kH, He, delta);
    if (p.Hh0 == 0.0) {
        HE = T(1) - He;
        Delta = delta;
    } else {
        soft_step<T>(kH * ((hf - h) - Hh0), kH, HE, Delta);
    }
    soft_step<T>(kH * (H - Hh0), kH, HH, unused);
    const T Hpos = HH * H;
    d = HE * rho * h * rrhow + He * Hpos;
    dd = HE * rho * rrhow - Delta * h * rho * rrhow + delta * Hpos;
}

// Physics of one integration point in arithmetic type T: fills the coefficient record c[kNC].
// The nodal values and shape tables are FP64. For T = float (precision form) the thickness,
// surface, bed and density are still interpolated in FP64, gradients are taken from node
// differences (exact in FP64), and the floatation terms and the hydrostatic combinations
// built on them are evaluated in FP64 at every point: on floating ice rho h - rhow d cancels
// to zero, which FP32 cannot resolve.
// x^y as MATLAB's x.^y for x >= 0: 0^0 = 1, 0^y = 0 (y > 0) or Inf (y < 0)
template <typename T>
__device__ __forceinline__ T pow_nonneg(T x, T y) {
    return y == T(0) ? T(1) : Fn<T>::exp(y * Fn<T>::log(x));
}

// Sliding laws other than Weertman (BasalDrag.m:182-257): Budd (W-N0, helper :448-498),
// Coulomb (:365-408), Tsai (minCW-N0: the Coulomb values where |tau_C|^2 < |tau_W|^2),
// Cornford (rpCW-N0, rpCWN0.m), Umbi (rCW-N0, rCWN0.m) and Joughin (rCW-v0, rCWV0.m). They
// replace the Weertman values of tau and its derivatives; beta2 and Db are Weertman's, He and
// delta are clamped at zero. The effective pressure (N0, BasalDrag.m:500-521) is
// N = rho g Dh + Nzero with Dh = max(h - hf, 0) and hf = max(rhow H / rho, 0) (both cut at
// eps); dN/dh = rho g where Dh > 0, 0 elsewhere. Everything here is FP64, also in the precision
// form: (mu N)^m of Cornford underflows FP32 on floating ice.
//
// With U = |(u, v, u0)|, C' = C + C0 and tau = f (u, v), the u, v derivatives follow from
// f'(U): d tau_u/du = f + u^2 f'/U, d tau_u/dv = u v f'/U. The h-derivatives are the
// reference's: Cornford's and Umbi's delta terms are taken at He = 1/2 (the value of the step
// where its derivative lives), and Umbi's dN/dh term carries He once.
template <int NOD, typename T>
__device__ __forceinline__ void other_sliding_law(const double *__restrict__ nv, const double *__restrict__ tN,
                                                  const Params &p, double hd, double Hd, double rhod, T uT, T vT,
                                                  T mT, T U2T, T beta2T, T DbT, T HecT, T dlcT, T &taux, T &tauy,
                                                  T &dtxdu, T &dtydv, T &dtxdv, T &dtxdh, T &dtydh) {
    constexpr double kEps = 2.220446049250313e-16;
    double q = 0, muk = 0, V0 = 0, C = 0;
    for (int i = 0; i < NOD; ++i) {
        const double Ni = tN[i];
        q += Ni * nv[NV_Q * NOD + i];
        muk += Ni * nv[NV_MUK * NOD + i];
        V0 += Ni * nv[NV_JV0 * NOD + i];
        C += Ni * nv[NV_C * NOD + i];
    }
    if (p.uv && C < p.cmin) C = p.cmin;
    const double u = double(uT), v = double(vT), m = double(mT), U2 = double(U2T);
    const double beta2 = double(beta2T), Db = double(DbT), He = double(HecT), dl = double(dlcT);
    double tx = double(taux), ty = double(tauy), txu = double(dtxdu), tyv = double(dtydv), txv = double(dtxdv);
    double txh = double(dtxdh), tyh = double(dtydh);
    double hf = p.rhow * Hd / rhod;
    if (hf < kEps) hf = 0.0;
    double Dh = hd - hf;
    const bool dry = Dh < kEps;   // no grounded ice above floatation
    if (dry) Dh = 0.0;
    const double N = rhod * p.g * Dh + p.nzero;
    const double dNdh = dry ? 0.0 : rhod * p.g;
    const double U = sqrt(U2), rm = 1.0 / m, Cp = C + p.czero;
    // tau = f (u, v) with f'(U) / U = g f: the u, v derivatives
    auto set_uv = [&](double f, double g) {
        tx = f * u; ty = f * v;
        txu = f * (1.0 + u * u * g); tyv = f * (1.0 + v * v * g); txv = f * u * v * g;
    };
    switch (p.law) {
        case 1: {   // Budd: tau = He N^(q/m) beta2 (u, v)
            const double qm = q / m, Nqm = pow_nonneg(N, qm), Tb = He * Nqm * beta2;
            tx = Tb * u; ty = Tb * v;
            txu = He * Nqm * (beta2 + Db * u * u);
            tyv = He * Nqm * (beta2 + Db * v * v);
            txv = He * Nqm * Db * u * v;
            // the reference's form (:489): delta times the whole of T, so He enters twice
            const double E = dl * Tb + He * qm * pow_nonneg(N, qm - 1.0) * dNdh * beta2;
            txh = E * u; tyh = E * v;
        } break;
        case 2:
        case 3: {   // Coulomb: tau = muk N (u, v) / U; Tsai keeps Weertman where that is smaller
            const double Tau = muk * N;
            const double cx = Tau * u / U, cy = Tau * v / U;
            if (p.law == 3 && !(cx * cx + cy * cy < tx * tx + ty * ty)) break;
            const double cube = U * U * U;
            tx = cx; ty = cy;
            txu = Tau * (1.0 / U - u * u / cube);
            tyv = Tau * (1.0 / U - v * v / cube);
            txv = -Tau * u * v / cube;
            const double E = dry ? 0.0 : muk * dNdh / U;
            txh = E * u; tyh = E * v;
        } break;
        case 4: {   // Cornford: |tau| = tauC tauW / (tauW^m + tauC^m)^(1/m), tauC = muk N, tauW = He w
            const double w = exp((log(U) - log(Cp)) * rm);        // C'^(-1/m) U^(1/m)
            const double P = pow_nonneg(muk * N, m), wm = U / Cp;  // (muk N)^m, w^m
            const double Q = pow_nonneg(He, m) * wm;
            const double S = P + Q;
            const double f = muk * N * He * w * pow_nonneg(S, -rm) / U;
            set_uv(f, (P / (m * S) - 1.0) / U2);
            const double dHe = pow_nonneg(2.0, m + 1.0) * pow_nonneg(muk * N, m + 1.0) * w *
                               pow_nonneg(pow_nonneg(2.0, m) * P + wm, -(m + 1.0) * rm);
            const double dN = muk * pow_nonneg(He, m + 1.0) * pow_nonneg(w, m + 1.0) * pow_nonneg(S, -(m + 1.0) * rm);
            const double E = (dl * dHe + dNdh * dN) / U;
            txh = E * u; tyh = E * v;
        } break;
        case 5: {   // Umbi: |tau| = tauC tauW / (tauW + tauC)
            const double Um = exp(log(U) * rm);                    // U^(1/m)
            const double a = He * Um, b = muk * N * exp(log(Cp) * rm);
            const double f = muk * N * He * Um / ((a + b) * U);
            set_uv(f, (b / (m * (a + b)) - 1.0) / U2);
            const double h2 = 2.0 * b + Um;
            const double E = (muk * dNdh * He * Um * Um / ((a + b) * (a + b)) +
                              4.0 * muk * N * b * dl * Um / (h2 * h2)) / U;
            txh = E * u; tyh = E * v;
        } break;
        case 6: {   // Joughin: |tau| = He C'^(-1/m) (U / (U + V0))^(1/m)
            const double base = exp(-log(Cp) * rm) * exp(-log(U + V0) * rm) * exp((rm - 1.0) * log(U));
            const double f = He * base;
            set_uv(f, -(m * U + (m - 1.0) * V0) / (m * U2 * (U + V0)));
            const double E = dl * base;
            txh = E * u; tyh = E * v;
        } break;
        default: break;
    }
    taux = T(tx); tauy = T(ty); dtxdu = T(txu); dtydv = T(tyv); dtxdv = T(txv); dtxdh = T(txh); dtydh = T(tyh);
}

// SL: the sliding laws other than Weertman are compiled in (a separate instantiation, so
// that their registers do not weigh on the Weertman kernels)
template <int NOD, typename T, bool SL = false>
__device__ void ip_physics(const double *__restrict__ nv,  // [NV_COUNT][NOD] in LDS
                           const double *__restrict__ tN, const double *__restrict__ tXi,
                           const double *__restrict__ tEta, double wq, const Params &p,
                           T *__restrict__ c) {
    // Jacobian and inverse (UaMain/derivVector.m:96-107). Coordinates are taken relative to
    // the element's first node: sum_i dN_i = 0, so J is unchanged, and the differences are
    // exact, which avoids cancelling ~1e5 m coordinates down to ~1e2 m element sizes.
    T J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    const double xo = nv[NV_X * NOD], yo = nv[NV_Y * NOD];
#pragma unroll
    for (int i = 1; i < NOD; ++i) {
        const T xi = T(nv[NV_X * NOD + i] - xo), yi = T(nv[NV_Y * NOD + i] - yo);
        const T dxi = T(tXi[i]), deta = T(tEta[i]);
        J11 += dxi * xi; J12 += dxi * yi;
        J21 += deta * xi; J22 += deta * yi;
    }
    const T det = J11 * J22 - J12 * J21;
    const T rdet = rcp_nr(det);
    const T i11 = J22 * rdet, i12 = -J12 * rdet, i21 = -J21 * rdet, i22 = J11 * rdet;

    // interpolation and gradients (uvhAssemblyIntPointImplicitSUPG.m:62-126,206-208,272-300);
    // gradients of h, h0, b and rho from differences to node 0 (sum_i dN_i = 0)
    double hd = 0, h0d = 0, Bd = 0, Sd = 0, rhod = 0;
    T u = 0, v = 0, C = 0, m = 0, A = 0, n = 0, u0 = 0, v0 = 0;
    T as0 = 0, ab0 = 0, as1 = 0, ab1 = 0, dadh = 0;
    T dhdx = 0, dhdy = 0, dh0dx = 0, dh0dy = 0, exx0 = 0, eyy0 = 0, dbdx = 0, dbdy = 0;
    T drhodx = 0, drhody = 0, exx = 0, eyy = 0, exy2 = 0;
    const double hr = nv[NV_H * NOD], h0r = nv[NV_H0 * NOD], br = nv[NV_BL * NOD], rhor = nv[NV_RHO * NOD];
#pragma unroll 1
    for (int i = 0; i < NOD; ++i) {
        const double Nd = tN[i];
        const T Ni = T(Nd);
        const T nx = i11 * T(tXi[i]) + i12 * T(tEta[i]);
        const T ny = i21 * T(tXi[i]) + i22 * T(tEta[i]);
        const double hid = nv[NV_H * NOD + i], h0id = nv[NV_H0 * NOD + i], rhoid = nv[NV_RHO * NOD + i];
        const T dh = T(hid - hr), dh0 = T(h0id - h0r), dbl = T(nv[NV_BL * NOD + i] - br), drho = T(rhoid - rhor);
        const T ui = T(nv[NV_U * NOD + i]), vi = T(nv[NV_V * NOD + i]);
        const T u0i = T(nv[NV_U0 * NOD + i]), v0i = T(nv[NV_V0 * NOD + i]);
        hd += Nd * hid; h0d += Nd * h0id; u += Ni * ui; v += Ni * vi;
        C += Ni * T(nv[NV_C * NOD + i]); m += Ni * T(nv[NV_M * NOD + i]);
        A += Ni * T(nv[NV_A * NOD + i]); n += Ni * T(nv[NV_N * NOD + i]);
        u0 += Ni * u0i; v0 += Ni * v0i;
        as0 += Ni * T(nv[NV_AS0 * NOD + i]); ab0 += Ni * T(nv[NV_AB0 * NOD + i]);
        as1 += Ni * T(nv[NV_AS1 * NOD + i]); ab1 += Ni * T(nv[NV_AB1 * NOD + i]);
        dadh += Ni * T(nv[NV_DADH * NOD + i]);
        Bd += Nd * nv[NV_B * NOD + i]; Sd += Nd * nv[NV_S * NOD + i];
        rhod += Nd * rhoid;
        dhdx += nx * dh; dhdy += ny * dh;
        dh0dx += nx * dh0; dh0dy += ny * dh0;
        exx0 += nx * u0i; eyy0 += ny * v0i;
        dbdx += nx * dbl; dbdy += ny * dbl;
        drhodx += nx * drho; drhody += ny * drho;
        exx += nx * ui; eyy += ny * vi;
        exy2 += nx * vi + ny * ui;
    }
    const T h = T(hd), h0 = T(h0d), rho = T(rhod);
    const T exy = T(0.5) * exy2;
    if (p.uv) {   // uvMatrixAssemblySSTREAM.m:185,211
        if (C < T(p.cmin)) C = T(p.cmin);
        if (A < T(p.aglenmin)) A = T(p.aglenmin);
    }
    const T a1 = as1 + ab1, a0 = as0 + ab0;

    // floatation (uvhAssemblyIntPointImplicitSUPG.m:209-245), FP64, and the combinations of
    // the water depth d that enter the driving stress and its derivative
    const double Hd = Sd - Bd;
    const double hfd = p.rhow * Hd * rcp_nr(rhod);
    double He64, delta64, d64, dd64;
    floatation<double>(hd, hfd, Hd, rhod, p, He64, delta64, d64, dd64);
    const T He = T(He64), delta = T(delta64);
    const T hyd1 = T(rhod * hd - p.rhow * d64);             // rho h - rhow d
    const T hyd2 = T(rhod * hd * hd - p.rhow * d64 * d64);  // rho h^2 - rhow d^2
    const T hyd3 = T(rhod - p.rhow * dd64);                 // rho - rhow dd
    const T hyd4 = T(rhod * hd - p.rhow * d64 * dd64);      // rho h - rhow d dd
    const T dh0h = T(h0d - hd);                             // h0 - h

    // Glen's flow law (EffectiveViscositySSTREAM.m:7-9) from logarithms:
    //   A^(-1/n) e^((1-n)/n) = exp(-log(A)/n + (1-n)/n * log(e)),  E = (1-n)/(4n) * that / e^2
    const T eps0 = T(p.eps0);
    const T e2 = eps0 * eps0 + exx * exx + eyy * eyy + exx * eyy + exy * exy;
    const T rn = rcp_nr(n);
    const T visc = Fn<T>::exp(-Fn<T>::log(A) * rn + (T(1) - n) * rn * (T(0.5) * Fn<T>::log(e2)));
    const T eta = T(0.5) * visc + T(p.eta0);
    const T E = T(0.25) * (T(1) - n) * rn * visc * rcp_nr(e2);

    // Weertman sliding (BasalDrag.m:54-77,410-443):
    //   beta2 = (C+C0)^(-1/m) U^(1/m-1),  Dbeta2 = (1/m-1) beta2 / U^2
    const T Hec = He < T(0) ? T(0) : He, dlc = delta < T(0) ? T(0) : delta;
    const T uz = T(p.uz);
    const T U2 = u * u + v * v + uz * uz;
    const T rm = rcp_nr(m);
    const T beta2 = Fn<T>::exp(-Fn<T>::log(C + T(p.czero)) * rm + (rm - T(1)) * (T(0.5) * Fn<T>::log(U2)));
    const T Db = (rm - T(1)) * beta2 * rcp_nr(U2);
    T taux = Hec * beta2 * u, tauy = Hec * beta2 * v;
    T dtxdu = Hec * (beta2 + Db * u * u), dtydv = Hec * (beta2 + Db * v * v);
    T dtxdv = Hec * Db * u * v;
    T dtxdh = dlc * beta2 * u, dtydh = dlc * beta2 * v;
    if constexpr (SL) {
        if (p.law != 0)   // the same law at every point: a uniform branch
            other_sliding_law<NOD, T>(nv, tN, p, hd, Hd, rhod, u, v, m, U2, beta2, Db, Hec, dlc, taux, tauy, dtxdu,
                                      dtydv, dtxdv, dtxdh, dtydh);
    }

    // SUPG (uvhAssemblyIntPointImplicitSUPG.m:323-390, SUPGtau.m:24-77)
    int c0i, c1i, c2i;
    corner_ids<NOD>(c0i, c1i, c2i);
    const double X1 = nv[NV_X * NOD + c0i], X2 = nv[NV_X * NOD + c1i], X3 = nv[NV_X * NOD + c2i];
    const double Y1 = nv[NV_Y * NOD + c0i], Y2 = nv[NV_Y * NOD + c1i], Y3 = nv[NV_Y * NOD + c2i];
    const T area = T(0.5 * fabs((X1 - X3) * (Y2 - Y1) - (X1 - X2) * (Y3 - Y1)));
    const T l = Fn<T>::sqrt(T(2) * area);
    const T speed0 = Fn<T>::sqrt(u0 * u0 + v0 * v0 + uz * uz);
    const T taut = T(p.dt / 2.0 + 2.220446049250313e-16);
    const T taus = l * rcp_nr(T(2) * speed0 + uz);
    T tau;
    switch (p.tau_opt) {
        case 1: tau = taut; break;
        case 2: {
            const T ecn = taut / taus;
            const T K = ecn < T(1e-6) ? ecn / T(3) : T(1) / Fn<T>::tanh(ecn) - T(1) / ecn;
            tau = K * taus;
        } break;
        case 3: tau = T(p.dt / 2.0) * T(1) / (T(1) + taut / taus); break;
        case 4: tau = T(0); break;
        default: tau = taus;
    }
    const T tau0 = T(p.beta0) * (T(p.tau_mult) * tau);
    const T W = det * T(wq);

    // fluxes (uvhNodalLoopSSTREAM.m:39-44)
    const T qx1 = rho * exx * h + rho * u * dhdx + drhodx * u * h;
    const T qy1 = rho * eyy * h + rho * v * dhdy + drhody * v * h;
    const T qx0 = rho * exx0 * h0 + rho * u0 * dh0dx + drhodx * u0 * h0;
    const T qy0 = rho * eyy0 * h0 + rho * v0 * dh0dy + drhody * v0 * h0;
    const T th = T(p.theta), dt = T(p.dt), g = T(p.g), ca = T(p.ca), sa = T(p.sa);
    const T dtth = dt * th;

    c[C_I11] = i11; c[C_I12] = i12; c[C_I21] = i21; c[C_I22] = i22;
    c[C_TU] = T(0.5) * tau0 * u0; c[C_TV] = T(0.5) * tau0 * v0;
    c[C_WHETA] = W * h * eta; c[C_WHE2] = T(2) * W * h * E * T(p.dvisk);
    c[C_S1] = T(2) * exx + eyy; c[C_S2] = T(2) * eyy + exx; c[C_EXY] = exy;
    c[C_WDTXDU] = W * dtxdu; c[C_WDTYDV] = W * dtydv; c[C_WDTXDV] = W * dtxdv;
    const T c2 = ca * g * hyd4;
    const T cx1 = dtxdh + ca * g * hyd3 * dbdx - sa * g * rho;
    const T cy1 = dtydh + ca * g * hyd3 * dbdy;
    c[C_PX] = W * (eta * (T(4) * exx + T(2) * eyy) - c2);
    c[C_PXY] = W * (eta * T(2) * exy);
    c[C_PCX] = W * cx1;
    c[C_PY] = W * (eta * (T(4) * eyy + T(2) * exx) - c2);
    c[C_PCY] = W * cy1;
    c[C_W] = W;
    c[C_AX] = dtth * (rho * dhdx + drhodx * h);
    c[C_AY] = dtth * (rho * dhdy + drhody * h);
    c[C_RH] = dtth * (rho * h);
    c[C_GAM] = rho - dtth * rho * dadh + dtth * (rho * exx + drhodx * u + rho * eyy + drhody * v);
    c[C_RU] = dtth * (rho * u);
    c[C_RV] = dtth * (rho * v);
    c[C_WTX] = W * taux; c[C_WTY] = W * tauy;
    c[C_WF1X] = W * (-ca * g * hyd1 * dbdx + rho * g * h * sa);
    c[C_WF1Y] = W * (-ca * g * hyd1 * dbdy);
    c[C_WF2] = W * (T(0.5) * ca * g * hyd2);
    c[C_TH] = -W * (rho * dh0h);
    c[C_FH] = W * (dt * rho * ((T(1) - th) * a0 + th * a1) - dt * (th * qx1 + (T(1) - th) * qx0 + th * qy1 + (T(1) - th) * qy0));
}

// Gather of node a of element e into the LDS staging record nv[NV_COUNT][NOD].
template <int NOD, bool SL = true>
__device__ __forceinline__ void gather_node_at(int k, int a, const double *__restrict__ xn,
                                               const double *__restrict__ yn, const FieldPtrs &f,
                                               const double *__restrict__ bnod, const Params &p,
                                               double *__restrict__ nv);

template <int NOD, bool SL>
__device__ __forceinline__ void gather_node(const int32_t *__restrict__ conn, const double *__restrict__ xn,
                                            const double *__restrict__ yn, const FieldPtrs &f,
                                            const double *__restrict__ bnod, const Params &p, int64_t e, int a,
                                            double *__restrict__ nv) {
    gather_node_at<NOD, SL>(conn[e * NOD + a], a, xn, yn, f, bnod, p, nv);
}

// Gather of node k, local node a of its element, into the LDS staging record nv[NV_COUNT][NOD].
// Every load is issued before any loaded value is used or any branch is taken, so the gather
// costs one memory round trip: an optional field that is absent (a null pointer) is read
// through the nodal b instead and replaced by zero afterwards. q, muk and V0 are read only
// where the sliding laws other than Weertman are compiled in (SL).
template <int NOD, bool SL>
__device__ __forceinline__ void gather_node_at(int k, int a, const double *__restrict__ xn,
                                               const double *__restrict__ yn, const FieldPtrs &f,
                                               const double *__restrict__ bnod, const Params &p,
                                               double *__restrict__ nv) {
    const double *pb = bnod ? bnod : f.h;   // a valid pointer: b itself or a stand-in
    const double *pd = f.dadh ? f.dadh : pb;
    double h = f.h[k], u = f.u[k], v = f.v[k], u0 = f.u0[k], v0 = f.v0[k];
    double as1 = f.as1[k], ab1 = f.ab1[k], as0 = f.as0[k], ab0 = f.ab0[k];
    const double h0 = f.h0[k], S = f.S[k], B = f.B[k], rho = f.rho[k], C = f.C[k], m = f.m[k], n = f.n[k];
    const double A = f.A[k], x = xn[k], y = yn[k];
    double bl = pb[k];
    double dadh = pd[k], q = 0.0, muk = 0.0, V0 = 0.0;
    if constexpr (SL) {
        const double *pq = f.q ? f.q : pb, *pm = f.muk ? f.muk : pb, *pv = f.V0 ? f.V0 : pb;
        q = pq[k]; muk = pm[k]; V0 = pv[k];
        if (!f.q) q = 0.0;
        if (!f.muk) muk = 0.0;
        if (!f.V0) V0 = 0.0;
    }
    if (!f.dadh) dadh = 0.0;
    if (p.zero_fields) {  // uvhMatrixAssembly.m:73-87; no mass-balance feedback term (:143, :168)
        u = 0.0; v = 0.0; u0 = 0.0; v0 = 0.0; h = h0;
        as1 = fabs(ab1) + 1.0; ab1 = fabs(ab1);
        as0 = fabs(ab0) + 1.0; ab0 = fabs(ab0);
        dadh = 0.0;
    }
    // without a nodal b (the uvh assembly), b is computed here from the gathered values, as
    // k_nodal_b does per node
    if (!bnod) bl = lower_surface(h, S, B, rho, p);
    nv[NV_H * NOD + a] = h; nv[NV_U * NOD + a] = u; nv[NV_V * NOD + a] = v;
    nv[NV_S * NOD + a] = S; nv[NV_B * NOD + a] = B; nv[NV_RHO * NOD + a] = rho;
    nv[NV_C * NOD + a] = C; nv[NV_M * NOD + a] = m; nv[NV_N * NOD + a] = n;
    nv[NV_A * NOD + a] = A; nv[NV_AS1 * NOD + a] = as1; nv[NV_AB1 * NOD + a] = ab1;
    nv[NV_H0 * NOD + a] = h0; nv[NV_U0 * NOD + a] = u0; nv[NV_V0 * NOD + a] = v0;
    nv[NV_AS0 * NOD + a] = as0; nv[NV_AB0 * NOD + a] = ab0;
    nv[NV_DADH * NOD + a] = dadh;
    nv[NV_BL * NOD + a] = bl;
    nv[NV_X * NOD + a] = x; nv[NV_Y * NOD + a] = y;
    if constexpr (SL) {
        nv[NV_Q * NOD + a] = q;
        nv[NV_MUK * NOD + a] = muk;
        nv[NV_JV0 * NOD + a] = V0;
    }
}

// Row i, block row BI (0: x, 1: y, 2: h) of the element: T and F of that row and its
// 3 x NOD Jacobian blocks, accumulated over the integration points in type T.
// The terms of uvhNodalLoopSSTREAM.m:62-179 (2025 form) are regrouped per (i, point) so that
// each (i, j, point) costs three FMAs per uv block: with a_j = S1 Nx_j + EXY Ny_j and
// b_j = S2 Ny_j + EXY Nx_j,
//   Kxu_ij = (4 wh eta Nx_i + wE a_i S1) Nx_j + (wh eta Ny_i + wE a_i EXY) Ny_j + W dtaux/du N_i N_j
//   Kxv_ij = (2 wh eta Nx_i + wE a_i S2) Ny_j + (wh eta Ny_i + wE a_i EXY) Nx_j + W dtaux/dv N_i N_j
//   Kyu_ij = (2 wh eta Ny_i + wE b_i S1) Nx_j + (wh eta Nx_i + wE b_i EXY) Ny_j + W dtaux/dv N_i N_j
//   Kyv_ij = (4 wh eta Ny_i + wE b_i S2) Ny_j + (wh eta Nx_i + wE b_i EXY) Nx_j + W dtauy/dv N_i N_j
// (wh eta = W h eta, wE = 2 W h E), summed over the points.
// The row's own shape values (index i, different in every lane) come from the LDS tables
// sN/sXi/sEta ([nip][NOD]); the column values (index j, the same in every lane) are read
// from the global tables g, which the compiler turns into scalar loads. In element_row_b the
// inverse Jacobian is folded into the per-(row, point) coefficients, so the column loop
// multiplies the reference-element derivatives directly (7-9 FMAs per column instead of 11
// instructions).
template <typename T>
struct RowTab;
template <>
struct RowTab<double> { using type = ShapeTables; };
template <>
struct RowTab<float> { using type = ShapeTables32; };

template <int NOD, int BI, typename T>
__device__ __forceinline__ void element_row_b(int i, int nip, const T *__restrict__ cr, const T *__restrict__ sN,
                                              const T *__restrict__ sXi, const T *__restrict__ sEta,
                                              const typename RowTab<T>::type *__restrict__ g, T &Tr, T &Fr,
                                              T (&K)[3][NOD]) {
    T Ts = 0, Fs = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int j = 0; j < NOD; ++j) K[b][j] = T(0);
    for (int q = 0; q < nip; ++q) {
        const T *r = cr + q * kNC;
        const T *qN = g->N[q], *qXi = g->dxi[q], *qEta = g->deta[q];
        const T I11 = r[C_I11], I12 = r[C_I12], I21 = r[C_I21], I22 = r[C_I22];
        const T Ni = sN[q * NOD + i];
        const T Nxi = I11 * sXi[q * NOD + i] + I12 * sEta[q * NOD + i];
        const T Nyi = I21 * sXi[q * NOD + i] + I22 * sEta[q * NOD + i];
        if (BI < 2) {
            const T S1 = r[C_S1], S2 = r[C_S2], EXY = r[C_EXY];
            const T wheta = r[C_WHETA], whE2 = r[C_WHE2];
            T cxx, cyy, cmix, cn0, cn1, ch;
            if (BI == 0) {
                const T ai = S1 * Nxi + EXY * Nyi;
                Ts += T(2) * wheta * ai + r[C_WTX] * Ni;
                Fs += r[C_WF1X] * Ni + r[C_WF2] * Nxi;
                const T wa = whE2 * ai;
                cxx = T(4) * wheta * Nxi + wa * S1;    // Kxu: Nx_j
                cmix = wheta * Nyi + wa * EXY;         // Kxu: Ny_j and Kxv: Nx_j
                cyy = T(2) * wheta * Nxi + wa * S2;    // Kxv: Ny_j
                cn0 = r[C_WDTXDU] * Ni;
                cn1 = r[C_WDTXDV] * Ni;
                ch = r[C_PX] * Nxi + r[C_PXY] * Nyi + r[C_PCX] * Ni;
            } else {
                const T bi = S2 * Nyi + EXY * Nxi;
                Ts += T(2) * wheta * bi + r[C_WTY] * Ni;
                Fs += r[C_WF1Y] * Ni + r[C_WF2] * Nyi;
                const T wb = whE2 * bi;
                cxx = T(2) * wheta * Nyi + wb * S1;    // Kyu: Nx_j
                cmix = wheta * Nxi + wb * EXY;         // Kyu: Ny_j and Kyv: Nx_j
                cyy = T(4) * wheta * Nyi + wb * S2;    // Kyv: Ny_j
                cn0 = r[C_WDTXDV] * Ni;
                cn1 = r[C_WDTYDV] * Ni;
                ch = r[C_PY] * Nyi + r[C_PXY] * Nxi + r[C_PCY] * Ni;
            }
            // the inverse Jacobian folded into the row coefficients: with Nx_j = I11 xi_j + I12 eta_j
            // and Ny_j = I21 xi_j + I22 eta_j, c Nx_j + c' Ny_j = (c I11 + c' I21) xi_j + (c I12 + c' I22) eta_j
            const T k0x = cxx * I11 + cmix * I21, k0e = cxx * I12 + cmix * I22;
            const T k1x = cyy * I21 + cmix * I11, k1e = cyy * I22 + cmix * I12;
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const T Nj = qN[j], Xj = qXi[j], Ej = qEta[j];
                K[0][j] = fma(k0x, Xj, fma(k0e, Ej, fma(cn0, Nj, K[0][j])));
                K[1][j] = fma(k1x, Xj, fma(k1e, Ej, fma(cn1, Nj, K[1][j])));
                K[2][j] = fma(ch, Nj, K[2][j]);
            }
        } else {
            const T psi = Ni + r[C_TU] * Nxi + r[C_TV] * Nyi;
            Ts += r[C_TH] * psi;
            Fs += r[C_FH] * psi;
            const T Wpsi = r[C_W] * psi;
            const T hAX = Wpsi * r[C_AX], hAY = Wpsi * r[C_AY], hRH = Wpsi * r[C_RH];
            const T hG = Wpsi * r[C_GAM], hRU = Wpsi * r[C_RU], hRV = Wpsi * r[C_RV];
            const T k0x = hRH * I11, k0e = hRH * I12, k1x = hRH * I21, k1e = hRH * I22;
            const T k2x = hRU * I11 + hRV * I21, k2e = hRU * I12 + hRV * I22;
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const T Nj = qN[j], Xj = qXi[j], Ej = qEta[j];
                K[0][j] = fma(k0x, Xj, fma(k0e, Ej, fma(hAX, Nj, K[0][j])));
                K[1][j] = fma(k1x, Xj, fma(k1e, Ej, fma(hAY, Nj, K[1][j])));
                K[2][j] = fma(k2x, Xj, fma(k2e, Ej, fma(hG, Nj, K[2][j])));
            }
        }
    }
    Tr = Ts; Fr = Fs;
}

// Block rows x and y of row i in one pass: T and F of both and their 6 x NOD Jacobian blocks,
// with the inverse Jacobian folded into the per-(row, point) coefficients as in element_row_b
// (the same regrouped terms); the prefactors that depend only on the row's point values
// (Nx_i, Ny_i, a_i, b_i) are formed once for both rows. Without K (RONLY) only T and F.
template <int NOD, typename T, bool RONLY>
__device__ __forceinline__ void element_row_uv(int i, int nip, const T *__restrict__ cr, const T *__restrict__ sN,
                                               const T *__restrict__ sXi, const T *__restrict__ sEta,
                                               const typename RowTab<T>::type *__restrict__ g, T (&TF)[4],
                                               T (&K)[RONLY ? 1 : 6][NOD]) {
    T Tx = 0, Fx = 0, Ty = 0, Fy = 0;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 6; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) K[b][j] = T(0);
    }
    for (int q = 0; q < nip; ++q) {
        const T *r = cr + q * kNC;
        const T *qN = g->N[q], *qXi = g->dxi[q], *qEta = g->deta[q];
        const T I11 = r[C_I11], I12 = r[C_I12], I21 = r[C_I21], I22 = r[C_I22];
        const T Ni = sN[q * NOD + i];
        const T Nxi = I11 * sXi[q * NOD + i] + I12 * sEta[q * NOD + i];
        const T Nyi = I21 * sXi[q * NOD + i] + I22 * sEta[q * NOD + i];
        const T S1 = r[C_S1], S2 = r[C_S2], EXY = r[C_EXY];
        const T wheta = r[C_WHETA];
        const T ai = S1 * Nxi + EXY * Nyi;
        const T bi = S2 * Nyi + EXY * Nxi;
        Tx += T(2) * wheta * ai + r[C_WTX] * Ni;
        Fx += r[C_WF1X] * Ni + r[C_WF2] * Nxi;
        Ty += T(2) * wheta * bi + r[C_WTY] * Ni;
        Fy += r[C_WF1Y] * Ni + r[C_WF2] * Nyi;
        if (!RONLY) {
            const T whE2 = r[C_WHE2];
            const T wa = whE2 * ai, wb = whE2 * bi;
            const T xxx = T(4) * wheta * Nxi + wa * S1, xmix = wheta * Nyi + wa * EXY, xyy = T(2) * wheta * Nxi + wa * S2;
            const T yxx = T(2) * wheta * Nyi + wb * S1, ymix = wheta * Nxi + wb * EXY, yyy = T(4) * wheta * Nyi + wb * S2;
            const T dxu = r[C_WDTXDU] * Ni, dxv = r[C_WDTXDV] * Ni, dyv = r[C_WDTYDV] * Ni;
            const T xh = r[C_PX] * Nxi + r[C_PXY] * Nyi + r[C_PCX] * Ni;
            const T yh = r[C_PY] * Nyi + r[C_PXY] * Nxi + r[C_PCY] * Ni;
            const T a0x = xxx * I11 + xmix * I21, a0e = xxx * I12 + xmix * I22;   // Kxu
            const T a1x = xyy * I21 + xmix * I11, a1e = xyy * I22 + xmix * I12;   // Kxv
            const T b0x = yxx * I11 + ymix * I21, b0e = yxx * I12 + ymix * I22;   // Kyu
            const T b1x = yyy * I21 + ymix * I11, b1e = yyy * I22 + ymix * I12;   // Kyv
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const T Nj = qN[j], Xj = qXi[j], Ej = qEta[j];
                K[0][j] = fma(a0x, Xj, fma(a0e, Ej, fma(dxu, Nj, K[0][j])));
                K[1][j] = fma(a1x, Xj, fma(a1e, Ej, fma(dxv, Nj, K[1][j])));
                K[2][j] = fma(xh, Nj, K[2][j]);
                K[3][j] = fma(b0x, Xj, fma(b0e, Ej, fma(dxv, Nj, K[3][j])));
                K[4][j] = fma(b1x, Xj, fma(b1e, Ej, fma(dyv, Nj, K[4][j])));
                K[5][j] = fma(yh, Nj, K[5][j]);
            }
        }
    }
    TF[0] = Tx; TF[1] = Fx; TF[2] = Ty; TF[3] = Fy;
}

// Row i of the element in one pass: T and F of the three block rows and all nine Jacobian
// blocks, with the column derivatives Nx_j, Ny_j formed once per (j, point) for all blocks
// (the same regrouped terms as element_row_b). Used where the 9 x NOD accumulators fit the
// register budget (3- and 6-node elements).
template <int NOD, typename T, bool RONLY>
__device__ __forceinline__ void element_row_full(int i, int nip, const T *__restrict__ cr, const T *__restrict__ sN,
                                                 const T *__restrict__ sXi, const T *__restrict__ sEta,
                                                 const typename RowTab<T>::type *__restrict__ g, T (&TF)[6],
                                                 T (&K)[RONLY ? 1 : 9][NOD]) {
    T Tx = 0, Fx = 0, Ty = 0, Fy = 0, Th = 0, Fh = 0;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 9; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) K[b][j] = T(0);
    }
    for (int q = 0; q < nip; ++q) {
        const T *r = cr + q * kNC;
        const T *qN = g->N[q], *qXi = g->dxi[q], *qEta = g->deta[q];
        const T I11 = r[C_I11], I12 = r[C_I12], I21 = r[C_I21], I22 = r[C_I22];
        const T Ni = sN[q * NOD + i];
        const T Nxi = I11 * sXi[q * NOD + i] + I12 * sEta[q * NOD + i];
        const T Nyi = I21 * sXi[q * NOD + i] + I22 * sEta[q * NOD + i];
        const T S1 = r[C_S1], S2 = r[C_S2], EXY = r[C_EXY];
        const T wheta = r[C_WHETA], whE2 = r[C_WHE2];
        const T ai = S1 * Nxi + EXY * Nyi;
        const T bi = S2 * Nyi + EXY * Nxi;
        const T psi = Ni + r[C_TU] * Nxi + r[C_TV] * Nyi;
        Tx += T(2) * wheta * ai + r[C_WTX] * Ni;
        Fx += r[C_WF1X] * Ni + r[C_WF2] * Nxi;
        Ty += T(2) * wheta * bi + r[C_WTY] * Ni;
        Fy += r[C_WF1Y] * Ni + r[C_WF2] * Nyi;
        Th += r[C_TH] * psi;
        Fh += r[C_FH] * psi;
        if (!RONLY) {
            const T wa = whE2 * ai, wb = whE2 * bi;
            const T xxx = T(4) * wheta * Nxi + wa * S1, xmix = wheta * Nyi + wa * EXY, xyy = T(2) * wheta * Nxi + wa * S2;
            const T yxx = T(2) * wheta * Nyi + wb * S1, ymix = wheta * Nxi + wb * EXY, yyy = T(4) * wheta * Nyi + wb * S2;
            const T dxu = r[C_WDTXDU] * Ni, dxv = r[C_WDTXDV] * Ni, dyv = r[C_WDTYDV] * Ni;
            const T xh = r[C_PX] * Nxi + r[C_PXY] * Nyi + r[C_PCX] * Ni;
            const T yh = r[C_PY] * Nyi + r[C_PXY] * Nxi + r[C_PCY] * Ni;
            const T Wpsi = r[C_W] * psi;
            const T hAX = Wpsi * r[C_AX], hAY = Wpsi * r[C_AY], hRH = Wpsi * r[C_RH];
            const T hG = Wpsi * r[C_GAM], hRU = Wpsi * r[C_RU], hRV = Wpsi * r[C_RV];
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const T Nj = qN[j];
                const T Nxj = I11 * qXi[j] + I12 * qEta[j];
                const T Nyj = I21 * qXi[j] + I22 * qEta[j];
                K[0][j] = fma(xxx, Nxj, fma(xmix, Nyj, fma(dxu, Nj, K[0][j])));   // Kxu
                K[1][j] = fma(xyy, Nyj, fma(xmix, Nxj, fma(dxv, Nj, K[1][j])));   // Kxv
                K[2][j] = fma(xh, Nj, K[2][j]);                                   // Kxh
                K[3][j] = fma(yxx, Nxj, fma(ymix, Nyj, fma(dxv, Nj, K[3][j])));   // Kyu
                K[4][j] = fma(yyy, Nyj, fma(ymix, Nxj, fma(dyv, Nj, K[4][j])));   // Kyv
                K[5][j] = fma(yh, Nj, K[5][j]);                                   // Kyh
                K[6][j] = fma(hAX, Nj, fma(hRH, Nxj, K[6][j]));                   // Khu
                K[7][j] = fma(hAY, Nj, fma(hRH, Nyj, K[7][j]));                   // Khv
                K[8][j] = fma(hG, Nj, fma(hRU, Nxj, fma(hRV, Nyj, K[8][j])));     // Khh
            }
        }
    }
    TF[0] = Tx; TF[1] = Fx; TF[2] = Ty; TF[3] = Fy; TF[4] = Th; TF[5] = Fh;
}

// Per-element strides of the LDS staging (doubles) and of the coefficient records (values of
// type T): one value of padding each, so that the records of the ~11 elements a wave reads at
// once fall on different LDS banks (an unpadded stride of 7 x 36 values repeats every 4th
// element on gfx950's 32 four-byte banks).
template <int NOD>
__host__ __device__ constexpr int nv_stride() { return NV_COUNT * NOD + 1; }
constexpr int nv_stride_host(int nod) { return NV_COUNT * nod + 1; }
__host__ __device__ constexpr int rec_stride(int nip) { return nip * kNC + 1; }

// Shape tables in LDS for the row stage, in type T: the FP64 tables themselves for the plain
// form, FP32 copies for the precision form.
template <typename T>
__device__ __forceinline__ int row_table_doubles(int tstride) {
    return sizeof(T) == sizeof(double) ? 0 : (3 * tstride + 1) / 2;
}

// Exchange of a double between lanes through DPP, control CTRL. Every lane of the lane's row
// of 16 must be active.
template <int CTRL>
__device__ __forceinline__ double lane_dpp(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The value of lane ^ MASK (MASK 1, 2 or 4): quad_perm swaps for 1 and 2; for 4 the half-row
// mirror (lane i of 8 takes lane 7 - i) followed by the quad reversal (i takes i ^ 3).
template <int MASK>
__device__ __forceinline__ double xor_dpp(double x) {
    if constexpr (MASK == 1) return lane_dpp<0xB1>(x);
    else if constexpr (MASK == 2) return lane_dpp<0x4E>(x);
    else return lane_dpp<0x1B>(lane_dpp<0x141>(x));
}

// One butterfly stage of the transpose below: for each slot pair (a, a + MASK), the lane with
// bit MASK set keeps the upper slot and takes its partner's lower slot into its lower one, the
// other lane keeps the lower slot and takes its partner's upper slot into its upper one.
template <int G, int MASK>
__device__ __forceinline__ void row_butterfly(double (&s)[G][2], int q) {
    const bool h = (q & MASK) != 0;
#pragma unroll
    for (int a = 0; a < G; ++a) {
        if (a & MASK) continue;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const double r = xor_dpp<MASK>(h ? s[a][w] : s[a + MASK][w]);
            s[a][w] = h ? r : s[a][w];
            s[a + MASK][w] = h ? s[a + MASK][w] : r;
        }
    }
}

// Stores each lane's NV-value row piece v at EB[row * stride + off] (FP64 element rows), the G
// lanes (4 or 8) of an aligned lane group together. The pieces are cut into 16-byte units; each
// run of G units is transposed across the group (log2 G butterfly stages of DPP exchanges), and
// then one store instruction per source lane t writes units G m .. G m + G - 1 of lane t's row
// from the group's G lanes, 16 G contiguous bytes, instead of G lanes writing 16 bytes each to
// G rows. Units past the last whole run are stored by their own lane. A lane with row < 0 has
// nothing stored. Every lane of the group must be active, and row * stride + off must be even.
template <int G, int NV>
__device__ __forceinline__ void group_store_rows(double *__restrict__ EB, int stride, int row, int off,
                                                const double *v) {
    static_assert(NV % 2 == 0 && (G == 4 || G == 8), "16-byte units, groups of 4 or 8 lanes");
    constexpr int NU = NV / 2, NC = NU / G;
    const int lane = __lane_id(), q = lane & (G - 1);
    int rows[G];
    if constexpr (G == 4) {
        rows[0] = __builtin_amdgcn_mov_dpp(row, 0x00, 0xF, 0xF, false);
        rows[1] = __builtin_amdgcn_mov_dpp(row, 0x55, 0xF, 0xF, false);
        rows[2] = __builtin_amdgcn_mov_dpp(row, 0xAA, 0xF, 0xF, false);
        rows[3] = __builtin_amdgcn_mov_dpp(row, 0xFF, 0xF, 0xF, false);
    } else {
#pragma unroll
        for (int t = 0; t < G; ++t) rows[t] = __shfl(row, (lane & ~(G - 1)) + t, 64);
    }
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        double s[G][2];
#pragma unroll
        for (int c = 0; c < G; ++c) {
            s[c][0] = v[2 * G * m + 2 * c];
            s[c][1] = v[2 * G * m + 2 * c + 1];
        }
        row_butterfly<G, 1>(s, q);
        row_butterfly<G, 2>(s, q);
        if constexpr (G == 8) row_butterfly<G, 4>(s, q);
        // slot t now holds unit G m + q of lane t's piece
#pragma unroll
        for (int t = 0; t < G; ++t)
            if (rows[t] >= 0) {
                *reinterpret_cast<double2 *>(EB + (int64_t)rows[t] * stride + off + 2 * G * m + 2 * q) =
                    make_double2(s[t][0], s[t][1]);
            }
    }
    if (row >= 0) {
#pragma unroll
        for (int k = 2 * G * NC; k < NV; k += 2) {
            *reinterpret_cast<double2 *>(EB + (int64_t)row * stride + off + k) = make_double2(v[k], v[k + 1]);
        }
    }
}

// lanes per group of the element kernel's row stores (group_store_rows); groups of 8 (128-byte
// runs, a third exchange stage) measured slower at 7 points per element
constexpr int kRowGroup = 4;

// Per-element kernel: EPB elements per 256-thread block (runtime), dense thread mappings
// for the three stages: (element, node) gathers, (element, point) physics, (element, row)
// rows. The caller guarantees epb*NOD <= 256 and epb*nip <= 256. T is the arithmetic type of
// the physics and the element rows (double, or float for the precision form); the rows are
// stored in T at their position in the node-incidence order.
template <int NOD, typename T, bool RONLY, bool FULL, bool SL = false>
__device__ __forceinline__ void element_block(double *lds, int nele, int nip, int epb, int lb, int nthr,
                                              const int32_t *__restrict__ conn, const double *__restrict__ xn,
                                              const double *__restrict__ yn, const FieldPtrs &f,
                                              const double *__restrict__ bnod, const ShapeTables *__restrict__ tab,
                                              const typename RowTab<T>::type *__restrict__ rtab, const Params &p,
                                              const int32_t *__restrict__ inc_pos, T *__restrict__ EB,
                                              T *__restrict__ RE) {
    const int tstride = nip * NOD;
    double *sN = lds;                       // [nip][NOD]
    double *sXi = sN + tstride;
    double *sEta = sXi + tstride;
    double *sW = sEta + tstride;            // [nip]
    T *rN = reinterpret_cast<T *>(sN), *rXi = reinterpret_cast<T *>(sXi), *rEta = reinterpret_cast<T *>(sEta);
    const int ttab = (3 * tstride + nip + 1) & ~1;
    if (sizeof(T) != sizeof(double)) {
        rN = reinterpret_cast<T *>(lds + ttab);
        rXi = rN + tstride;
        rEta = rXi + tstride;
    }
    double *work = lds + ttab + ((row_table_doubles<T>(tstride) + 1) & ~1);  // 16-byte aligned union:
                                            // nodal [epb][NV_COUNT][NOD] / coef [epb][nip][kNC]
    const int tid = threadIdx.x;
    // The shape-table entry of this thread is loaded before the gathers and stored after them,
    // so that the tile's global loads take two round trips in all (tables and connectivity,
    // then the nodal values) instead of one per loop.
    double vN = 0, vXi = 0, vEta = 0, vW = 0;
    if (tid < tstride) {
        const int q = tid / NOD, i = tid - q * NOD;
        vN = tab->N[q][i]; vXi = tab->dxi[q][i]; vEta = tab->deta[q][i];
    }
    if (tid < nip) vW = tab->w[tid];
    const int64_t e0 = (int64_t)lb * epb;

    // Stages 1 and 2 leave the highest waves of the block idle (epb*NOD and epb*nip < nthr);
    // which waves those are rotates with the tile, so that the several blocks a CU holds do
    // not all leave the same SIMD idle. vt is the thread's index within the rotated order.
    const int nwav = nthr >> 6;
    const int vt = nwav > 1 ? (tid + 64 * (lb % nwav)) % nthr : tid;

    // stage 1: gathers, one (element, node) per thread
    if (vt < epb * NOD) {
        const int el = vt / NOD, a = vt - el * NOD;
        if (e0 + el < nele) gather_node<NOD, SL>(conn, xn, yn, f, bnod, p, e0 + el, a, work + el * nv_stride<NOD>());
    }
    for (int k = tid; k < tstride; k += nthr) {
        if (k != tid) {
            const int q = k / NOD, i = k - q * NOD;
            vN = tab->N[q][i]; vXi = tab->dxi[q][i]; vEta = tab->deta[q][i];
        }
        sN[k] = vN; sXi[k] = vXi; sEta[k] = vEta;
        if (sizeof(T) != sizeof(double)) {
            rN[k] = T(vN); rXi[k] = T(vXi); rEta[k] = T(vEta);
        }
    }
    for (int k = tid; k < nip; k += nthr) sW[k] = k == tid ? vW : tab->w[k];
    __syncthreads();

    // stage 2: physics, one (element, integration point) per thread
    T c[kNC];
    const int pel = vt / nip, pq = vt - pel * nip;
    const bool phys = vt < epb * nip && e0 + pel < nele;
    if (phys) ip_physics<NOD, T, SL>(work + pel * nv_stride<NOD>(), sN + pq * NOD, sXi + pq * NOD, sEta + pq * NOD,
                                 sW[pq], p, c);
    __syncthreads();  // the nodal staging is dead from here; the coefficient records reuse it
    T *rec = reinterpret_cast<T *>(work);
    if (phys) {
        T *dst = rec + pel * rec_stride(nip) + pq * kNC;
#pragma unroll
        for (int k = C_COUNT; k < kNC; ++k) c[k] = T(0);
#pragma unroll
        for (int k = 0; k < kNC; ++k) dst[k] = c[k];
    }
    __syncthreads();

    const int nrow = epb * NOD;
    if (FULL) {
        // stage 3, one pass: (element, row) items, all nine blocks per item
        for (int it = tid; it < nrow; it += nthr) {
            const int el = it / NOD, i = it - el * NOD;
            const int64_t e = e0 + el;
            if (e >= nele) continue;
            const int64_t row = inc_pos[e * NOD + i];
            T TF[6], Kf[RONLY ? 1 : 9][NOD];
            element_row_full<NOD, T, RONLY>(i, nip, rec + el * rec_stride(nip), rN, rXi, rEta, rtab, TF, Kf);
#pragma unroll
            for (int k = 0; k < 6; ++k) RE[row * 6 + k] = TF[k];
            if (!RONLY) {
                T *eb = EB + row * eb_stride<NOD, T>();
#pragma unroll
                for (int b = 0; b < (RONLY ? 1 : 9); ++b)
#pragma unroll
                    for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = Kf[b][j];
            }
        }
        return;
    }
    // stage 3: element rows. FP64 3- and 6-node elements: (element, row) items for the x and y
    // block rows together, then (element, row) items for the h block row; else (block row,
    // element, row) items. Each accumulates its Jacobian terms and T, F over the points in
    // registers. Each kind of item takes a segment of the threads' item space that starts at a
    // wave boundary (seg: nrow rounded up to whole waves), so that no wave runs two kinds of row
    // arithmetic in one round, unless that takes the block one more round of items.
    const int nbr = p.uv ? 2 : 3;   // the uv assembly has no thickness rows
    constexpr bool kUV2 = NOD <= 6 && sizeof(T) == sizeof(double);
    const int nkind = kUV2 ? (p.uv ? 1 : 2) : nbr;
    const int wseg = ((nrow + 63) / 64) * 64;
    const int seg = (nkind * wseg + nthr - 1) / nthr <= (nkind * nrow + nthr - 1) / nthr ? wseg : nrow;
    if constexpr (kUV2 && NOD == 6 && !RONLY) {
        // 6-node rows stored a lane group at a time (group_store_rows): the lanes of a group
        // share a kind of item (seg is a multiple of the group) and a trip of the loop (each
        // wave runs the same trips); a lane without an item computes the tile's first row and
        // stores nothing
        if (seg % kRowGroup == 0 && (nthr & 63) == 0) {
            for (int vb = tid & ~63; vb < nkind * seg; vb += nthr) {
                const int v = vb + (tid & 63);
                const int kind = v / seg, rr = v - kind * seg;
                int el = rr / NOD, i = rr - el * NOD;
                const bool item = v < nkind * seg && rr < nrow && e0 + el < nele;
                if (!item) el = i = 0;
                const int row = item ? inc_pos[(e0 + el) * NOD + i] : -1;
                const T *crr = rec + el * rec_stride(nip);
                if (kind == 0) {
                    T TF[4], Kb[6][NOD];
                    element_row_uv<NOD, T, false>(i, nip, crr, rN, rXi, rEta, rtab, TF, Kb);
                    if (item) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) RE[(int64_t)row * 6 + k] = TF[k];
                    }
                    group_store_rows<kRowGroup, 6 * NOD>(EB, eb_stride<NOD, T>(), row, 0, &Kb[0][0]);
                } else {
                    T Tr, Fr, Kb[3][NOD];
                    element_row_b<NOD, 2, T>(i, nip, crr, rN, rXi, rEta, rtab, Tr, Fr, Kb);
                    if (item) {
                        RE[(int64_t)row * 6 + 4] = Tr;
                        RE[(int64_t)row * 6 + 5] = Fr;
                    }
                    group_store_rows<kRowGroup, 3 * NOD>(EB, eb_stride<NOD, T>(), row, 6 * NOD, &Kb[0][0]);
                }
            }
            return;
        }
    }
    if constexpr (kUV2) {
        for (int v = tid; v < nkind * seg; v += nthr) {
            const int kind = v / seg, rr = v - kind * seg;
            if (rr >= nrow) continue;
            const int el = rr / NOD, i = rr - el * NOD;
            const int64_t e = e0 + el;
            if (e >= nele) continue;
            const int64_t row = inc_pos[e * NOD + i];
            const T *crr = rec + el * rec_stride(nip);
            if (kind == 0) {
                T TF[4], Kb[RONLY ? 1 : 6][NOD];
                element_row_uv<NOD, T, RONLY>(i, nip, crr, rN, rXi, rEta, rtab, TF, Kb);
#pragma unroll
                for (int k = 0; k < 4; ++k) RE[row * 6 + k] = TF[k];
                if (!RONLY) {
                    T *eb = EB + row * eb_stride<NOD, T>();
#pragma unroll
                    for (int b = 0; b < (RONLY ? 1 : 6); ++b)
#pragma unroll
                        for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = Kb[b][j];
                }
            } else {
                T Tr, Fr, Kb[3][NOD];
                element_row_b<NOD, 2, T>(i, nip, crr, rN, rXi, rEta, rtab, Tr, Fr, Kb);
                RE[row * 6 + 4] = Tr;
                RE[row * 6 + 5] = Fr;
                if (!RONLY) {
                    T *eb = EB + row * eb_stride<NOD, T>() + 6 * NOD;
#pragma unroll
                    for (int b = 0; b < 3; ++b)
#pragma unroll
                        for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = Kb[b][j];
                }
            }
        }
        return;
    }
    for (int v = tid; v < nbr * seg; v += nthr) {
        const int bi = v / seg;
        const int rr = v - bi * seg;
        if (rr >= nrow) continue;
        const int el = rr / NOD, i = rr - el * NOD;
        const int64_t e = e0 + el;
        if (e >= nele) continue;
        // the row is stored at its position in the node-incidence order
        const int64_t row = inc_pos[e * NOD + i];
        const T *crr = rec + el * rec_stride(nip);
        T Tr, Fr, Kb[3][NOD];
        if (bi == 0) element_row_b<NOD, 0, T>(i, nip, crr, rN, rXi, rEta, rtab, Tr, Fr, Kb);
        else if (bi == 1) element_row_b<NOD, 1, T>(i, nip, crr, rN, rXi, rEta, rtab, Tr, Fr, Kb);
        else element_row_b<NOD, 2, T>(i, nip, crr, rN, rXi, rEta, rtab, Tr, Fr, Kb);
        RE[row * 6 + 2 * bi] = Tr;
        RE[row * 6 + 2 * bi + 1] = Fr;
        if (!RONLY) {
            T *eb = EB + row * eb_stride<NOD, T>() + bi * 3 * NOD;
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = Kb[b][j];
        }
    }
}

template <int NOD, typename T, bool RONLY, bool FULL, bool SL = false>
__global__ __launch_bounds__(256) void k_element(int nele, int nip, int epb, const int32_t *__restrict__ conn,
                                                 const double *__restrict__ xn, const double *__restrict__ yn,
                                                 FieldPtrs f, const double *__restrict__ bnod,
                                                 const ShapeTables *__restrict__ tab,
                                                 const typename RowTab<T>::type *__restrict__ rtab, Params p,
                                                 const int32_t *__restrict__ inc_pos,
                                                 T *__restrict__ EB, T *__restrict__ RE, int lb0, int lb1,
                                                 int *__restrict__ status) {
    extern __shared__ double lds[];
    const int lb = lb0 + xcd_block(blockIdx.x, gridDim.x);   // tiles lb0 .. lb1-1 of this launch
    if (lb >= lb1) return;  // whole block
    // without a nodal pre-pass (bnod null), the block of the first tile clears the assembly's
    // NaN flags; the row kernels that set them run after this kernel
    if (status && lb == 0 && threadIdx.x == 0) *status = 0;
    element_block<NOD, T, RONLY, FULL, SL>(lds, nele, nip, epb, lb, blockDim.x, conn, xn, yn, f, bnod, tab, rtab, p,
                                           inc_pos, EB, RE);
}

// Node rows: row_lanes(NOD) lanes per node, 64 / row_lanes nodes per wave. The incident element
// rows (stored contiguously per node, in incidence order, by k_element) are added, in
// ascending element order, into an LDS buffer that holds the node's three CSR rows
// (9*deg values); within one element row no two values share a column, and the adds of one
// lane group are issued in program order, so every sum has a fixed order and the result is
// bitwise reproducible. The buffer is then written to K, contiguous along the rows. Element
// rows of type TE (float in the precision form) are accumulated in FP64.

// NBK = 3: the uvh system (u, v, h rows and columns); NBK = 2: the uv system, which reads
// the (u, v) part of the same element rows and writes R (2 Nn) and the 2 x 2-block K.
template <int NOD, typename TE, int NBK = 3, bool KNT = false>
__device__ __forceinline__ void rows_block(double *lds, int nn, int k0, const int32_t *__restrict__ inc_ptr,
                                           const uint8_t *__restrict__ pos, const int64_t *__restrict__ rowptr_n,
                                           const uint8_t *__restrict__ diag, int64_t nnzN,
                                           const TE *__restrict__ EB, const TE *__restrict__ RE,
                                           double *__restrict__ R, double *__restrict__ Tint,
                                           double *__restrict__ Fext, double *__restrict__ Kval, int check_nan,
                                           int *__restrict__ status) {
    constexpr int kRowLanes = row_lanes(NOD), kRowNodesPerWave = 64 / kRowLanes;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / kRowLanes, gl = lane - sub * kRowLanes;
    const int slot = wave * kRowNodesPerWave + sub;
    const int k = k0 + slot;   // the block's nodes are consecutive
    if (k >= nn) return;  // whole lane group; no barriers or cross-group shuffles below
    const int64_t rp = rowptr_n[k];
    // the block's row buffers are packed: node k's starts at NBK^2 (rowptr_n[k] - rowptr_n[k0]);
    // the launch provides the largest such total over the blocks
    double *buf = lds + NBK * NBK * (rp - rowptr_n[k0]);
    const int t0 = inc_ptr[k], t1 = inc_ptr[k + 1];
    const int deg = (int)(rowptr_n[k + 1] - rp);
    const int dg = deg > 0 ? diag[k] : 0;   // loaded now, used after the sums
    int bad = 0;
    // lane 2*bi (< 2 NBK): T of block row bi, lane 2*bi+1: F, summed in ascending element order;
    // their loads go out with the element rows of the same chunk, so that the node's residual
    // rows cost no round trips of their own
    const bool tf_lane = gl < 2 * NBK;
    double tf = 0.0;
    for (int v = gl; v < NBK * NBK * deg; v += kRowLanes) buf[v] = 0.0;
    // the node's element rows t0..t1-1 are contiguous (incidence order); CH at a time:
    // all loads first, then the adds in ascending element order. Value u of the NBK^2 x NOD
    // used ones is block (bi, bj) = (u / (NBK NOD), u / NOD mod NBK), column j = u mod NOD,
    // at (3 bi + bj) NOD + j of the element row
    constexpr int NV = NBK * NBK * NOD;                   // values used per element row
    constexpr int VPL = (NV + kRowLanes - 1) / kRowLanes; // values per lane
    // element rows loaded per round: 4 for 3- and 6-node elements, 2 for 10-node elements
    // (measured on 400k-element P1, 200k P2 and 500k P3 meshes: P1 and P2 rows within 2 %
    // for 2-4; P3 rows 1.38 ms with 2 against 1.60 ms with 4 and 2.77 ms with 8)
    constexpr int CH = NOD >= 10 ? 2 : 4;
    constexpr int PST = pos_inc_stride(NOD);
    constexpr bool kPairs = NBK == 3 ? (9 * NOD) % 2 == 0 && (sizeof(TE) == sizeof(double) || NOD <= 6)
                                     : NOD == 6 && sizeof(TE) == sizeof(double);
    if constexpr (kPairs) {
        // element rows of 6- and 10-node elements hold an even number of values (54 or 90), so
        // each row is aligned to a pair (16 bytes in FP64, 8 in FP32), and so is each block row
        // of the (u, v) blocks the uv system reads (NBK = 2, 6-node elements): each lane loads
        // pairs of values (2 gl + 2 kRowLanes u of the NV used ones), and the row's columns with
        // one 8-byte load per 8 columns, 4 element rows at a time. Every LDS address still takes
        // its adds in ascending element order, so the sums are those of the one-value-per-load
        // form bit for bit. (FP32 rows of 10-node elements keep one value per load: 90 values in
        // pairs fill 45 of 64 lane slots, one by one 90 of 96, measured 1.46 against 1.40 ms at
        // 500k P3 elements; so do the uv system's FP32 rows: 134 against 120 us at 200k P2.)
        using V2 = typename std::conditional<sizeof(TE) == sizeof(double), double2, float2>::type;
        constexpr int VP2 = (NV + 2 * kRowLanes - 1) / (2 * kRowLanes);   // pairs per lane
        constexpr int PW = PST / 8;                                      // column words per row
        for (int c0 = t0; c0 < t1; c0 += CH) {
            const int cnt = min(CH, t1 - c0);
            V2 vals[CH][VP2];
            uint64_t cw[CH][PW];
            double re[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (c < cnt) {
                    const V2 *row = reinterpret_cast<const V2 *>(EB + (int64_t)(c0 + c) * eb_stride<NOD, TE>());
#pragma unroll
                    for (int u = 0; u < VP2; ++u) {
                        const int v = 2 * (gl + kRowLanes * u);
                        if (v < NV) {
                            // value v of the used ones: block (bk / NBK, bk mod NBK), column j,
                            // at (3 bi + bj) NOD + j of the row (v itself when NBK = 3)
                            const int bk = v / NOD, j = v - bk * NOD;
                            const int src = ((bk / NBK) * 3 + bk % NBK) * NOD + j;
                            vals[c][u] = row[src >> 1];
                        }
                    }
                    const uint64_t *pw = reinterpret_cast<const uint64_t *>(pos + (int64_t)(c0 + c) * PST);
#pragma unroll
                    for (int w = 0; w < PW; ++w) cw[c][w] = pw[w];
                    if (tf_lane) re[c] = double(RE[(int64_t)(c0 + c) * 6 + gl]);
                }
            }
#pragma unroll
            for (int c = 0; c < CH; ++c)
                if (c < cnt && tf_lane) tf += re[c];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (c >= cnt) continue;
#pragma unroll
                for (int u = 0; u < VP2; ++u) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int v = 2 * (gl + kRowLanes * u) + h;
                        if (v < NV) {
                            const int bk = v / NOD, j = v - bk * NOD;
                            const int bi = bk / NBK, bj = bk - bi * NBK;
                            const int col = int((cw[c][j >> 3] >> (8 * (j & 7))) & 0xFF);
                            lds_add(buf + bi * NBK * deg + bj * deg + col, double(h ? vals[c][u].y : vals[c][u].x));
                        }
                    }
                }
            }
        }
    } else
    for (int c0 = t0; c0 < t1; c0 += CH) {
        const int cnt = min(CH, t1 - c0);
        TE vals[CH][VPL];
        int cols[CH][VPL];
        double re[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
            for (int u = 0; u < VPL; ++u) {
                const int v = gl + kRowLanes * u;
                if (c < cnt && v < NV) {
                    const int bk = v / NOD, j = v - bk * NOD;
                    const int src = NBK == 3 ? v : ((bk / NBK) * 3 + bk % NBK) * NOD + j;
                    vals[c][u] = EB[(int64_t)(c0 + c) * eb_stride<NOD, TE>() + src];
                    cols[c][u] = pos[(int64_t)(c0 + c) * PST + j];
                }
            }
            if (c < cnt && tf_lane) re[c] = double(RE[(int64_t)(c0 + c) * 6 + gl]);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c < cnt && tf_lane) tf += re[c];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
            for (int u = 0; u < VPL; ++u) {
                const int v = gl + kRowLanes * u;
                if (c < cnt && v < NV) {
                    const int bk = v / NOD;
                    const int bi = bk / NBK, bj = bk - bi * NBK;
                    lds_add(buf + bi * NBK * deg + bj * deg + cols[c][u], double(vals[c][u]));
                }
            }
        }
    }
    if (tf_lane) {
        const int bi = gl >> 1;
        if (gl & 1) {
            if (Fext) Fext[(int64_t)bi * nn + k] = tf;
        } else {
            if (Tint) Tint[(int64_t)bi * nn + k] = tf;
        }
        const double other = __shfl_xor(tf, 1, 64);
        if (!(gl & 1)) {
            const double r = tf - other;
            R[(int64_t)bi * nn + k] = r;
            if (check_nan && isnan(r)) bad |= 1;
        }
    }
    // K. The wave's nodes are consecutive, so their parts of one block row of K are one
    // contiguous range; when the wave holds all its kRowNodesPerWave nodes (uniform over the
    // wave, every lane group is here) its 64 lanes store that range together, 64 consecutive
    // values per instruction, else each lane group stores its node's part. In-order LDS
    // operations within the wave make the other groups' sums visible to these reads.
    // With KNT the stores are non-temporal: K is not read again by the assembly, and kept out
    // of the caches it leaves the element rows not yet read in the Infinity Cache (200k P2 elements:
    // rows 200 against 208 us; the batched row kernels, whose element rows far exceed that
    // cache, store K plainly: 64 meshes in FP64 2.75-2.78 against 2.80 ms).
    const int kw = k0 + wave * kRowNodesPerWave;
    if (kw + kRowNodesPerWave <= nn) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int64_t rpw = __shfl(rp, 0, 64);
        int start[kRowNodesPerWave + 1], dgs[kRowNodesPerWave];
#pragma unroll
        for (int s = 0; s < kRowNodesPerWave; ++s) {
            start[s] = NBK * (int)(__shfl(rp, s * kRowLanes, 64) - rpw);
            dgs[s] = __shfl(deg, s * kRowLanes, 64);
        }
        start[kRowNodesPerWave] = start[kRowNodesPerWave - 1] + NBK * dgs[kRowNodesPerWave - 1];
        const double *wbuf = lds + NBK * NBK * (rpw - rowptr_n[k0]);   // the wave's first node's buffer
#pragma unroll
        for (int bi = 0; bi < NBK; ++bi) {
            double *dst = Kval + (int64_t)bi * NBK * nnzN + NBK * rpw;
            for (int d = lane; d < start[kRowNodesPerWave]; d += 64) {
                int s = 0;
#pragma unroll
                for (int t = 1; t < kRowNodesPerWave; ++t) s += d >= start[t];
                int st = start[0], dgv = dgs[0];
#pragma unroll
                for (int t = 1; t < kRowNodesPerWave; ++t)
                    if (s == t) {
                        st = start[t];
                        dgv = dgs[t];
                    }
                // node s's buffer starts NBK * start[s] values after the wave's first
                const double x = wbuf[NBK * st + bi * NBK * dgv + (d - st)];
                if constexpr (KNT) __builtin_nontemporal_store(x, &dst[d]);
                else dst[d] = x;
            }
        }
    } else {
#pragma unroll
        for (int bi = 0; bi < NBK; ++bi) {
            double *dst = Kval + (int64_t)bi * NBK * nnzN + NBK * rp;
            const double *src = buf + bi * NBK * deg;
            for (int v = gl; v < NBK * deg; v += kRowLanes) {
                if constexpr (KNT) __builtin_nontemporal_store(src[v], &dst[v]);
                else dst[v] = src[v];
            }
        }
    }
    // a node no element references has an empty row (R = 0, no diagonal to check)
    if (check_nan && deg > 0 && gl < NBK && isnan(buf[gl * NBK * deg + gl * deg + dg])) bad |= 2;
    if (bad) atomicOr(status, bad);
}

template <int NOD, typename TE, int NBK = 3>
__global__ __launch_bounds__(256) void k_rows(int nn, const int32_t *__restrict__ inc_ptr,
                                              const uint8_t *__restrict__ pos,
                                              const int64_t *__restrict__ rowptr_n, const uint8_t *__restrict__ diag,
                                              int64_t nnzN, const TE *__restrict__ EB,
                                              const TE *__restrict__ RE, double *__restrict__ R,
                                              double *__restrict__ Tint, double *__restrict__ Fext,
                                              double *__restrict__ Kval, int check_nan, int *__restrict__ status,
                                              const int32_t *__restrict__ blocks, int nblk) {
    extern __shared__ double lds[];
    // node blocks of this launch: blocks[0 .. nblk-1], or 0 .. nblk-1 when blocks is null, the
    // latter taken in descending order: the element kernel writes the element rows with its
    // tiles in ascending node order, so the rows it wrote last, the likeliest to be still in
    // the memory-side Infinity Cache, are read first (200k P2 elements: 210 against 215 us)
    const int i = xcd_block(blockIdx.x, gridDim.x);
    if (i >= nblk) return;
    const int k0 = (blocks ? blocks[i] : nblk - 1 - i) * row_nodes_per_block(NOD);
    rows_block<NOD, TE, NBK, true>(lds, nn, k0, inc_ptr, pos, rowptr_n, diag, nnzN, EB, RE, R, Tint, Fext, Kval,
                             check_nan, status);
}

// Node rows, gather form: one thread per node-graph entry (node n, column c). It sums the
// nine K values of that entry over n's incident element rows in ascending element order (the
// inverse column map says which local node of each element is adj(n)[c]), so every sum has
// the same fixed order as in k_rows and the results agree bit for bit; no LDS, no atomics.
// The thread of column 0 also sums the node's residual rows.
template <int NOD, typename TE>
__global__ __launch_bounds__(256) void k_rows_gather(int nn, int64_t nnzN, const int32_t *__restrict__ nzrow,
                                                     const int64_t *__restrict__ rowptr_n,
                                                     const int32_t *__restrict__ inc_ptr,
                                                     const int64_t *__restrict__ imap_ptr,
                                                     const uint8_t *__restrict__ imap,
                                                     const uint8_t *__restrict__ diag,
                                                     const TE *__restrict__ EB, const TE *__restrict__ RE,
                                                     double *__restrict__ R, double *__restrict__ Tint,
                                                     double *__restrict__ Fext, double *__restrict__ Kval,
                                                     int check_nan, int *__restrict__ status) {
    const int64_t x = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (x >= nnzN) return;
    const int n = nzrow[x];
    const int64_t rp = rowptr_n[n];
    const int deg = (int)(rowptr_n[n + 1] - rp);
    const int c = (int)(x - rp);
    const int t0 = inc_ptr[n], t1 = inc_ptr[n + 1];
    const uint8_t *im = imap + imap_ptr[n] + c;
    double acc[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) acc[b] = 0.0;
    for (int t = t0; t < t1; ++t) {
        const int j = im[(int64_t)(t - t0) * deg];
        if (j != 0xFF) {
            const TE *row = EB + (int64_t)t * (9 * NOD) + j;
#pragma unroll
            for (int b = 0; b < 9; ++b) acc[b] += double(row[b * NOD]);
        }
    }
    int bad = 0;
#pragma unroll
    for (int bi = 0; bi < 3; ++bi)
#pragma unroll
        for (int bj = 0; bj < 3; ++bj) Kval[(int64_t)bi * 3 * nnzN + 3 * rp + bj * deg + c] = acc[bi * 3 + bj];
    if (check_nan && c == diag[n] && (isnan(acc[0]) || isnan(acc[4]) || isnan(acc[8]))) bad |= 2;
    if (c == 0) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int t = t0; t < t1; ++t)
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] += double(RE[(int64_t)t * 6 + k]);
#pragma unroll
        for (int bi = 0; bi < 3; ++bi) {
            const double r = v[2 * bi] - v[2 * bi + 1];
            R[(int64_t)bi * nn + n] = r;
            if (Tint) Tint[(int64_t)bi * nn + n] = v[2 * bi];
            if (Fext) Fext[(int64_t)bi * nn + n] = v[2 * bi + 1];
            if (check_nan && isnan(r)) bad |= 1;
        }
    }
    if (bad) atomicOr(status, bad);
}

// T and F of a node's block rows from its incident element residual rows t0 .. t1-1
// (contiguous, 6 values each), added in ascending element order; the rows are loaded four at a
// time before they are added, so that a node costs one memory round trip per four elements.
template <typename TE>
__device__ __forceinline__ void sum_residual_rows(const TE *__restrict__ RE, int t0, int t1, int nbk, double (&T)[3],
                                                  double (&F)[3]) {
    constexpr int CH = 4;
    for (int c0 = t0; c0 < t1; c0 += CH) {
        const int cnt = min(CH, t1 - c0);
        TE v[CH][6];
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c < cnt) {
#pragma unroll
                for (int x = 0; x < 6; ++x) v[c][x] = RE[(int64_t)(c0 + c) * 6 + x];
            }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c < cnt) {
#pragma unroll
                for (int bi = 0; bi < 3; ++bi)
                    if (bi < nbk) {
                        T[bi] += double(v[c][2 * bi]);
                        F[bi] += double(v[c][2 * bi + 1]);
                    }
            }
    }
}

// Residual only: one thread per node sums its incident element residual rows (contiguous).
template <typename TE>
__global__ __launch_bounds__(256) void k_rows_r(int nn, int nbk, const int32_t *__restrict__ inc_ptr,
                                                const TE *__restrict__ RE, double *__restrict__ R,
                                                double *__restrict__ Tint, double *__restrict__ Fext, int check_nan,
                                                int *__restrict__ status) {
    // nbk block rows: 3 (uvh) or 2 (uv, whose element rows carry no thickness terms)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double T[3] = {0, 0, 0}, F[3] = {0, 0, 0};
    sum_residual_rows(RE, inc_ptr[k], inc_ptr[k + 1], nbk, T, F);
    bool bad = false;
#pragma unroll
    for (int bi = 0; bi < 3; ++bi)
        if (bi < nbk) {
            const double r = T[bi] - F[bi];
            R[(int64_t)bi * nn + k] = r;
            if (Tint) Tint[(int64_t)bi * nn + k] = T[bi];
            if (Fext) Fext[(int64_t)bi * nn + k] = F[bi];
            bad = bad || isnan(r);
        }
    if (check_nan && bad) atomicOr(status, 1);
}

// ------------------------------------------------------------------ batched assembly
// Many independent meshes (ensemble members, mixed element types) in one launch per stage.
// The work is a flattened list of (member, block) items; persistent workgroups take items
// from an atomic counter until the list is exhausted, so large and small members balance
// across the CUs. Every item writes only its own element rows or node rows, so the order in
// which workgroups take items does not change any result.

// The member table as the batched kernels read it: the same layout as BatchMember, with every
// pointer qualified as global memory (address space 1). A pointer loaded from memory is
// otherwise generic, and generic (flat) accesses need 64-bit vector addresses and lose the
// scalar loads of wave-uniform data; through this view the compiler emits global and scalar
// loads, as for kernel arguments.
#define UVH_G __attribute__((address_space(1)))
struct FieldPtrsG {
    const UVH_G double *h, *u, *v, *S, *B, *rho, *C, *m, *n, *A, *as1, *ab1;
    const UVH_G double *h0, *u0, *v0, *as0, *ab0;
    const UVH_G double *dadh;
    const UVH_G double *q, *muk, *V0;
};
struct BatchMemberG {
    const UVH_G int32_t *conn;
    const UVH_G double *x, *y;
    const UVH_G ShapeTables *tab;
    const UVH_G ShapeTables32 *tab32;
    const UVH_G int32_t *inc_ptr, *inc_pos;
    const UVH_G uint8_t *pos_inc, *diag;
    const UVH_G int64_t *rowptr_n;
    UVH_G double *b;
    UVH_G void *EB, *RE;
    int64_t nnzN;
    int32_t nele, nn, nod, nip, epb, pad;
    FieldPtrsG f;
    UVH_G double *R, *Tint, *Fext, *Kval;
};
static_assert(sizeof(BatchMemberG) == sizeof(BatchMember), "member table views differ");
static_assert(sizeof(FieldPtrsG) == sizeof(FieldPtrs), "field table views differ");

__device__ __forceinline__ const UVH_G BatchMemberG *member(const BatchMember *mem, int i) {
    return (const UVH_G BatchMemberG *)mem + i;
}

__device__ __forceinline__ FieldPtrs generic_fields(const UVH_G FieldPtrsG *g) {
    return FieldPtrs{(const double *)g->h, (const double *)g->u, (const double *)g->v, (const double *)g->S,
                     (const double *)g->B, (const double *)g->rho, (const double *)g->C, (const double *)g->m,
                     (const double *)g->n, (const double *)g->A, (const double *)g->as1, (const double *)g->ab1,
                     (const double *)g->h0, (const double *)g->u0, (const double *)g->v0, (const double *)g->as0,
                     (const double *)g->ab0, (const double *)g->dadh, (const double *)g->q,
                     (const double *)g->muk, (const double *)g->V0};
}

// next item of a persistent workgroup: thread 0 takes it, the block reads it from LDS
__device__ __forceinline__ int next_item(int *counter, int *s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1);
    __syncthreads();
    // the same value in every lane: make it a scalar, so the member's table entries and
    // pointers are read with scalar loads and kept in scalar registers
    const int it = __builtin_amdgcn_readfirstlane(*s_item);
    __syncthreads();  // all threads have read it before thread 0 takes the next one
    return it;
}

// items: (member, node block of 256)
// Block 0 also zeroes the work counters and the per-member status words, which only the
// later launches of the same assembly read.
__global__ __launch_bounds__(256) void k_nodal_b_batch(const BatchMember *__restrict__ mem,
                                                       const int2 *__restrict__ items, Params p, int *counters,
                                                       int ncounters, int *status, int nstatus) {
    if (blockIdx.x == 0) {
        if (threadIdx.x < ncounters) counters[threadIdx.x] = 0;
        for (int k = threadIdx.x; k < nstatus; k += blockDim.x) status[k] = 0;
    }
    const int2 w = items[blockIdx.x];
    const BatchMember &M = mem[w.x];
    const int k = w.y * 256 + threadIdx.x;
    if (k >= M.nn) return;
    const FieldPtrs &f = M.f;
    M.b[k] = lower_surface(p.zero_fields ? f.h0[k] : f.h[k], f.S[k], f.B[k], f.rho[k], p);
}

// items: (member, element block of the member's epb elements), all members with NOD nodes
template <int NOD, typename T, bool RONLY>
__global__ __launch_bounds__(256) void k_element_batch(const BatchMember *__restrict__ mem,
                                                       const int2 *__restrict__ items, int nitems,
                                                       int *__restrict__ counter, Params p) {
    extern __shared__ double lds[];
    __shared__ int s_item;
    for (;;) {
        const int it = next_item(counter, &s_item);
        if (it >= nitems) break;
        const int2 w = items[it];
        const UVH_G BatchMemberG *M = member(mem, w.x);
        // the shape tables are not written while any kernel runs: read them as constant memory
        // (address space 4), so their wave-uniform entries come through scalar loads
        const typename RowTab<T>::type *rtab;
        if constexpr (sizeof(T) == sizeof(double))
            rtab = (const ShapeTables *)(const __attribute__((address_space(4))) ShapeTables *)M->tab;
        else
            rtab = (const ShapeTables32 *)(const __attribute__((address_space(4))) ShapeTables32 *)M->tab32;
        const FieldPtrs f = generic_fields(&M->f);
        element_block<NOD, T, RONLY, (NOD <= 6 && sizeof(T) == sizeof(float))>(
            lds, M->nele, M->nip, M->epb, w.y, 256, (const int32_t *)M->conn, (const double *)M->x,
            (const double *)M->y, f, (const double *)M->b, (const ShapeTables *)M->tab, rtab, p,
            (const int32_t *)M->inc_pos, (T *)M->EB, (T *)M->RE);
        __syncthreads();  // the LDS of this item is dead before the next one overwrites it
    }
}

// The same work list as k_element_batch, one workgroup per item instead of persistent
// workgroups: item blockIdx (XCD-aware), no work counter, the register budget of the
// single-mesh kernel. SL: the sliding laws other than Weertman compiled in.
template <int NOD, typename T, bool RONLY, bool SL = false>
__global__ __launch_bounds__(256) void k_element_batch_flat(const BatchMember *__restrict__ mem,
                                                            const int2 *__restrict__ items, int nitems, Params p) {
    extern __shared__ double lds[];
    const int it = xcd_block(blockIdx.x, gridDim.x);
    if (it >= nitems) return;
    const int2 w = items[it];
    const UVH_G BatchMemberG *M = member(mem, w.x);
    const typename RowTab<T>::type *rtab;
    if constexpr (sizeof(T) == sizeof(double))
        rtab = (const ShapeTables *)(const __attribute__((address_space(4))) ShapeTables *)M->tab;
    else
        rtab = (const ShapeTables32 *)(const __attribute__((address_space(4))) ShapeTables32 *)M->tab32;
    const FieldPtrs f = generic_fields(&M->f);
    element_block<NOD, T, RONLY, (NOD <= 6 && sizeof(T) == sizeof(float)), SL>(
        lds, M->nele, M->nip, M->epb, w.y, 256, (const int32_t *)M->conn, (const double *)M->x, (const double *)M->y, f,
        (const double *)M->b, (const ShapeTables *)M->tab, rtab, p, (const int32_t *)M->inc_pos, (T *)M->EB,
        (T *)M->RE);
}

template <int NOD, typename TE>
__global__ __launch_bounds__(256) void k_rows_batch_flat(const BatchMember *__restrict__ mem,
                                                         const int2 *__restrict__ items, int nitems, int check_nan,
                                                         int *__restrict__ status) {
    extern __shared__ double lds[];
    const int it = xcd_block(blockIdx.x, gridDim.x);
    if (it >= nitems) return;
    const int2 w = items[it];
    const UVH_G BatchMemberG *M = member(mem, w.x);
    rows_block<NOD, TE>(lds, M->nn, w.y * row_nodes_per_block(NOD), (const int32_t *)M->inc_ptr,
                        (const uint8_t *)M->pos_inc, (const int64_t *)M->rowptr_n, (const uint8_t *)M->diag, M->nnzN,
                        (const TE *)M->EB, (const TE *)M->RE, (double *)M->R, (double *)M->Tint, (double *)M->Fext,
                        (double *)M->Kval, check_nan, status + w.x);
}

// items: (member, node block of row_nodes_per_block(NOD) nodes), all members with NOD nodes;
// status has one word per member
template <int NOD, typename TE>
__global__ __launch_bounds__(256) void k_rows_batch(const BatchMember *__restrict__ mem,
                                                    const int2 *__restrict__ items, int nitems,
                                                    int *__restrict__ counter, int check_nan,
                                                    int *__restrict__ status) {
    extern __shared__ double lds[];
    __shared__ int s_item;
    for (;;) {
        const int it = next_item(counter, &s_item);
        if (it >= nitems) break;
        const int2 w = items[it];
        const UVH_G BatchMemberG *M = member(mem, w.x);
        rows_block<NOD, TE>(lds, M->nn, w.y * row_nodes_per_block(NOD), (const int32_t *)M->inc_ptr,
                            (const uint8_t *)M->pos_inc, (const int64_t *)M->rowptr_n, (const uint8_t *)M->diag,
                            M->nnzN, (const TE *)M->EB, (const TE *)M->RE, (double *)M->R, (double *)M->Tint,
                            (double *)M->Fext, (double *)M->Kval, check_nan, status + w.x);
        __syncthreads();
    }
}

// R only: items (member, node block of 256), one thread per node
template <typename TE>
__global__ __launch_bounds__(256) void k_rows_r_batch(const BatchMember *__restrict__ mem,
                                                      const int2 *__restrict__ items, int check_nan,
                                                      int *__restrict__ status) {
    const int2 w = items[blockIdx.x];
    const BatchMember &M = mem[w.x];
    const int k = w.y * 256 + threadIdx.x;
    if (k >= M.nn) return;
    const TE *RE = static_cast<const TE *>(M.RE);
    double T[3] = {0, 0, 0}, F[3] = {0, 0, 0};
    sum_residual_rows(RE, M.inc_ptr[k], M.inc_ptr[k + 1], 3, T, F);
    const double T0 = T[0], T1 = T[1], T2 = T[2], F0 = F[0], F1 = F[1], F2 = F[2];
    const int nn = M.nn;
    const double r0 = T0 - F0, r1 = T1 - F1, r2 = T2 - F2;
    M.R[k] = r0; M.R[nn + k] = r1; M.R[2 * (int64_t)nn + k] = r2;
    if (M.Tint) { M.Tint[k] = T0; M.Tint[nn + k] = T1; M.Tint[2 * (int64_t)nn + k] = T2; }
    if (M.Fext) { M.Fext[k] = F0; M.Fext[nn + k] = F1; M.Fext[2 * (int64_t)nn + k] = F2; }
    if (check_nan && (isnan(r0) || isnan(r1) || isnan(r2))) atomicOr(status + w.x, 1);
}

// Fused assembly: one 256-thread workgroup per patch, two workgroups per CU. Per batch of
// `epb` elements, in three phases separated by barriers:
//   A  zero the window slots of rows that start in the batch; add the previous batch's staged
//      ordered sums (fixups) into the window
//   B  integration-point physics into coefficient records; write out the rows whose last
//      element was in the previous batch (K, R, Tint, Fext)
//   C  row items ((element, local node) of owned nodes, per block row): the element's row in
//      registers, added into the owning node's window slot, ordered values staged where the
//      node has several items in the batch; on the remaining lanes, the gathers of the next
//      batch's nodal values (their latency hides behind the row arithmetic)
// The schedule records each carry what a thread needs in one load, and each phase's records
// are loaded at the top of the iteration, so no phase waits on a chain of index loads.
// Every K entry off a node's diagonal block, and every value of a node with at most two
// incident elements, is a sum of at most two terms and is added with ds_add_f64 in any order
// (a + b = b + a exactly); the diagonal block and T, F of a node with three or more incident
// elements take one add per batch, in batch order. Results are bitwise reproducible.
constexpr int kFusedThreads = 256;
constexpr int kFlushLanes = 16;

template <int NOD>
__device__ __forceinline__ void gather_batch(const MeshArgs *__restrict__ MA, const FieldPtrs &f, const Params &p,
                                             int b, int epb, int g, double *__restrict__ nodal) {
    const int k = MA->batch_node[int64_t(b) * epb * NOD + g];
    if (k >= 0) gather_node_at<NOD>(k, g % NOD, MA->x, MA->y, f, MA->b, p, nodal + (g / NOD) * nv_stride<NOD>());
}

template <int NOD, bool SL, bool RONLY>
__global__ __launch_bounds__(kFusedThreads) void k_fused(const MeshArgs *__restrict__ MA,
                                                         const CallArgs *__restrict__ CA, int nip, int epb,
                                                         int table_doubles, int nodal_doubles, int rec_doubles,
                                                         int stage_items, int npatch, int64_t nnzN, int nn,
                                                         int check_nan) {
    extern __shared__ double lds[];
    const int patch = xcd_block(blockIdx.x, gridDim.x);
    if (patch >= npatch) return;   // whole workgroup
    const int tstride = nip * NOD;
    double *sN = lds;
    double *sXi = sN + tstride;
    double *sEta = sXi + tstride;
    double *sW = sEta + tstride;
    double *nodal = lds + table_doubles;                 // [epb][nv_stride]
    double *rec = nodal + nodal_doubles;                 // [epb][rec_stride]
    double *stage = rec + rec_doubles;                   // [stage_items][kStageValues]
    double *win = stage + int64_t(stage_items) * kStageValues;
    const int tid = threadIdx.x;
    const ShapeTables *tab = MA->tab;
    for (int k = tid; k < tstride; k += kFusedThreads) {
        const int q = k / NOD, i = k - q * NOD;
        sN[k] = tab->N[q][i]; sXi[k] = tab->dxi[q][i]; sEta[k] = tab->deta[q][i];
    }
    if (tid < nip) sW[tid] = tab->w[tid];
    const Params &p = CA->p;      // read through the pointer (scalar loads), not held in registers
    const FieldPtrs &f = CA->f;
    constexpr int nvs = nv_stride<NOD>();
    const int rs = rec_stride(nip);
    const int gl = tid % kFlushLanes, grp = tid / kFlushLanes;
    int bad = 0;
    const int b0 = MA->patch_batch_ptr[patch], b1 = MA->patch_batch_ptr[patch + 1];
    if (b0 < b1)   // the first batch's gathers (later batches are gathered in the phase C before)
        for (int g = tid; g < epb * NOD; g += kFusedThreads) gather_batch<NOD>(MA, f, p, b0, epb, g, nodal);
    BatchHdr H{}, Hp{};
    if (b0 < b1) H = MA->hdr[b0];
    for (int b = b0; b <= b1; ++b) {
        const bool has = b < b1;
        const BatchHdr Hn = b + 1 < b1 ? MA->hdr[b + 1] : BatchHdr{};
        const int ni = H.item1 - H.item0;
        // records of this iteration's phases, loaded before any of them runs
        PatchItem it{};
        if (tid < 3 * ni) it = MA->items[H.item0 + tid % ni];
        PatchFlush fr{};
        const bool fl0 = b > b0 && Hp.flush0 + grp < Hp.flush1;
        if (fl0) fr = MA->flushes[Hp.flush0 + grp];
        // ---- phase A
        if (has) {
            const int lane = tid & 63, wave = tid >> 6;
            for (int z = H.zero0 + wave; z < H.zero1; z += kFusedThreads / 64) {
                const PatchZero zr = MA->zeros[z];
                double *w = win + zr.slot;
                for (int k = lane; k < zr.len; k += 64) w[k] = 0.0;
            }
        }
        if (b > b0) {   // the previous batch's ordered sums, in item order, one add per value
            for (int t = tid; t < (Hp.fix1 - Hp.fix0) * kStageValues; t += kFusedThreads) {
                const int gi = t / kStageValues, v = t - gi * kStageValues;
                if (RONLY && v % 5 < 3) continue;   // R only: no K values staged
                const PatchFix fx = MA->fixes[Hp.fix0 + gi];
                double sum = 0.0;
                for (int k = 0; k < fx.count; ++k) sum += stage[(fx.stage0 + k) * kStageValues + v];
                const int deg = fx.deg_dcol & 0xFFFF, dcol = fx.deg_dcol >> 16;
                const int bi = v / 5, wv = v - bi * 5;
                const int off = wv < 3 ? (3 * bi + wv) * deg + dcol : 9 * deg + (wv == 3 ? bi : 3 + bi);
                win[fx.slot + off] += sum;
            }
        }
        __syncthreads();
        // ---- phase B
        if (has) {
            for (int t = tid; t < epb * nip; t += kFusedThreads) {
                const int el = t / nip, q = t - el * nip;
                if (MA->batch_node[(int64_t(b) * epb + el) * NOD] < 0) continue;   // padding slot
                double c[kNC];
                ip_physics<NOD, double, SL>(nodal + el * nvs, sN + q * NOD, sXi + q * NOD, sEta + q * NOD, sW[q], p, c);
                double *dst = rec + el * rs + q * kNC;
#pragma unroll
                for (int k = C_COUNT; k < kNC; ++k) c[k] = 0.0;
#pragma unroll
                for (int k = 0; k < kNC; ++k) dst[k] = c[k];
            }
        }
        if (b > b0) {   // rows whose last element was in the previous batch
            double *Kval = CA->Kval;
            for (int fl = Hp.flush0 + grp; fl < Hp.flush1; fl += kFusedThreads / kFlushLanes) {
                const PatchFlush r = fl == Hp.flush0 + grp ? fr : MA->flushes[fl];
                const double *w = win + r.slot;
                const int dg = r.deg;
                if (!RONLY) {
#pragma unroll
                    for (int bk = 0; bk < 3; ++bk) {
                        double *dst = Kval + int64_t(bk) * 3 * nnzN + r.rp3;
                        const double *src = w + bk * 3 * dg;
                        for (int k = gl; k < 3 * dg; k += kFlushLanes) dst[k] = src[k];
                    }
                    // diagonal entry of block (gl, gl) (uvhMatrixAssembly.m:575-579)
                    if (check_nan && gl < 3 && isnan(w[gl * 3 * dg + gl * dg + r.dcol])) bad |= 2;
                }
                if (gl < 3) {
                    const double Tv = w[9 * dg + gl], Fv = w[9 * dg + 3 + gl];
                    const double rv = Tv - Fv;
                    CA->R[int64_t(gl) * nn + r.node] = rv;
                    if (CA->Tint) CA->Tint[int64_t(gl) * nn + r.node] = Tv;
                    if (CA->Fext) CA->Fext[int64_t(gl) * nn + r.node] = Fv;
                    if (check_nan && isnan(rv)) bad |= 1;
                }
            }
        }
        __syncthreads();
        // ---- phase C: row items of this batch, then (other lanes) the next batch's gathers
        const int ngat = b + 1 < b1 ? epb * NOD : 0;
        for (int t = tid; t < 3 * ni + ngat; t += kFusedThreads) {
            if (t >= 3 * ni) {
                gather_batch<NOD>(MA, f, p, b + 1, epb, t - 3 * ni, nodal);
                continue;
            }
            const int bi = t / ni;
            const PatchItem r = t == tid ? it : MA->items[H.item0 + (t - bi * ni)];
            const int el = r.ea / NOD, a = r.ea - el * NOD;
            const int deg = r.deg, st = r.stage;
            const double *crr = rec + el * rs;
            double Tr, Fr, Kb[3][NOD];
            if (bi == 0) element_row_b<NOD, 0, double>(a, nip, crr, sN, sXi, sEta, tab, Tr, Fr, Kb);
            else if (bi == 1) element_row_b<NOD, 1, double>(a, nip, crr, sN, sXi, sEta, tab, Tr, Fr, Kb);
            else element_row_b<NOD, 2, double>(a, nip, crr, sN, sXi, sEta, tab, Tr, Fr, Kb);
            double *w = win + r.slot + bi * 3 * deg;
            double *sg = stage + st * kStageValues + bi * 5;
            if (!RONLY) {   // (with R only the Jacobian terms above are dead code)
#pragma unroll
                for (int j = 0; j < NOD; ++j) {
                    const int cj = r.pos[j];
#pragma unroll
                    for (int bj = 0; bj < 3; ++bj) {
                        if (st >= 0 && j == a) sg[bj] = Kb[bj][j];
                        else lds_add(w + bj * deg + cj, Kb[bj][j]);
                    }
                }
            }
            if (st >= 0) {
                sg[3] = Tr;
                sg[4] = Fr;
            } else {
                lds_add(win + r.slot + 9 * deg + bi, Tr);
                lds_add(win + r.slot + 9 * deg + 3 + bi, Fr);
            }
        }
        __syncthreads();
        Hp = H;
        H = Hn;
    }
    if (bad) atomicOr(MA->status, bad);
}

// ------------------------------------------------------------------ batched residual cost
// Per member of a batch: rr = sum_i (w_i R_i)^2 and d2 = -sum_i w_i R_i d_i (the numerator of
// CalcCostFunctionNRuvh.m's rForce and its D2), with w_i 0 at fixed dofs. Chunks of kCostChunk
// entries per workgroup, a fixed-order tree in LDS per workgroup, then one thread per member
// adds its chunks in order: deterministic, no atomics, nothing to clear (graph-capturable).
constexpr int kCostChunk = 32768;

__global__ __launch_bounds__(256) void k_cost_partial(const CostMember *__restrict__ mem, const int2 *__restrict__ items,
                                                      double2 *__restrict__ partial) {
    __shared__ double2 red[256];
    const int2 it = items[blockIdx.x];
    const CostMember M = mem[it.x];
    const int64_t i0 = int64_t(it.y) * kCostChunk;
    const int64_t i1 = i0 + kCostChunk < M.n ? i0 + kCostChunk : M.n;
    double rr = 0.0, d2 = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const double r = M.w ? M.w[i] * M.R[i] : M.R[i];
        rr = fma(r, r, rr);
        d2 = fma(-r, M.d[i], d2);
    }
    red[threadIdx.x] = make_double2(rr, d2);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[threadIdx.x].x += red[threadIdx.x + h].x;
            red[threadIdx.x].y += red[threadIdx.x + h].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_cost_final(const CostMember *__restrict__ mem, int nmem, const double2 *__restrict__ partial,
                             double *__restrict__ rr, double *__restrict__ d2) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    const CostMember M = mem[m];
    double a = 0.0, b = 0.0;
    for (int c = 0; c < M.nchunk; ++c) {
        a += partial[M.first + c].x;
        b += partial[M.first + c].y;
    }
    rr[m] = a;
    d2[m] = b;
}

hipError_t launch_cost(const CostMember *mem, int nmem, const int2 *items, int nitems, double2 *partial, double *rr,
                       double *d2, hipStream_t s) {
    if (nitems > 0) hipLaunchKernelGGL(k_cost_partial, dim3(nitems), dim3(256), 0, s, mem, items, partial);
    hipLaunchKernelGGL(k_cost_final, dim3((nmem + 63) / 64), dim3(64), 0, s, mem, nmem, partial, rr, d2);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launchers

namespace {
constexpr size_t kLdsLimit = 64 * 1024;

// Raises a kernel's dynamic-LDS limit once (not at every launch, so that launches can be
// captured into a HIP graph without other runtime calls in between).
void lds_attr(const void *kernel, size_t lds) {
    static std::mutex mu;
    static std::map<const void *, size_t> done;
    if (lds <= kLdsLimit) return;
    std::lock_guard<std::mutex> g(mu);
    size_t &cur = done[kernel];
    if (cur >= lds) return;
    hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    cur = lds;
}
// Default LDS cap of the element kernel: two 256-thread workgroups per CU (160 KiB per CU).
constexpr size_t kElementLds = 80 * 1024;
constexpr size_t kElementLdsSmall = 52 * 1024;

size_t element_table_bytes(int nod, int nip, bool fp32) {
    const int tstride = nip * nod;
    size_t d = size_t((3 * tstride + nip + 1) & ~1);
    if (fp32) d += size_t(((3 * tstride + 1) / 2 + 1) & ~1);
    return d * sizeof(double);
}

size_t element_per_el_bytes(int nod, int nip, bool fp32) {
    return std::max(size_t(NV_COUNT * nod + 1) * sizeof(double),
                    size_t(rec_stride(nip)) * (fp32 ? sizeof(float) : sizeof(double)));
}

template <int NOD, typename T>
hipError_t launch_element_t(int nele, int nip, int epb, int nthr, bool full, const int32_t *conn, const double *x,
                            const double *y, const FieldPtrs &f, const double *b, const ShapeTables *tab,
                            const ShapeTables32 *tab32, const Params &p, const int32_t *inc_pos, bool ronly,
                            void *EB, void *RE, int lb0, int lb1, int *status, hipStream_t s) {
    const typename RowTab<T>::type *rtab;
    if constexpr (sizeof(T) == sizeof(double)) rtab = tab;
    else rtab = tab32;
    const int nblocks = lb1 - lb0;
    if (nblocks <= 0) return hipSuccess;
    const int grid = ((nblocks + kXcds - 1) / kXcds) * kXcds;  // whole rounds over the XCDs
    const bool fp32 = sizeof(T) == sizeof(float);
    const size_t lds = element_table_bytes(NOD, nip, fp32) + element_per_el_bytes(NOD, nip, fp32) * epb;
    constexpr bool can_full = NOD <= 6;
    auto kern = ronly ? (full && can_full ? k_element<NOD, T, true, can_full> : k_element<NOD, T, true, false>)
                      : (full && can_full ? k_element<NOD, T, false, can_full> : k_element<NOD, T, false, false>);
    if (p.law != 0)
        kern = ronly ? (full && can_full ? k_element<NOD, T, true, can_full, true> : k_element<NOD, T, true, false, true>)
                     : (full && can_full ? k_element<NOD, T, false, can_full, true>
                                         : k_element<NOD, T, false, false, true>);
    lds_attr(reinterpret_cast<const void *>(kern), lds);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(nthr), lds, s, nele, nip, epb, conn, x, y, f, b, tab, rtab, p, inc_pos,
                       static_cast<T *>(EB), static_cast<T *>(RE), lb0, lb1, status);
    return hipGetLastError();
}
}  // namespace

// Elements per block of `nthr` threads: every stage maps at most one item per thread and the
// LDS (shape tables + the larger of nodal staging and coefficient records) stays within
// `lds_cap` bytes (at most 160 KiB, the gfx950 workgroup limit).
int element_epb_max(int nod, int nip, bool fp32, size_t lds_cap, int nthr) {
    int epb = std::min(nthr / nip, nthr / nod);
    const size_t tables = element_table_bytes(nod, nip, fp32);
    if (lds_cap <= tables) return 0;
    epb = std::min<int>(epb, int((lds_cap - tables) / element_per_el_bytes(nod, nip, fp32)));
    return epb;
}

// FP64 tiles of 3- and 6-node elements under 52 KiB, so that three workgroups share a CU (the
// register budget allows three waves per SIMD): 5 % faster on 6-node meshes than two
// workgroups of 80 KiB; 10-node and FP32 tiles keep the larger cap (measured).
int element_epb_default(int nod, int nip, bool fp32, int nthr) {
    return element_epb_max(nod, nip, fp32, (!fp32 && nod <= 6) ? kElementLdsSmall : kElementLds, nthr);
}

// Tiles of the uvh assembly: FP64 3- and 6-node tiles are further capped so that the row stage
// (two items per element row: the x and y block rows together, then the h block row) takes a
// single round of the block's threads, 21 6-node elements per 256 threads instead of 27
// (200k P2 elements: element kernel 253-257 against 265-268 us, the assembly 482-489 against 498-503;
// 16-20 and 24-30 elements measured slower). The uv assembly, one item per element row, and the
// R-only calls, whose rows are cheap, keep the larger tiles (21 elements: uv 222 against 197
// us, R-only 158 against 143 us).
int element_epb_uvh(int nod, int nip, bool fp32, int nthr) {
    const int epb = element_epb_default(nod, nip, fp32, nthr);
    return (!fp32 && nod <= 6) ? std::min(epb, nthr / (2 * nod)) : epb;
}

hipError_t launch_nodal_b(int nn, const FieldPtrs &f, const Params &p, double *b, int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_nodal_b, dim3((nn + 255) / 256), dim3(256), 0, s, nn, f, p, b, status);
    return hipGetLastError();
}

hipError_t launch_nodal_s(int nn, const double *sfc, const double *h, double *b, int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_nodal_s, dim3((nn + 255) / 256), dim3(256), 0, s, nn, sfc, h, b, status);
    return hipGetLastError();
}

// Element tiles lb0 .. lb1-1 (tiles of epb elements); lb1 < 0 means all tiles.
hipError_t launch_element(int nod, int nele, int nip, int epb, int nthr, bool fp32, bool full, const int32_t *conn,
                          const double *x, const double *y, const FieldPtrs &f, const double *b,
                          const ShapeTables *tab, const ShapeTables32 *tab32, const Params &p,
                          const int32_t *inc_pos, bool ronly, void *EB, void *RE, int lb0, int lb1, int *status,
                          hipStream_t s) {
    if (nthr < 64 || nthr > 256 || nthr % 64 != 0) return hipErrorInvalidConfiguration;
    if (epb < 1 || epb * nip > nthr || epb * nod > nthr) return hipErrorInvalidConfiguration;
    const int ntiles = (nele + epb - 1) / epb;
    if (lb1 < 0 || lb1 > ntiles) lb1 = ntiles;
    if (lb0 < 0) lb0 = 0;
#define UVH_EL(NODV)                                                                                           \
    return fp32 ? launch_element_t<NODV, float>(nele, nip, epb, nthr, full, conn, x, y, f, b, tab, tab32, p, inc_pos, ronly, EB, RE, lb0, lb1, status, s) \
                : launch_element_t<NODV, double>(nele, nip, epb, nthr, full, conn, x, y, f, b, tab, tab32, p, inc_pos, ronly, EB, RE, lb0, lb1, status, s)
    switch (nod) {
        case 3: UVH_EL(3);
        case 6: UVH_EL(6);
        case 10: UVH_EL(10);
        default: return hipErrorInvalidValue;
    }
#undef UVH_EL
}

size_t element_row_stride(int nod, bool) { return size_t(9 * nod); }

// Node rows of the node blocks `blocks[0 .. nblk-1]` (all blocks when blocks is null).
hipError_t launch_rows(int nod, int nn, int64_t block_vals, bool fp32, const int32_t *inc_ptr, const uint8_t *pos,
                       const int64_t *rowptr_n, const uint8_t *diag, int64_t nnzN, const void *EB, const void *RE,
                       double *R, double *Tint, double *Fext, double *Kval, int check_nan, int *status,
                       const int32_t *blocks, int nblk, hipStream_t s, int nbk) {
    if (nbk != 2 && nbk != 3) return hipErrorInvalidValue;
    if (!Kval) {
        if (fp32)
            hipLaunchKernelGGL(k_rows_r<float>, dim3((nn + 255) / 256), dim3(256), 0, s, nn, nbk, inc_ptr,
                               static_cast<const float *>(RE), R, Tint, Fext, check_nan, status);
        else
            hipLaunchKernelGGL(k_rows_r<double>, dim3((nn + 255) / 256), dim3(256), 0, s, nn, nbk, inc_ptr,
                               static_cast<const double *>(RE), R, Tint, Fext, check_nan, status);
        return hipGetLastError();
    }
    const int nblocks = blocks ? nblk : (nn + row_nodes_per_block(nod) - 1) / row_nodes_per_block(nod);
    if (nblocks <= 0) return hipSuccess;
    const dim3 grid(((nblocks + kXcds - 1) / kXcds) * kXcds), block(256);
    const size_t lds = size_t(nbk * nbk) * size_t(block_vals) * sizeof(double);
    if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
#define UVH_ROWS(NODV, TE)                                                                                     \
    do {                                                                                                       \
        auto kern = nbk == 3 ? k_rows<NODV, TE, 3> : k_rows<NODV, TE, 2>;                                      \
        lds_attr(reinterpret_cast<const void *>(kern), lds);                                                   \
        hipLaunchKernelGGL(kern, grid, block, lds, s, nn, inc_ptr, pos, rowptr_n, diag, nnzN,                   \
                           static_cast<const TE *>(EB), static_cast<const TE *>(RE), R, Tint, Fext, Kval,      \
                           check_nan, status, blocks, nblocks);                                                \
    } while (0)
    switch (nod) {
        case 3: if (fp32) UVH_ROWS(3, float); else UVH_ROWS(3, double); break;
        case 6: if (fp32) UVH_ROWS(6, float); else UVH_ROWS(6, double); break;
        case 10: if (fp32) UVH_ROWS(10, float); else UVH_ROWS(10, double); break;
        default: return hipErrorInvalidValue;
    }
#undef UVH_ROWS
    return hipGetLastError();
}

size_t element_lds_bytes(int nod, int nip, int epb, bool fp32) {
    return element_table_bytes(nod, nip, fp32) + element_per_el_bytes(nod, nip, fp32) * size_t(epb);
}

// Batched launches. `ctr` holds 2 work counters; `status` one word per member.
// Workgroups of a persistent launch: every CU filled to the kernel's own occupancy at this LDS size.
// Computed once per (device, kernel, LDS size), so that captured launches make no queries.
int persistent_grid(const void *kernel, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void *, size_t>, int> cache;
    int dev = 0;
    hipGetDevice(&dev);
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({dev, kernel, lds});
    if (it != cache.end()) return it->second;
    int cus = 0, per_cu = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int grid = std::max(1, cus) * per_cu;
    cache[{dev, kernel, lds}] = grid;
    return grid;
}

hipError_t launch_rows_gather(int nod, int nn, int64_t nnzN, bool fp32, const int32_t *nzrow, const int64_t *rowptr_n,
                              const int32_t *inc_ptr, const int64_t *imap_ptr, const uint8_t *imap,
                              const uint8_t *diag, const void *EB, const void *RE, double *R, double *Tint,
                              double *Fext, double *Kval, int check_nan, int *status, hipStream_t s) {
    const int64_t nblocks = (nnzN + 255) / 256;
    const dim3 grid(unsigned(((nblocks + kXcds - 1) / kXcds) * kXcds)), block(256);
#define UVH_RG(NODV, TE)                                                                                   \
    hipLaunchKernelGGL((k_rows_gather<NODV, TE>), grid, block, 0, s, nn, nnzN, nzrow, rowptr_n, inc_ptr, imap_ptr, \
                       imap, diag, static_cast<const TE *>(EB), static_cast<const TE *>(RE), R, Tint, Fext, Kval,  \
                       check_nan, status)
    switch (nod) {
        case 3: if (fp32) UVH_RG(3, float); else UVH_RG(3, double); break;
        case 6: if (fp32) UVH_RG(6, float); else UVH_RG(6, double); break;
        case 10: if (fp32) UVH_RG(10, float); else UVH_RG(10, double); break;
        default: return hipErrorInvalidValue;
    }
#undef UVH_RG
    return hipGetLastError();
}

// Batched launches, one persistent element launch and one persistent row launch per element
// type present (each instantiation keeps its own register budget). Items are grouped by type:
// type k (3, 6, 10 nodes) owns elem_items[eb[k] .. eb[k+1]) and row_items[rb[k] .. rb[k+1]).
// `ctr` holds 6 work counters; `status` one word per member.
hipError_t launch_batch(const BatchMember *mem, int nmem, const int2 *nodal_items, int n_nodal,
                        const int2 *elem_items, const int *eb, const int2 *row_items, const int *rb,
                        const size_t *elem_lds, const size_t *row_lds, bool fp32, bool ronly, int check_nan,
                        int *ctr, int *status, const Params &p, bool flat, hipStream_t s) {
    hipLaunchKernelGGL(k_nodal_b_batch, dim3(n_nodal), dim3(256), 0, s, mem, nodal_items, p, ctr, 6, status, nmem);
    for (int k = 0; k < 3; ++k)
        if (elem_lds[k] > 156 * 1024 || row_lds[k] > 156 * 1024) return hipErrorInvalidConfiguration;
#define UVH_EB(NODV, T, RO, K)                                                                             \
    do {                                                                                                   \
        const int n_it = eb[K + 1] - eb[K];                                                                \
        if (flat) {                                                                                        \
            auto kern = p.law ? k_element_batch_flat<NODV, T, RO, true> : k_element_batch_flat<NODV, T, RO>; \
            lds_attr(reinterpret_cast<const void *>(kern), elem_lds[K]);                                   \
            const int grid = ((n_it + kXcds - 1) / kXcds) * kXcds;                                         \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), elem_lds[K], s, mem, elem_items + eb[K], n_it, p); \
        } else {                                                                                           \
            auto kern = k_element_batch<NODV, T, RO>;                                                      \
            lds_attr(reinterpret_cast<const void *>(kern), elem_lds[K]);                                   \
            const int grid = persistent_grid(reinterpret_cast<const void *>(kern), elem_lds[K]);           \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), elem_lds[K], s, mem, elem_items + eb[K], n_it, \
                               ctr + K, p);                                                                \
        }                                                                                                  \
    } while (0)
#define UVH_EB_T(NODV, K)                                                                                  \
    do {                                                                                                   \
        if (eb[K + 1] == eb[K]) break;                                                                     \
        if (fp32) { if (ronly) UVH_EB(NODV, float, true, K); else UVH_EB(NODV, float, false, K); }        \
        else { if (ronly) UVH_EB(NODV, double, true, K); else UVH_EB(NODV, double, false, K); }           \
    } while (0)
    UVH_EB_T(3, 0);
    UVH_EB_T(6, 1);
    UVH_EB_T(10, 2);
#undef UVH_EB_T
#undef UVH_EB
    if (ronly) {
        if (fp32)
            hipLaunchKernelGGL(k_rows_r_batch<float>, dim3(n_nodal), dim3(256), 0, s, mem, nodal_items, check_nan, status);
        else
            hipLaunchKernelGGL(k_rows_r_batch<double>, dim3(n_nodal), dim3(256), 0, s, mem, nodal_items, check_nan, status);
        return hipGetLastError();
    }
#define UVH_RB(NODV, TE, K)                                                                                \
    do {                                                                                                   \
        const int n_it = rb[K + 1] - rb[K];                                                                \
        if (flat) {                                                                                        \
            auto kern = k_rows_batch_flat<NODV, TE>;                                                       \
            lds_attr(reinterpret_cast<const void *>(kern), row_lds[K]);                                    \
            const int grid = ((n_it + kXcds - 1) / kXcds) * kXcds;                                         \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), row_lds[K], s, mem, row_items + rb[K], n_it,   \
                               check_nan, status);                                                         \
        } else {                                                                                           \
            auto kern = k_rows_batch<NODV, TE>;                                                            \
            lds_attr(reinterpret_cast<const void *>(kern), row_lds[K]);                                    \
            const int grid = persistent_grid(reinterpret_cast<const void *>(kern), row_lds[K]);            \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), row_lds[K], s, mem, row_items + rb[K], n_it,   \
                               ctr + 3 + K, check_nan, status);                                            \
        }                                                                                                  \
    } while (0)
#define UVH_RB_T(NODV, K)                                                                                  \
    do {                                                                                                   \
        if (rb[K + 1] == rb[K]) break;                                                                     \
        if (fp32) UVH_RB(NODV, float, K); else UVH_RB(NODV, double, K);                                    \
    } while (0)
    UVH_RB_T(3, 0);
    UVH_RB_T(6, 1);
    UVH_RB_T(10, 2);
#undef UVH_RB_T
#undef UVH_RB
    return hipGetLastError();
}

// LDS layout of the fused kernel: shape tables, nodal staging and coefficient records of one
// batch, ordered staging, row window; within kFusedLdsBytes so that two workgroups share a CU.
// The batch takes about 45 % of the budget (P1: 24 elements, P2: 12, P3: 4 at the configs'
// rules); the ordered staging starts at three items per element (the corners) and grows if a
// schedule needs more (set_mesh rebuilds it with the smaller window).
FusedLayout fused_layout(int nod, int nip) {
    FusedLayout L{};
    L.table_doubles = (3 * nip * nod + nip + 1) & ~1;
    const int per_el = nv_stride_host(nod) + rec_stride(nip);
    const int64_t budget = kFusedLdsBytes / int64_t(sizeof(double)) - L.table_doubles;
    L.epb = std::max(2, std::min(64, int(0.45 * double(budget) / double(per_el))));
    if (const char *e = std::getenv("UVH_FUSED_EPB")) L.epb = std::max(2, std::min(64, std::atoi(e)));
    L.nodal_doubles = (L.epb * nv_stride_host(nod) + 1) & ~1;
    L.rec_doubles = (L.epb * rec_stride(nip) + 1) & ~1;
    L.stage_items = L.epb * std::min(nod, 3);
    L.window_doubles = ((budget - L.nodal_doubles - L.rec_doubles - int64_t(L.stage_items) * kStageValues) / 2) * 2;
    return L;
}

hipError_t launch_fused(int nod, int nip, int npatch, const FusedLayout &L, const MeshArgs *MA, const CallArgs *CA,
                        int64_t nnzN, int nn, bool sl, bool ronly, int check_nan, hipStream_t s) {
    if (npatch <= 0) return hipSuccess;
    const size_t lds = size_t(L.lds_bytes());
    const dim3 grid(((npatch + kXcds - 1) / kXcds) * kXcds), block(kFusedThreads);
#define UVH_FUSED(NODV, SLV)                                                                                  \
    do {                                                                                                      \
        auto kern = ronly ? k_fused<NODV, SLV, true> : k_fused<NODV, SLV, false>;                             \
        lds_attr(reinterpret_cast<const void *>(kern), lds);                                                  \
        hipLaunchKernelGGL(kern, grid, block, lds, s, MA, CA, nip, L.epb, L.table_doubles, L.nodal_doubles,   \
                           L.rec_doubles, L.stage_items, npatch, nnzN, nn, check_nan);                        \
    } while (0)
    switch (nod) {
        case 3: if (sl) UVH_FUSED(3, true); else UVH_FUSED(3, false); break;
        case 6: if (sl) UVH_FUSED(6, true); else UVH_FUSED(6, false); break;
        case 10: if (sl) UVH_FUSED(10, true); else UVH_FUSED(10, false); break;
        default: return hipErrorInvalidValue;
    }
#undef UVH_FUSED
    return hipGetLastError();
}
}  // namespace uvh
