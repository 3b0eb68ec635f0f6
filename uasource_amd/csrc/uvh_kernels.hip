// HIP kernels of the uvh assembly for gfx950 (CDNA4, wave64).
//
// Three kernels per assembly:
//   k_nodal_b    per node: lower surface b from floatation (UaMain/Calc_bs_From_hBS.m:71-79,109)
//   k_element    per element: gather of nodal fields through the connectivity into LDS,
//                Gauss-point physics (UaMain/uvhAssemblyIntPointImplicitSUPG.m:62-395) and the
//                element residual rows and Jacobian rows (UaMain/uvhNodalLoopSSTREAM.m:39-228)
//   k_rows       per node: deterministic sum of the element rows into R (Tint - Fext) and into
//                the CSR rows of K (UaMain/uvhMatrixAssembly.m:416-509); incident elements are
//                summed in ascending element order, so results are bitwise reproducible.
//
// Thread mapping of k_element: a 256-thread block owns `epb` elements. Threads first gather
// (element, node) pairs into LDS, then compute the physics of (element, integration point)
// pairs and publish a record of per-point coefficients in LDS, then accumulate (element, row)
// pairs: 6 residual terms and 9 x nod Jacobian terms per row, over all points, in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "uvh_internal.h"

namespace uvh {

// nodal values staged per element in LDS
enum : int {
    NV_H, NV_U, NV_V, NV_S, NV_B, NV_RHO, NV_C, NV_M, NV_N, NV_A, NV_AS1, NV_AB1,
    NV_H0, NV_U0, NV_V0, NV_AS0, NV_AB0, NV_DADH, NV_BL, NV_X, NV_Y, NV_COUNT
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

// Element rows in HBM (two-pass path) are padded to whole 128-byte lines.
template <int NOD>
constexpr int eb_stride() { return ((9 * NOD + 15) / 16) * 16; }

__global__ void k_nodal_b(int nn, FieldPtrs f, Params p, double *__restrict__ b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double h = p.zero_fields ? f.h0[k] : f.h[k];
    if (p.reset_thickness && h < p.thick_min) h = p.thick_min;
    const double S = f.S[k], B = f.B[k], rho = f.rho[k];
    const double hf = p.rhow * (S - B) / rho;
    const double GF = 0.5 + 0.5 * tanh(p.kH * ((h - hf) - p.Hh0));
    const double bfloat = S - rho * h / p.rhow;
    double bb = GF * B + (1.0 - GF) * bfloat;
    if (p.enforce_b_above && bb < B) bb = B;
    b[k] = bb;
}

__global__ void k_clear_status(int *status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *status = 0;
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

// Smooth step and its derivative at x = k (z - z0) (HeavisideApprox.m:101, DiracDelta.m:29):
//   H = 1/2 + 1/2 tanh(x),   D = k/2 sech^2(x).
// One expm1 gives both: with t = exp(-2|x|), tanh|x| = (1 - t)/(1 + t) and
// sech^2 x = 4t/(1 + t)^2, each to a few ulp for every x.
__device__ __forceinline__ void soft_step(double x, double k, double &H, double &D) {
    const double em = expm1(-2.0 * fabs(x));  // t - 1, in (-1, 0]
    const double den = 2.0 + em;              // 1 + t
    const double th = copysign(-em / den, x);
    H = 0.5 + 0.5 * th;
    D = 0.5 * k * (4.0 * (1.0 + em) / (den * den));
}

// Physics of one integration point: fills the coefficient record c[kNC].
template <int NOD>
__device__ void ip_physics(const double *__restrict__ nv,  // [NV_COUNT][NOD] in LDS
                           const double *__restrict__ tN, const double *__restrict__ tXi,
                           const double *__restrict__ tEta, double wq, const Params &p,
                           double *__restrict__ c) {
    // Jacobian and inverse (UaMain/derivVector.m:96-107). Coordinates are taken relative to
    // the element's first node: sum_i dN_i = 0, so J is unchanged, and the differences are
    // exact, which avoids cancelling ~1e5 m coordinates down to ~1e2 m element sizes.
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    const double xo = nv[NV_X * NOD], yo = nv[NV_Y * NOD];
#pragma unroll
    for (int i = 1; i < NOD; ++i) {
        const double xi = nv[NV_X * NOD + i] - xo, yi = nv[NV_Y * NOD + i] - yo;
        J11 += tXi[i] * xi; J12 += tXi[i] * yi;
        J21 += tEta[i] * xi; J22 += tEta[i] * yi;
    }
    const double det = J11 * J22 - J12 * J21;
    const double rdet = 1.0 / det;
    const double i11 = J22 * rdet, i12 = -J12 * rdet, i21 = -J21 * rdet, i22 = J11 * rdet;

    // interpolation and gradients (uvhAssemblyIntPointImplicitSUPG.m:62-126,206-208,272-300)
    double h = 0, u = 0, v = 0, C = 0, m = 0, A = 0, n = 0, h0 = 0, u0 = 0, v0 = 0;
    double as0 = 0, ab0 = 0, as1 = 0, ab1 = 0, dadh = 0, Bq = 0, Sq = 0, rho = 0;
    double dhdx = 0, dhdy = 0, dh0dx = 0, dh0dy = 0, exx0 = 0, eyy0 = 0, dbdx = 0, dbdy = 0;
    double drhodx = 0, drhody = 0, exx = 0, eyy = 0, exy2 = 0;
#pragma unroll 1
    for (int i = 0; i < NOD; ++i) {
        const double Ni = tN[i];
        const double nx = i11 * tXi[i] + i12 * tEta[i];
        const double ny = i21 * tXi[i] + i22 * tEta[i];
        const double hi = nv[NV_H * NOD + i], ui = nv[NV_U * NOD + i], vi = nv[NV_V * NOD + i];
        const double h0i = nv[NV_H0 * NOD + i], u0i = nv[NV_U0 * NOD + i], v0i = nv[NV_V0 * NOD + i];
        const double rhoi = nv[NV_RHO * NOD + i], bi = nv[NV_BL * NOD + i];
        h += Ni * hi; u += Ni * ui; v += Ni * vi;
        C += Ni * nv[NV_C * NOD + i]; m += Ni * nv[NV_M * NOD + i];
        A += Ni * nv[NV_A * NOD + i]; n += Ni * nv[NV_N * NOD + i];
        h0 += Ni * h0i; u0 += Ni * u0i; v0 += Ni * v0i;
        as0 += Ni * nv[NV_AS0 * NOD + i]; ab0 += Ni * nv[NV_AB0 * NOD + i];
        as1 += Ni * nv[NV_AS1 * NOD + i]; ab1 += Ni * nv[NV_AB1 * NOD + i];
        dadh += Ni * nv[NV_DADH * NOD + i];
        Bq += Ni * nv[NV_B * NOD + i]; Sq += Ni * nv[NV_S * NOD + i];
        rho += Ni * rhoi;
        dhdx += nx * hi; dhdy += ny * hi;
        dh0dx += nx * h0i; dh0dy += ny * h0i;
        exx0 += nx * u0i; eyy0 += ny * v0i;
        dbdx += nx * bi; dbdy += ny * bi;
        drhodx += nx * rhoi; drhody += ny * rhoi;
        exx += nx * ui; eyy += ny * vi;
        exy2 += nx * vi + ny * ui;
    }
    const double exy = 0.5 * exy2;
    const double a1 = as1 + ab1, a0 = as0 + ab0;
    const double rrhow = 1.0 / p.rhow;

    // floatation (uvhAssemblyIntPointImplicitSUPG.m:209-243): He = H(h - hf), HE = H(hf - h),
    // delta, Delta their derivatives; with Hh0 = 0 the second pair mirrors the first.
    const double H = Sq - Bq;
    const double hf = p.rhow * H / rho;
    double He, delta, HE, Delta, HH, unused;
    soft_step(p.kH * ((h - hf) - p.Hh0), p.kH, He, delta);
    if (p.Hh0 == 0.0) {
        HE = 1.0 - He;
        Delta = delta;
    } else {
        soft_step(p.kH * ((hf - h) - p.Hh0), p.kH, HE, Delta);
    }
    soft_step(p.kH * (H - p.Hh0), p.kH, HH, unused);
    const double Hpos = HH * H;
    const double d = HE * rho * h * rrhow + He * Hpos;
    const double dd = HE * rho * rrhow - Delta * h * rho * rrhow + delta * Hpos;

    // Glen's flow law (EffectiveViscositySSTREAM.m:7-9) from logarithms:
    //   A^(-1/n) e^((1-n)/n) = exp(-log(A)/n + (1-n)/n * log(e)),  E = (1-n)/(4n) * that / e^2
    const double e2 = p.eps0 * p.eps0 + exx * exx + eyy * eyy + exx * eyy + exy * exy;
    const double rn = 1.0 / n;
    const double visc = exp(-log(A) * rn + (1.0 - n) * rn * (0.5 * log(e2)));
    const double eta = 0.5 * visc + p.eta0;
    const double E = 0.25 * (1.0 - n) * rn * visc / e2;

    // Weertman sliding (BasalDrag.m:54-77,410-443):
    //   beta2 = (C+C0)^(-1/m) U^(1/m-1),  Dbeta2 = (1/m-1) beta2 / U^2
    const double Hec = He < 0 ? 0.0 : He, dlc = delta < 0 ? 0.0 : delta;
    const double U2 = u * u + v * v + p.uz * p.uz;
    const double rm = 1.0 / m;
    const double beta2 = exp(-log(C + p.czero) * rm + (rm - 1.0) * (0.5 * log(U2)));
    const double Db = (rm - 1.0) * beta2 / U2;
    const double taux = Hec * beta2 * u, tauy = Hec * beta2 * v;
    const double dtxdu = Hec * (beta2 + Db * u * u), dtydv = Hec * (beta2 + Db * v * v);
    const double dtxdv = Hec * Db * u * v;
    const double dtxdh = dlc * beta2 * u, dtydh = dlc * beta2 * v;

    // SUPG (uvhAssemblyIntPointImplicitSUPG.m:323-390, SUPGtau.m:24-77)
    int c0i, c1i, c2i;
    corner_ids<NOD>(c0i, c1i, c2i);
    const double X1 = nv[NV_X * NOD + c0i], X2 = nv[NV_X * NOD + c1i], X3 = nv[NV_X * NOD + c2i];
    const double Y1 = nv[NV_Y * NOD + c0i], Y2 = nv[NV_Y * NOD + c1i], Y3 = nv[NV_Y * NOD + c2i];
    const double area = 0.5 * fabs((X1 - X3) * (Y2 - Y1) - (X1 - X2) * (Y3 - Y1));
    const double l = sqrt(2.0 * area);
    const double speed0 = sqrt(u0 * u0 + v0 * v0 + p.uz * p.uz);
    const double taut = p.dt / 2.0 + 2.220446049250313e-16;
    const double taus = l / (2.0 * speed0 + p.uz);
    double tau;
    switch (p.tau_opt) {
        case 1: tau = taut; break;
        case 2: {
            const double ecn = taut / taus;
            const double K = ecn < 1e-6 ? ecn / 3.0 : 1.0 / tanh(ecn) - 1.0 / ecn;
            tau = K * taus;
        } break;
        case 3: tau = (p.dt / 2.0) * 1.0 / (1.0 + taut / taus); break;
        case 4: tau = 0.0; break;
        default: tau = taus;
    }
    const double tau0 = p.beta0 * (p.tau_mult * tau);
    const double W = det * wq;

    // fluxes (uvhNodalLoopSSTREAM.m:39-44)
    const double qx1 = rho * exx * h + rho * u * dhdx + drhodx * u * h;
    const double qy1 = rho * eyy * h + rho * v * dhdy + drhody * v * h;
    const double qx0 = rho * exx0 * h0 + rho * u0 * dh0dx + drhodx * u0 * h0;
    const double qy0 = rho * eyy0 * h0 + rho * v0 * dh0dy + drhody * v0 * h0;
    const double th = p.theta, dt = p.dt, g = p.g, ca = p.ca, sa = p.sa, rhow = p.rhow;
    const double dtth = dt * th;

    c[C_I11] = i11; c[C_I12] = i12; c[C_I21] = i21; c[C_I22] = i22;
    c[C_TU] = 0.5 * tau0 * u0; c[C_TV] = 0.5 * tau0 * v0;
    c[C_WHETA] = W * h * eta; c[C_WHE2] = 2.0 * W * h * E;
    c[C_S1] = 2.0 * exx + eyy; c[C_S2] = 2.0 * eyy + exx; c[C_EXY] = exy;
    c[C_WDTXDU] = W * dtxdu; c[C_WDTYDV] = W * dtydv; c[C_WDTXDV] = W * dtxdv;
    const double c2 = ca * g * (rho * h - rhow * d * dd);
    const double cx1 = dtxdh + ca * g * (rho - rhow * dd) * dbdx - sa * g * rho;
    const double cy1 = dtydh + ca * g * (rho - rhow * dd) * dbdy;
    c[C_PX] = W * (eta * (4.0 * exx + 2.0 * eyy) - c2);
    c[C_PXY] = W * (eta * 2.0 * exy);
    c[C_PCX] = W * cx1;
    c[C_PY] = W * (eta * (4.0 * eyy + 2.0 * exx) - c2);
    c[C_PCY] = W * cy1;
    c[C_W] = W;
    c[C_AX] = dtth * (rho * dhdx + drhodx * h);
    c[C_AY] = dtth * (rho * dhdy + drhody * h);
    c[C_RH] = dtth * (rho * h);
    c[C_GAM] = rho - dtth * rho * dadh + dtth * (rho * exx + drhodx * u + rho * eyy + drhody * v);
    c[C_RU] = dtth * (rho * u);
    c[C_RV] = dtth * (rho * v);
    c[C_WTX] = W * taux; c[C_WTY] = W * tauy;
    c[C_WF1X] = W * (-ca * g * (rho * h - rhow * d) * dbdx + rho * g * h * sa);
    c[C_WF1Y] = W * (-ca * g * (rho * h - rhow * d) * dbdy);
    c[C_WF2] = W * (0.5 * ca * g * (rho * h * h - rhow * d * d));
    c[C_TH] = -W * (rho * (h0 - h));
    c[C_FH] = W * (dt * rho * ((1.0 - th) * a0 + th * a1) - dt * (th * qx1 + (1.0 - th) * qx0 + th * qy1 + (1.0 - th) * qy0));
}

// Gather of node a of element e into the LDS staging record nv[NV_COUNT][NOD].
template <int NOD>
__device__ __forceinline__ void gather_node(const int32_t *__restrict__ conn, const double *__restrict__ xn,
                                            const double *__restrict__ yn, const FieldPtrs &f,
                                            const double *__restrict__ bnod, const Params &p, int64_t e, int a,
                                            double *__restrict__ nv) {
    const int k = conn[e * NOD + a];
    double h = f.h[k], u = f.u[k], v = f.v[k], u0 = f.u0[k], v0 = f.v0[k];
    double as1 = f.as1[k], ab1 = f.ab1[k], as0 = f.as0[k], ab0 = f.ab0[k];
    const double h0 = f.h0[k];
    if (p.zero_fields) {  // uvhMatrixAssembly.m:73-87
        u = 0.0; v = 0.0; u0 = 0.0; v0 = 0.0; h = h0;
        as1 = fabs(ab1) + 1.0; ab1 = fabs(ab1);
        as0 = fabs(ab0) + 1.0; ab0 = fabs(ab0);
    }
    nv[NV_H * NOD + a] = h; nv[NV_U * NOD + a] = u; nv[NV_V * NOD + a] = v;
    nv[NV_S * NOD + a] = f.S[k]; nv[NV_B * NOD + a] = f.B[k]; nv[NV_RHO * NOD + a] = f.rho[k];
    nv[NV_C * NOD + a] = f.C[k]; nv[NV_M * NOD + a] = f.m[k]; nv[NV_N * NOD + a] = f.n[k];
    nv[NV_A * NOD + a] = f.A[k]; nv[NV_AS1 * NOD + a] = as1; nv[NV_AB1 * NOD + a] = ab1;
    nv[NV_H0 * NOD + a] = h0; nv[NV_U0 * NOD + a] = u0; nv[NV_V0 * NOD + a] = v0;
    nv[NV_AS0 * NOD + a] = as0; nv[NV_AB0 * NOD + a] = ab0;
    nv[NV_DADH * NOD + a] = f.dadh ? f.dadh[k] : 0.0;
    nv[NV_BL * NOD + a] = bnod[k];
    nv[NV_X * NOD + a] = xn[k]; nv[NV_Y * NOD + a] = yn[k];
}

// Row i of the element: 6 residual terms and (unless RONLY) 9 x NOD Jacobian terms,
// accumulated over the integration points from the coefficient records cr[nip][kNC].
template <int NOD, bool RONLY>
__device__ __forceinline__ void element_row(int i, int nip, const double *__restrict__ cr,
                                            const double *__restrict__ sN, const double *__restrict__ sXi,
                                            const double *__restrict__ sEta, double (&Rr)[6],
                                            double (&K)[RONLY ? 1 : 9][NOD]) {
    double Tx = 0, Fx = 0, Ty = 0, Fy = 0, Th = 0, Fh = 0;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 9; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) K[b][j] = 0.0;
    }
    for (int q = 0; q < nip; ++q) {
        const double *r = cr + q * kNC;
        const double *qN = sN + q * NOD, *qXi = sXi + q * NOD, *qEta = sEta + q * NOD;
        const double I11 = r[C_I11], I12 = r[C_I12], I21 = r[C_I21], I22 = r[C_I22];
        const double S1 = r[C_S1], S2 = r[C_S2], EXY = r[C_EXY];
        const double Ni = qN[i];
        const double Nxi = I11 * qXi[i] + I12 * qEta[i];
        const double Nyi = I21 * qXi[i] + I22 * qEta[i];
        const double ai = S1 * Nxi + EXY * Nyi;
        const double bi = S2 * Nyi + EXY * Nxi;
        const double psi = Ni + r[C_TU] * Nxi + r[C_TV] * Nyi;
        const double wheta = r[C_WHETA];
        Tx += 2.0 * wheta * ai + r[C_WTX] * Ni;
        Ty += 2.0 * wheta * bi + r[C_WTY] * Ni;
        Fx += r[C_WF1X] * Ni + r[C_WF2] * Nxi;
        Fy += r[C_WF1Y] * Ni + r[C_WF2] * Nyi;
        Th += r[C_TH] * psi;
        Fh += r[C_FH] * psi;
        if (!RONLY) {
            const double whE2 = r[C_WHE2];
            const double l1 = 4.0 * wheta * Nxi, l2 = wheta * Nyi, l3 = whE2 * ai, l4 = r[C_WDTXDU] * Ni;
            const double m1 = 4.0 * wheta * Nyi, m2 = wheta * Nxi, m3 = whE2 * bi, m4 = r[C_WDTYDV] * Ni;
            const double n1 = 2.0 * wheta * Nxi, n4 = r[C_WDTXDV] * Ni;
            const double o1 = 2.0 * wheta * Nyi;
            const double Xi = r[C_PX] * Nxi + r[C_PXY] * Nyi + r[C_PCX] * Ni;
            const double Yi = r[C_PY] * Nyi + r[C_PXY] * Nxi + r[C_PCY] * Ni;
            const double Wpsi = r[C_W] * psi;
            const double hAX = Wpsi * r[C_AX], hAY = Wpsi * r[C_AY], hRH = Wpsi * r[C_RH];
            const double hG = Wpsi * r[C_GAM], hRU = Wpsi * r[C_RU], hRV = Wpsi * r[C_RV];
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const double Nj = qN[j];
                const double Nxj = I11 * qXi[j] + I12 * qEta[j];
                const double Nyj = I21 * qXi[j] + I22 * qEta[j];
                const double aj = S1 * Nxj + EXY * Nyj;
                const double bj = S2 * Nyj + EXY * Nxj;
                K[0][j] += l1 * Nxj + l2 * Nyj + l3 * aj + l4 * Nj;   // Kxu
                K[1][j] += n1 * Nyj + l2 * Nxj + l3 * bj + n4 * Nj;   // Kxv
                K[2][j] += Xi * Nj;                                    // Kxh
                K[3][j] += o1 * Nxj + m2 * Nyj + m3 * aj + n4 * Nj;   // Kyu
                K[4][j] += m1 * Nyj + m2 * Nxj + m3 * bj + m4 * Nj;   // Kyv
                K[5][j] += Yi * Nj;                                    // Kyh
                K[6][j] += hAX * Nj + hRH * Nxj;                       // Khu
                K[7][j] += hAY * Nj + hRH * Nyj;                       // Khv
                K[8][j] += hG * Nj + hRU * Nxj + hRV * Nyj;            // Khh
            }
        }
    }
    Rr[0] = Tx; Rr[1] = Fx; Rr[2] = Ty; Rr[3] = Fy; Rr[4] = Th; Rr[5] = Fh;
}

// Row i, block row BI (0: x, 1: y, 2: h) of the element: T and F of that row and its
// 3 x NOD Jacobian blocks, accumulated over the integration points.
template <int NOD, int BI>
__device__ __forceinline__ void element_row_b(int i, int nip, const double *__restrict__ cr,
                                              const double *__restrict__ sN, const double *__restrict__ sXi,
                                              const double *__restrict__ sEta, double &Tr, double &Fr,
                                              double (&K)[3][NOD]) {
    double T = 0, F = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int j = 0; j < NOD; ++j) K[b][j] = 0.0;
    for (int q = 0; q < nip; ++q) {
        const double *r = cr + q * kNC;
        const double *qN = sN + q * NOD, *qXi = sXi + q * NOD, *qEta = sEta + q * NOD;
        const double I11 = r[C_I11], I12 = r[C_I12], I21 = r[C_I21], I22 = r[C_I22];
        const double S1 = r[C_S1], S2 = r[C_S2], EXY = r[C_EXY];
        const double Ni = qN[i];
        const double Nxi = I11 * qXi[i] + I12 * qEta[i];
        const double Nyi = I21 * qXi[i] + I22 * qEta[i];
        const double wheta = r[C_WHETA];
        if (BI == 0) {
            const double ai = S1 * Nxi + EXY * Nyi;
            T += 2.0 * wheta * ai + r[C_WTX] * Ni;
            F += r[C_WF1X] * Ni + r[C_WF2] * Nxi;
            const double whE2 = r[C_WHE2];
            const double l1 = 4.0 * wheta * Nxi, l2 = wheta * Nyi, l3 = whE2 * ai, l4 = r[C_WDTXDU] * Ni;
            const double n1 = 2.0 * wheta * Nxi, n4 = r[C_WDTXDV] * Ni;
            const double Xi = r[C_PX] * Nxi + r[C_PXY] * Nyi + r[C_PCX] * Ni;
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const double Nj = qN[j];
                const double Nxj = I11 * qXi[j] + I12 * qEta[j];
                const double Nyj = I21 * qXi[j] + I22 * qEta[j];
                const double aj = S1 * Nxj + EXY * Nyj;
                const double bj = S2 * Nyj + EXY * Nxj;
                K[0][j] += l1 * Nxj + l2 * Nyj + l3 * aj + l4 * Nj;
                K[1][j] += n1 * Nyj + l2 * Nxj + l3 * bj + n4 * Nj;
                K[2][j] += Xi * Nj;
            }
        } else if (BI == 1) {
            const double bi = S2 * Nyi + EXY * Nxi;
            T += 2.0 * wheta * bi + r[C_WTY] * Ni;
            F += r[C_WF1Y] * Ni + r[C_WF2] * Nyi;
            const double whE2 = r[C_WHE2];
            const double m1 = 4.0 * wheta * Nyi, m2 = wheta * Nxi, m3 = whE2 * bi, m4 = r[C_WDTYDV] * Ni;
            const double o1 = 2.0 * wheta * Nyi, n4 = r[C_WDTXDV] * Ni;
            const double Yi = r[C_PY] * Nyi + r[C_PXY] * Nxi + r[C_PCY] * Ni;
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const double Nj = qN[j];
                const double Nxj = I11 * qXi[j] + I12 * qEta[j];
                const double Nyj = I21 * qXi[j] + I22 * qEta[j];
                const double aj = S1 * Nxj + EXY * Nyj;
                const double bj = S2 * Nyj + EXY * Nxj;
                K[0][j] += o1 * Nxj + m2 * Nyj + m3 * aj + n4 * Nj;
                K[1][j] += m1 * Nyj + m2 * Nxj + m3 * bj + m4 * Nj;
                K[2][j] += Yi * Nj;
            }
        } else {
            const double psi = Ni + r[C_TU] * Nxi + r[C_TV] * Nyi;
            T += r[C_TH] * psi;
            F += r[C_FH] * psi;
            const double Wpsi = r[C_W] * psi;
            const double hAX = Wpsi * r[C_AX], hAY = Wpsi * r[C_AY], hRH = Wpsi * r[C_RH];
            const double hG = Wpsi * r[C_GAM], hRU = Wpsi * r[C_RU], hRV = Wpsi * r[C_RV];
#pragma unroll
            for (int j = 0; j < NOD; ++j) {
                const double Nj = qN[j];
                const double Nxj = I11 * qXi[j] + I12 * qEta[j];
                const double Nyj = I21 * qXi[j] + I22 * qEta[j];
                K[0][j] += hAX * Nj + hRH * Nxj;
                K[1][j] += hAY * Nj + hRH * Nyj;
                K[2][j] += hG * Nj + hRU * Nxj + hRV * Nyj;
            }
        }
    }
    Tr = T; Fr = F;
}

// Per-element kernel: EPB elements per 256-thread block (runtime), dense thread mappings
// for the three stages: (element, node) gathers, (element, point) physics, (element, row)
// rows. The caller guarantees epb*NOD <= 256 and epb*nip <= 256.
template <int NOD, bool RONLY>
__global__ __launch_bounds__(256) void k_element(int nele, int nip, int epb, const int32_t *__restrict__ conn,
                                                 const double *__restrict__ xn, const double *__restrict__ yn,
                                                 FieldPtrs f, const double *__restrict__ bnod,
                                                 const ShapeTables *__restrict__ tab, Params p,
                                                 const int32_t *__restrict__ inc_pos,
                                                 double *__restrict__ EB, double *__restrict__ RE) {
    extern __shared__ double lds[];
    const int tstride = nip * NOD;
    double *sN = lds;                       // [nip][NOD]
    double *sXi = sN + tstride;
    double *sEta = sXi + tstride;
    double *sW = sEta + tstride;            // [nip]
    double *work = lds + ((3 * tstride + nip + 1) & ~1);  // 16-byte aligned union:
                                            // nodal [epb][NV_COUNT][NOD] / coef [epb][nip][kNC]
    const int tid = threadIdx.x;
    for (int k = tid; k < tstride; k += 256) {
        const int q = k / NOD, i = k - q * NOD;
        sN[k] = tab->N[q][i]; sXi[k] = tab->dxi[q][i]; sEta[k] = tab->deta[q][i];
    }
    if (tid < nip) sW[tid] = tab->w[tid];
    const int nblocks = (nele + epb - 1) / epb;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    if (lb >= nblocks) return;  // whole block: no barrier has been reached
    const int64_t e0 = (int64_t)lb * epb;

    // stage 1: gathers, one (element, node) per thread
    if (tid < epb * NOD) {
        const int el = tid / NOD, a = tid - el * NOD;
        if (e0 + el < nele) gather_node<NOD>(conn, xn, yn, f, bnod, p, e0 + el, a, work + el * (NV_COUNT * NOD));
    }
    __syncthreads();

    // stage 2: physics, one (element, integration point) per thread
    double c[kNC];
    const int pel = tid / nip, pq = tid - pel * nip;
    const bool phys = tid < epb * nip && e0 + pel < nele;
    if (phys) ip_physics<NOD>(work + pel * (NV_COUNT * NOD), sN + pq * NOD, sXi + pq * NOD, sEta + pq * NOD,
                              sW[pq], p, c);
    __syncthreads();  // the nodal staging is dead from here; the coefficient records reuse it
    if (phys) {
        double *dst = work + (pel * nip + pq) * kNC;
#pragma unroll
        for (int k = 0; k < kNC; k += 2) {
            double2 t;
            t.x = c[k];
            t.y = (k + 1 < C_COUNT) ? c[k + 1] : 0.0;
            *reinterpret_cast<double2 *>(dst + k) = t;
        }
    }
    __syncthreads();

    // stage 3: rows, one (element, row) per thread
    if (tid >= epb * NOD) return;
    const int el = tid / NOD, i = tid - el * NOD;
    const int64_t e = e0 + el;
    if (e >= nele) return;
    // the three block rows one after the other: 3*NOD accumulators instead of 9*NOD;
    // the row is stored at its position in the node-incidence order
    const int64_t row = inc_pos[e * NOD + i];
    double *re = RE + row * 6;
    double *eb = EB + row * eb_stride<NOD>();
    const double *crr = work + el * nip * kNC;
    double T, F, Kb[3][NOD];
    element_row_b<NOD, 0>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
    re[0] = T; re[1] = F;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = Kb[b][j];
    }
    element_row_b<NOD, 1>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
    re[2] = T; re[3] = F;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) eb[(3 + b) * NOD + j] = Kb[b][j];
    }
    element_row_b<NOD, 2>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
    re[4] = T; re[5] = F;
    if (!RONLY) {
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) eb[(6 + b) * NOD + j] = Kb[b][j];
    }
}

// Node rows: one wave per node. The incident element rows (stored contiguously per node, in
// incidence order, by k_element) are added, in ascending element order, into a per-wave LDS buffer that
// holds the node's three CSR rows (9*deg values); within one element row no two values
// share a column, so every sum has a fixed order and the result is bitwise reproducible.
// The buffer is then written to K coalesced along the rows.
template <int NOD>
__global__ __launch_bounds__(256) void k_rows(int nn, int maxdeg, const int32_t *__restrict__ inc_ptr,
                                              const int32_t *__restrict__ inc, const uint8_t *__restrict__ pos,
                                              const int64_t *__restrict__ rowptr_n, const uint8_t *__restrict__ diag,
                                              int64_t nnzN, const double *__restrict__ EB,
                                              const double *__restrict__ RE, double *__restrict__ R,
                                              double *__restrict__ Tint, double *__restrict__ Fext,
                                              double *__restrict__ Kval, int check_nan, int *__restrict__ status) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (k >= nn) return;  // whole wave; no block-wide barriers below
    double *buf = lds + wave * 9 * maxdeg;
    const int t0 = inc_ptr[k], t1 = inc_ptr[k + 1];
    int bad = 0;
    if (lane < 6) {
        // lane 2*bi: T of block row bi, lane 2*bi+1: F
        double v = 0;
        for (int t = t0; t < t1; ++t) v += RE[(int64_t)t * 6 + lane];
        const int bi = lane >> 1;
        if (lane & 1) {
            if (Fext) Fext[(int64_t)bi * nn + k] = v;
        } else {
            if (Tint) Tint[(int64_t)bi * nn + k] = v;
        }
        const double other = __shfl_xor(v, 1, 64);
        if (!(lane & 1)) {
            const double r = v - other;
            R[(int64_t)bi * nn + k] = r;
            if (check_nan && isnan(r)) bad |= 1;
        }
    }
    if (Kval) {
        const int64_t rp = rowptr_n[k];
        const int deg = (int)(rowptr_n[k + 1] - rp);
        for (int v = lane; v < 9 * deg; v += 64) buf[v] = 0.0;
        // the node's element rows t0..t1-1 are contiguous (incidence order); up to 8 at a
        // time: all loads first, then the adds in ascending element order
        constexpr int NV = 9 * NOD;               // values per element row
        constexpr int VPL = (NV + 63) / 64;       // values per lane
        for (int c0 = t0; c0 < t1; c0 += 8) {
            const int cnt = min(8, t1 - c0);
            double vals[8][VPL];
            int cols[8][VPL];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
#pragma unroll
                for (int u = 0; u < VPL; ++u) {
                    const int v = lane + 64 * u;
                    if (c < cnt && v < NV) {
                        const int bk = v / NOD, j = v - bk * NOD;
                        vals[c][u] = EB[(int64_t)(c0 + c) * eb_stride<NOD>() + v];
                        cols[c][u] = pos[(int64_t)(c0 + c) * NOD + j];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
#pragma unroll
                for (int u = 0; u < VPL; ++u) {
                    const int v = lane + 64 * u;
                    if (c < cnt && v < NV) {
                        const int bk = v / NOD;
                        const int bi = bk / 3, bj = bk - bi * 3;
                        lds_add(buf + bi * 3 * deg + bj * deg + cols[c][u], vals[c][u]);
                    }
                }
            }
        }
        for (int v = lane; v < 9 * deg; v += 64) {
            const int bi = v / (3 * deg);
            Kval[(int64_t)bi * 3 * nnzN + 3 * rp + (v - bi * 3 * deg)] = buf[v];
        }
        if (check_nan && lane < 3 && isnan(buf[lane * 3 * deg + lane * deg + diag[k]])) bad |= 2;
    }
    if (bad) atomicOr(status, bad);
}

// Residual only: one thread per node sums its incident element residual rows (contiguous).
__global__ __launch_bounds__(256) void k_rows_r(int nn, const int32_t *__restrict__ inc_ptr,
                                                const int32_t *__restrict__ inc, const double *__restrict__ RE,
                                                double *__restrict__ R, double *__restrict__ Tint,
                                                double *__restrict__ Fext, int check_nan, int *__restrict__ status) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double T0 = 0, T1 = 0, T2 = 0, F0 = 0, F1 = 0, F2 = 0;
    for (int t = inc_ptr[k]; t < inc_ptr[k + 1]; ++t) {
        const double *re = RE + (int64_t)t * 6;
        T0 += re[0]; F0 += re[1]; T1 += re[2]; F1 += re[3]; T2 += re[4]; F2 += re[5];
    }
    const double r0 = T0 - F0, r1 = T1 - F1, r2 = T2 - F2;
    R[k] = r0; R[nn + k] = r1; R[2 * (int64_t)nn + k] = r2;
    if (Tint) { Tint[k] = T0; Tint[nn + k] = T1; Tint[2 * (int64_t)nn + k] = T2; }
    if (Fext) { Fext[k] = F0; Fext[nn + k] = F1; Fext[2 * (int64_t)nn + k] = F2; }
    if (check_nan && (isnan(r0) || isnan(r1) || isnan(r2))) atomicOr(status, 1);
}

// Fused assembly: one 512-thread workgroup per patch. Per batch of `epb` elements: zero the
// window slots of rows that start here, gather nodal values, compute the integration-point
// physics, compute the element rows as (element, node, block row) items, add the rows of
// owned nodes into their window slots in rank order (one rank per barrier phase, so no two
// threads add to the same slot at once and every sum has a fixed order), then write the rows
// whose last element was in this batch to K, R, Tint and Fext.
constexpr int kFusedThreads = 512;
constexpr int kFusedWaves = kFusedThreads / 64;

template <int NOD, bool RONLY>
__global__ __launch_bounds__(kFusedThreads) void k_fused(const MeshArgs *__restrict__ MA,
                                                         const CallArgs *__restrict__ CA, int nip, int epb,
                                                         int work_doubles, int64_t nnzN, int nn, int check_nan) {
    extern __shared__ double lds[];
    const int tstride = nip * NOD;
    double *sN = lds;
    double *sXi = sN + tstride;
    double *sEta = sXi + tstride;
    double *sW = sEta + tstride;
    double *work = lds + ((3 * tstride + nip + 1) & ~1);
    double *win = work + work_doubles;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const ShapeTables *tab = MA->tab;
        for (int k = tid; k < tstride; k += kFusedThreads) {
            const int q = k / NOD, i = k - q * NOD;
            sN[k] = tab->N[q][i]; sXi[k] = tab->dxi[q][i]; sEta[k] = tab->deta[q][i];
        }
        if (tid < nip) sW[tid] = tab->w[tid];
    }
    int bad = 0;
    const int nrow = epb * NOD;  // (element, node) pairs per batch
    const int b0 = MA->s.patch_batch_ptr[blockIdx.x], b1 = MA->s.patch_batch_ptr[blockIdx.x + 1];
    for (int b = b0; b < b1; ++b) {
        // rows that start in this batch: zero their slots (one wave per row)
        {
            const int z0 = MA->s.zero_ptr[b], z1 = MA->s.zero_ptr[b + 1];
            for (int z = z0 + wave; z < z1; z += kFusedWaves) {
                double *w = win + MA->s.zero_slot[z];
                const int len = MA->s.zero_len[z];
                for (int k = lane; k < len; k += 64) w[k] = 0.0;
            }
        }
        // gathers
        const int32_t *belem = MA->s.batch_elem + (int64_t)b * epb;
        if (tid < nrow) {
            const int el = tid / NOD, a = tid - el * NOD;
            const int e = belem[el];
            if (e >= 0)
                gather_node<NOD>(MA->conn, MA->x, MA->y, CA->f, MA->b, CA->p, e, a, work + el * (NV_COUNT * NOD));
        }
        __syncthreads();
        // physics
        double c[kNC];
        const int pel = tid / nip, pq = tid - pel * nip;
        const bool phys = tid < epb * nip && belem[pel < epb ? pel : 0] >= 0;
        if (phys) {
            const Params p = CA->p;
            ip_physics<NOD>(work + pel * (NV_COUNT * NOD), sN + pq * NOD, sXi + pq * NOD, sEta + pq * NOD, sW[pq],
                            p, c);
        }
        __syncthreads();
        if (phys) {
            double *dst = work + (pel * nip + pq) * kNC;
#pragma unroll
            for (int k = 0; k < kNC; k += 2) {
                double2 t;
                t.x = c[k];
                t.y = (k + 1 < C_COUNT) ? c[k + 1] : 0.0;
                *reinterpret_cast<double2 *>(dst + k) = t;
            }
        }
        __syncthreads();
        // element rows of owned nodes: item = bi * nrow + (el * NOD + i), block row bi
        const int bi = tid / nrow;
        const int rr = tid - bi * nrow;
        const int el = rr / NOD, i = rr - el * NOD;
        int slot = -1, rank = 0, e = -1, deg = 0;
        if (bi < 3) {
            e = belem[el];
            const int64_t ea = ((int64_t)b * epb + el) * NOD + i;
            slot = e >= 0 ? MA->s.ea_slot[ea] : -1;
            rank = MA->s.ea_rank[ea];
        }
        double T = 0, F = 0, Kb[3][NOD];
        if (slot >= 0) {
            const double *crr = work + el * nip * kNC;
            if (bi == 0) element_row_b<NOD, 0>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
            else if (bi == 1) element_row_b<NOD, 1>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
            else element_row_b<NOD, 2>(i, nip, crr, sN, sXi, sEta, T, F, Kb);
            const int node = MA->conn[(int64_t)e * NOD + i];
            deg = (int)(MA->rowptr_n[node + 1] - MA->rowptr_n[node]);
        }
        const int maxrank = MA->s.batch_maxrank[b];
        const uint8_t *ps = MA->pos + ((int64_t)(e < 0 ? 0 : e) * NOD + i) * NOD;
        for (int r = 0; r <= maxrank; ++r) {
            if (slot >= 0 && rank == r) {
                double *w = win + slot;
                if (!RONLY) {
                    double *wb = w + bi * 3 * deg;
#pragma unroll
                    for (int j = 0; j < NOD; ++j) {
                        const int cj = ps[j];
#pragma unroll
                        for (int bj = 0; bj < 3; ++bj) lds_add(wb + bj * deg + cj, Kb[bj][j]);
                    }
                }
                lds_add(w + 9 * deg + bi, T);       // Tx Ty Th
                lds_add(w + 9 * deg + 3 + bi, F);   // Fx Fy Fh
            }
            __syncthreads();
        }
        // rows whose last element was in this batch: write out (one wave per row)
        {
            const int f0 = MA->s.flush_ptr[b], f1 = MA->s.flush_ptr[b + 1];
            double *Kval = CA->Kval;
            for (int fl = f0 + wave; fl < f1; fl += kFusedWaves) {
                const int node = MA->s.flush_node[fl];
                const double *w = win + MA->s.flush_slot[fl];
                const int64_t rp = MA->rowptr_n[node];
                const int dg = (int)(MA->rowptr_n[node + 1] - rp);
                if (!RONLY && Kval) {
#pragma unroll
                    for (int bk = 0; bk < 3; ++bk) {
                        double *dst = Kval + (int64_t)bk * 3 * nnzN + 3 * rp;
                        const double *src = w + bk * 3 * dg;
                        for (int k = lane; k < 3 * dg; k += 64) dst[k] = src[k];
                    }
                    // diagonal entry of block (lane, lane) (uvhMatrixAssembly.m:575-579)
                    if (check_nan && lane < 3 && isnan(w[lane * 3 * dg + lane * dg + MA->diag[node]])) bad |= 2;
                }
                if (lane < 3) {
                    const double Tv = w[9 * dg + lane], Fv = w[9 * dg + 3 + lane];
                    const double rv = Tv - Fv;
                    CA->R[(int64_t)lane * nn + node] = rv;
                    if (CA->Tint) CA->Tint[(int64_t)lane * nn + node] = Tv;
                    if (CA->Fext) CA->Fext[(int64_t)lane * nn + node] = Fv;
                    if (check_nan && isnan(rv)) bad |= 1;
                }
            }
        }
        __syncthreads();
    }
    if (bad) atomicOr(MA->status, bad);
}

// ------------------------------------------------------------------ launchers

namespace {
constexpr size_t kLdsLimit = 64 * 1024;

// Elements per 256-thread block: every stage maps at most one item per thread, and the
// LDS (shape tables + the larger of nodal staging and coefficient records) stays <= 64 KiB.
int element_epb(int nod, int nip) {
    const size_t tables = size_t((3 * nip * nod + nip + 1) & ~1) * sizeof(double);
    const size_t per_el = std::max(size_t(NV_COUNT) * nod, size_t(nip) * kNC) * sizeof(double);
    int epb = std::min(256 / nip, 256 / nod);
    epb = std::min<int>(epb, int((kLdsLimit - tables) / per_el));
    return epb;
}

size_t element_lds_bytes(int nod, int nip, int epb) {
    const size_t tables = size_t((3 * nip * nod + nip + 1) & ~1) * sizeof(double);
    const size_t per_el = std::max(size_t(NV_COUNT) * nod, size_t(nip) * kNC) * sizeof(double);
    return tables + per_el * epb;
}

template <int NOD>
hipError_t launch_element_nod(int nele, int nip, const int32_t *conn, const double *x, const double *y,
                              const FieldPtrs &f, const double *b, const ShapeTables *tab, const Params &p,
                              const int32_t *inc_pos, bool ronly, double *EB, double *RE, hipStream_t s) {
    const int epb = element_epb(NOD, nip);
    if (epb < 1) return hipErrorInvalidConfiguration;
    const int nblocks = (nele + epb - 1) / epb;
    const int grid = ((nblocks + kXcds - 1) / kXcds) * kXcds;  // whole rounds over the XCDs
    const size_t lds = element_lds_bytes(NOD, nip, epb);
    if (ronly)
        hipLaunchKernelGGL((k_element<NOD, true>), dim3(grid), dim3(256), lds, s, nele, nip, epb, conn, x, y, f, b, tab, p, inc_pos, EB, RE);
    else
        hipLaunchKernelGGL((k_element<NOD, false>), dim3(grid), dim3(256), lds, s, nele, nip, epb, conn, x, y, f, b, tab, p, inc_pos, EB, RE);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_nodal_b(int nn, const FieldPtrs &f, const Params &p, double *b, hipStream_t s) {
    hipLaunchKernelGGL(k_nodal_b, dim3((nn + 255) / 256), dim3(256), 0, s, nn, f, p, b);
    return hipGetLastError();
}

hipError_t launch_clear_status(int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, status);
    return hipGetLastError();
}

hipError_t launch_element(int nod, int nele, int nip, const int32_t *conn, const double *x, const double *y,
                          const FieldPtrs &f, const double *b, const ShapeTables *tab, const Params &p,
                          const int32_t *inc_pos, bool ronly, double *EB, double *RE, hipStream_t s) {
    switch (nod) {
        case 3: return launch_element_nod<3>(nele, nip, conn, x, y, f, b, tab, p, inc_pos, ronly, EB, RE, s);
        case 6: return launch_element_nod<6>(nele, nip, conn, x, y, f, b, tab, p, inc_pos, ronly, EB, RE, s);
        case 10: return launch_element_nod<10>(nele, nip, conn, x, y, f, b, tab, p, inc_pos, ronly, EB, RE, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows(int nod, int nn, int maxdeg, const int32_t *inc_ptr, const int32_t *inc,
                       const uint8_t *pos, const int64_t *rowptr_n, const uint8_t *diag, int64_t nnzN,
                       const double *EB, const double *RE, double *R, double *Tint, double *Fext, double *Kval,
                       int check_nan, int *status, hipStream_t s) {
    if (!Kval) {
        hipLaunchKernelGGL(k_rows_r, dim3((nn + 255) / 256), dim3(256), 0, s, nn, inc_ptr, inc, RE, R, Tint, Fext,
                           check_nan, status);
        return hipGetLastError();
    }
    const int nblocks = (nn + 3) / 4;
    const dim3 grid(((nblocks + kXcds - 1) / kXcds) * kXcds), block(256);
    const size_t lds = size_t(4) * 9 * maxdeg * sizeof(double);
    switch (nod) {
        case 3: hipLaunchKernelGGL(k_rows<3>, grid, block, lds, s, nn, maxdeg, inc_ptr, inc, pos, rowptr_n, diag, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        case 6: hipLaunchKernelGGL(k_rows<6>, grid, block, lds, s, nn, maxdeg, inc_ptr, inc, pos, rowptr_n, diag, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        case 10: hipLaunchKernelGGL(k_rows<10>, grid, block, lds, s, nn, maxdeg, inc_ptr, inc, pos, rowptr_n, diag, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// LDS layout of the fused kernel: shape tables, per-batch work area, row window.
FusedLayout fused_layout(int nod, int nip) {
    FusedLayout L{};
    L.table_doubles = (3 * nip * nod + nip + 1) & ~1;
    const int per_el_bytes = std::max(NV_COUNT * nod, nip * kNC) * int(sizeof(double));
    L.epb = std::min(kFusedThreads / nip, kFusedThreads / (3 * nod));
    L.epb = std::min<int>(L.epb, int((kLdsLimit - L.table_doubles * sizeof(double)) / per_el_bytes));
    const int per_el = std::max(NV_COUNT * nod, nip * kNC);
    L.work_doubles = (L.epb * per_el + 1) & ~1;
    const int64_t total = kFusedLdsBytes / int64_t(sizeof(double));
    L.window_doubles = ((total - L.table_doubles - L.work_doubles) / 2) * 2;
    return L;
}

hipError_t launch_fused(int nod, int nip, int npatch, const FusedLayout &L, const MeshArgs *MA, const CallArgs *CA,
                        int64_t nnzN, int nn, bool ronly, int check_nan, hipStream_t s) {
    if (npatch <= 0) return hipSuccess;
    const size_t lds = size_t(L.table_doubles + L.work_doubles + L.window_doubles) * sizeof(double);
    const dim3 grid(npatch), block(kFusedThreads);
#define UVH_FUSED(NODV, RO)                                                                                     \
    do {                                                                                                        \
        auto kern = k_fused<NODV, RO>;                                                                          \
        hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,  \
                            int(lds));                                                                          \
        hipLaunchKernelGGL(kern, grid, block, lds, s, MA, CA, nip, L.epb, L.work_doubles, nnzN, nn, check_nan);  \
    } while (0)
    switch (nod) {
        case 3: if (ronly) UVH_FUSED(3, true); else UVH_FUSED(3, false); break;
        case 6: if (ronly) UVH_FUSED(6, true); else UVH_FUSED(6, false); break;
        case 10: if (ronly) UVH_FUSED(10, true); else UVH_FUSED(10, false); break;
        default: return hipErrorInvalidValue;
    }
#undef UVH_FUSED
    return hipGetLastError();
}
}  // namespace uvh
