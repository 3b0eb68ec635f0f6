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
// Lane mapping of k_element: a group of G lanes (G = 4, 8, 16 or 32, a power of two >=
// max(nod, nip)) owns one element. Lanes q < nip compute the physics of integration point q
// and publish a record of per-point coefficients in LDS; lanes i < nod then accumulate row i
// of the element (6 residual terms, 9 x nod Jacobian terms) over all points in registers.
#include <hip/hip_runtime.h>

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
#pragma unroll
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

    // floatation (uvhAssemblyIntPointImplicitSUPG.m:209-243; HeavisideApprox.m, DiracDelta.m)
    const double H = Sq - Bq;
    const double hf = p.rhow * H / rho;
    const double x1 = p.kH * ((h - hf) - p.Hh0);
    const double x2 = p.kH * ((hf - h) - p.Hh0);
    const double He = 0.5 + 0.5 * tanh(x1);
    const double HE = 0.5 + 0.5 * tanh(x2);
    const double s1c = 1.0 / cosh(x1), s2c = 1.0 / cosh(x2);
    const double delta = 0.5 * p.kH * (s1c * s1c);
    const double Delta = 0.5 * p.kH * (s2c * s2c);
    const double Hpos = (0.5 + 0.5 * tanh(p.kH * (H - p.Hh0))) * H;
    const double d = HE * rho * h / p.rhow + He * Hpos;
    const double dd = HE * rho / p.rhow - Delta * h * rho / p.rhow + delta * Hpos;

    // Glen's flow law (EffectiveViscositySSTREAM.m:7-9)
    const double e = sqrt(p.eps0 * p.eps0 + exx * exx + eyy * eyy + exx * eyy + exy * exy);
    const double Am = pow(A, -1.0 / n);
    const double eta = 0.5 * Am * pow(e, (1.0 - n) / n) + p.eta0;
    const double E = (1.0 - n) / (4.0 * n) * Am * pow(e, (1.0 - 3.0 * n) / n);

    // Weertman sliding (BasalDrag.m:54-77,410-443)
    const double Hec = He < 0 ? 0.0 : He, dlc = delta < 0 ? 0.0 : delta;
    const double U2 = u * u + v * v + p.uz * p.uz;
    const double Cm = pow(C + p.czero, -1.0 / m);
    const double beta2 = Cm * pow(sqrt(U2), 1.0 / m - 1.0);
    const double Db = (1.0 / m - 1.0) * Cm * pow(U2, (1.0 - 3.0 * m) / (2.0 * m));
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

template <int NOD, int G, int BLOCK, bool RONLY>
__global__ __launch_bounds__(BLOCK) void k_element(int nele, int nip, const int32_t *__restrict__ conn,
                                                 const double *__restrict__ xn, const double *__restrict__ yn,
                                                 FieldPtrs f, const double *__restrict__ bnod,
                                                 const ShapeTables *__restrict__ tab, Params p,
                                                 double *__restrict__ EB, double *__restrict__ RE) {
    constexpr int EPB = BLOCK / G;
    extern __shared__ double lds[];
    const int tstride = nip * NOD;
    double *sN = lds;                       // [nip][NOD]
    double *sXi = sN + tstride;
    double *sEta = sXi + tstride;
    double *sW = sEta + tstride;            // [nip]
    double *work = lds + ((3 * tstride + nip + 1) & ~1);  // 16-byte aligned union:
                                            // nodal [EPB][NV_COUNT][NOD] / coef [EPB][nip][kNC]

    const int tid = threadIdx.x;
    for (int k = tid; k < nip * NOD; k += BLOCK) {
        const int q = k / NOD, i = k % NOD;
        sN[k] = tab->N[q][i]; sXi[k] = tab->dxi[q][i]; sEta[k] = tab->deta[q][i];
    }
    if (tid < nip) sW[tid] = tab->w[tid];

    const int el = tid / G, lane = tid % G;
    const int e = blockIdx.x * EPB + el;
    const bool valid = e < nele;

    // stage 1: gather nodal values (lane i loads node i of its element)
    double *nv = work + el * (NV_COUNT * NOD);
    if (valid && lane < NOD) {
        const int k = conn[(int64_t)e * NOD + lane];
        double h = f.h[k], u = f.u[k], v = f.v[k], u0 = f.u0[k], v0 = f.v0[k];
        double as1 = f.as1[k], ab1 = f.ab1[k], as0 = f.as0[k], ab0 = f.ab0[k];
        const double h0 = f.h0[k];
        if (p.zero_fields) {  // uvhMatrixAssembly.m:73-87
            u = 0.0; v = 0.0; u0 = 0.0; v0 = 0.0; h = h0;
            as1 = fabs(ab1) + 1.0; ab1 = fabs(ab1);
            as0 = fabs(ab0) + 1.0; ab0 = fabs(ab0);
        }
        nv[NV_H * NOD + lane] = h; nv[NV_U * NOD + lane] = u; nv[NV_V * NOD + lane] = v;
        nv[NV_S * NOD + lane] = f.S[k]; nv[NV_B * NOD + lane] = f.B[k]; nv[NV_RHO * NOD + lane] = f.rho[k];
        nv[NV_C * NOD + lane] = f.C[k]; nv[NV_M * NOD + lane] = f.m[k]; nv[NV_N * NOD + lane] = f.n[k];
        nv[NV_A * NOD + lane] = f.A[k]; nv[NV_AS1 * NOD + lane] = as1; nv[NV_AB1 * NOD + lane] = ab1;
        nv[NV_H0 * NOD + lane] = h0; nv[NV_U0 * NOD + lane] = u0; nv[NV_V0 * NOD + lane] = v0;
        nv[NV_AS0 * NOD + lane] = as0; nv[NV_AB0 * NOD + lane] = ab0;
        nv[NV_DADH * NOD + lane] = f.dadh ? f.dadh[k] : 0.0;
        nv[NV_BL * NOD + lane] = bnod[k];
        nv[NV_X * NOD + lane] = xn[k]; nv[NV_Y * NOD + lane] = yn[k];
    }
    __syncthreads();

    // stage 2: physics of integration point q = lane
    double c[kNC];
    const bool phys = valid && lane < nip;
    if (phys) ip_physics<NOD>(nv, sN + lane * NOD, sXi + lane * NOD, sEta + lane * NOD, sW[lane], p, c);
    __syncthreads();  // nodal staging is dead from here; the coefficient records reuse it
    double *cr = work + el * (nip * kNC);
    if (phys) {
#pragma unroll
        for (int k = 0; k < kNC; k += 2) {
            double2 t;
            t.x = c[k];
            t.y = (k + 1 < C_COUNT) ? c[k + 1] : 0.0;
            *reinterpret_cast<double2 *>(cr + lane * kNC + k) = t;
        }
    }
    __syncthreads();

    // stage 3: row lane i accumulates element row i over the integration points
    if (!valid || lane >= NOD) return;
    const int i = lane;
    double Tx = 0, Fx = 0, Ty = 0, Fy = 0, Th = 0, Fh = 0;
    double K[RONLY ? 1 : 9][NOD];
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
            const double AX = r[C_AX], AY = r[C_AY], RH = r[C_RH], GAM = r[C_GAM], RU = r[C_RU], RV = r[C_RV];
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
                K[6][j] += Wpsi * (AX * Nj + RH * Nxj);                // Khu
                K[7][j] += Wpsi * (AY * Nj + RH * Nyj);                // Khv
                K[8][j] += Wpsi * (GAM * Nj + RU * Nxj + RV * Nyj);    // Khh
            }
        }
    }
    const int64_t row = (int64_t)e * NOD + i;
    double *re = RE + row * 6;
    re[0] = Tx; re[1] = Fx; re[2] = Ty; re[3] = Fy; re[4] = Th; re[5] = Fh;
    if (!RONLY) {
        double *eb = EB + row * (9 * NOD);
#pragma unroll
        for (int b = 0; b < 9; ++b)
#pragma unroll
            for (int j = 0; j < NOD; ++j) eb[b * NOD + j] = K[b][j];
    }
}

// One thread per node: R, Tint, Fext and the three CSR rows (x, y, h) of K.
template <int NOD>
__global__ __launch_bounds__(256) void k_rows(int nn, const int32_t *__restrict__ inc_ptr,
                                              const int32_t *__restrict__ inc, const uint8_t *__restrict__ pos,
                                              const int64_t *__restrict__ rowptr_n, int64_t nnzN,
                                              const double *__restrict__ EB, const double *__restrict__ RE,
                                              double *__restrict__ R, double *__restrict__ Tint,
                                              double *__restrict__ Fext, double *__restrict__ Kval, int check_nan,
                                              int *__restrict__ status) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int t0 = inc_ptr[k], t1 = inc_ptr[k + 1];
    double T0 = 0, T1 = 0, T2 = 0, F0 = 0, F1 = 0, F2 = 0;
    for (int t = t0; t < t1; ++t) {
        const double *re = RE + (int64_t)inc[t] * 6;
        T0 += re[0]; F0 += re[1]; T1 += re[2]; F1 += re[3]; T2 += re[4]; F2 += re[5];
    }
    const double r0 = T0 - F0, r1 = T1 - F1, r2 = T2 - F2;
    R[k] = r0; R[nn + k] = r1; R[2 * (int64_t)nn + k] = r2;
    if (Tint) { Tint[k] = T0; Tint[nn + k] = T1; Tint[2 * (int64_t)nn + k] = T2; }
    if (Fext) { Fext[k] = F0; Fext[nn + k] = F1; Fext[2 * (int64_t)nn + k] = F2; }
    int bad = 0;
    if (check_nan && (isnan(r0) || isnan(r1) || isnan(r2))) bad |= 1;
    if (Kval) {
        const int64_t rp = rowptr_n[k];
        const int deg = (int)(rowptr_n[k + 1] - rp);
        double *rows[3];
#pragma unroll
        for (int bi = 0; bi < 3; ++bi) {
            rows[bi] = Kval + (int64_t)bi * 3 * nnzN + 3 * rp;
            for (int c = 0; c < 3 * deg; ++c) rows[bi][c] = 0.0;
        }
        int diag = 0;
        for (int t = t0; t < t1; ++t) {
            const int ea = inc[t];
            const uint8_t *ps = pos + (int64_t)ea * NOD;
            const double *eb = EB + (int64_t)ea * (9 * NOD);
            if (t == t0) diag = ps[ea % NOD];
#pragma unroll
            for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                for (int bj = 0; bj < 3; ++bj)
#pragma unroll
                    for (int j = 0; j < NOD; ++j) rows[bi][bj * deg + ps[j]] += eb[(bi * 3 + bj) * NOD + j];
        }
        if (check_nan) {
#pragma unroll
            for (int bi = 0; bi < 3; ++bi)
                if (isnan(rows[bi][bi * deg + diag])) bad |= 2;
        }
    }
    if (bad) atomicOr(status, bad);
}

// ------------------------------------------------------------------ launchers

namespace {
template <int NOD, int G>
size_t element_lds_bytes(int nip, int block) {
    const int EPB = block / G;
    const size_t tables = size_t((3 * nip * NOD + nip + 1) & ~1) * sizeof(double);
    const size_t nodal = size_t(EPB) * NV_COUNT * NOD * sizeof(double);
    const size_t coef = size_t(EPB) * nip * kNC * sizeof(double);
    return tables + (nodal > coef ? nodal : coef);
}

constexpr size_t kLdsLimit = 64 * 1024;

template <int NOD, int G>
hipError_t launch_element_g(int nele, int nip, const int32_t *conn, const double *x, const double *y,
                            const FieldPtrs &f, const double *b, const ShapeTables *tab, const Params &p,
                            bool ronly, double *EB, double *RE, hipStream_t s) {
    if (element_lds_bytes<NOD, G>(nip, 256) <= kLdsLimit) {
        constexpr int EPB = 256 / G;
        const int grid = (nele + EPB - 1) / EPB;
        const size_t lds = element_lds_bytes<NOD, G>(nip, 256);
        if (ronly)
            hipLaunchKernelGGL((k_element<NOD, G, 256, true>), dim3(grid), dim3(256), lds, s, nele, nip, conn, x, y, f, b, tab, p, EB, RE);
        else
            hipLaunchKernelGGL((k_element<NOD, G, 256, false>), dim3(grid), dim3(256), lds, s, nele, nip, conn, x, y, f, b, tab, p, EB, RE);
    } else {
        constexpr int EPB = 128 / G;
        const int grid = (nele + EPB - 1) / EPB;
        const size_t lds = element_lds_bytes<NOD, G>(nip, 128);
        if (lds > kLdsLimit) return hipErrorInvalidConfiguration;
        if (ronly)
            hipLaunchKernelGGL((k_element<NOD, G, 128, true>), dim3(grid), dim3(128), lds, s, nele, nip, conn, x, y, f, b, tab, p, EB, RE);
        else
            hipLaunchKernelGGL((k_element<NOD, G, 128, false>), dim3(grid), dim3(128), lds, s, nele, nip, conn, x, y, f, b, tab, p, EB, RE);
    }
    return hipGetLastError();
}

template <int NOD>
hipError_t launch_element_nod(int nele, int nip, const int32_t *conn, const double *x, const double *y,
                              const FieldPtrs &f, const double *b, const ShapeTables *tab, const Params &p,
                              bool ronly, double *EB, double *RE, hipStream_t s) {
    const int need = nip > NOD ? nip : NOD;
    if (need <= 4) return launch_element_g<NOD, (NOD <= 4 ? 4 : 32)>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
    if (need <= 8) return launch_element_g<NOD, (NOD <= 8 ? 8 : 32)>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
    if (need <= 16) return launch_element_g<NOD, 16>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
    return launch_element_g<NOD, 32>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
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
                          const FieldPtrs &f, const double *b, const ShapeTables *tab, const Params &p, bool ronly,
                          double *EB, double *RE, hipStream_t s) {
    switch (nod) {
        case 3: return launch_element_nod<3>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
        case 6: return launch_element_nod<6>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
        case 10: return launch_element_nod<10>(nele, nip, conn, x, y, f, b, tab, p, ronly, EB, RE, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows(int nod, int nn, const int32_t *inc_ptr, const int32_t *inc, const uint8_t *pos,
                       const int64_t *rowptr_n, int64_t nnzN, const double *EB, const double *RE, double *R,
                       double *Tint, double *Fext, double *Kval, int check_nan, int *status, hipStream_t s) {
    const dim3 grid((nn + 255) / 256), block(256);
    switch (nod) {
        case 3: hipLaunchKernelGGL(k_rows<3>, grid, block, 0, s, nn, inc_ptr, inc, pos, rowptr_n, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        case 6: hipLaunchKernelGGL(k_rows<6>, grid, block, 0, s, nn, inc_ptr, inc, pos, rowptr_n, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        case 10: hipLaunchKernelGGL(k_rows<10>, grid, block, 0, s, nn, inc_ptr, inc, pos, rowptr_n, nnzN, EB, RE, R, Tint, Fext, Kval, check_nan, status); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace uvh
