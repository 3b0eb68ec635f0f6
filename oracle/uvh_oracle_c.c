/*
 * Oracle O2 -- TEST INFRASTRUCTURE ONLY.
 *
 * An independent FP64 per-element C restatement of Úa's uvh assembly, used
 * (a) to cross-check oracle O1 (oracle/uvh_oracle.py, which mirrors the MATLAB
 * vectorisation), (b) as the checker for the HIP kernels at sizes where O1 is
 * slow, and (c) as the timed CPU baseline ("port": a restatement of Úa, not
 * MATLAB). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it; the product path never does.
 *
 * Where O1 mirrors MATLAB's array expressions, O2 is written per element and per
 * integration point, with shape functions generated from barycentric Lagrange
 * forms rather than the expanded polynomials of shape_fun.m/shape_der.m, so a
 * slip in one restatement shows up as an O1/O2 mismatch.
 *
 * Restated reference lines:
 *   zero-fields transform        UaMain/uvhMatrixAssembly.m:48-101
 *   floatation pre-pass (b)      UaMain/Calc_bs_From_hBS.m:71-79,109
 *   shape functions, Jacobian    UaMain/shape_fun.m:52-82, shape_der.m:54-120, derivVector.m:96-107
 *   integration-point physics    UaMain/uvhAssemblyIntPointImplicitSUPG.m:62-126,206-308,323-395
 *   viscosity                    UaMain/EffectiveViscositySSTREAM.m:7-9
 *   Weertman drag                UaMain/BasalDrag.m:54-77,410-443
 *   SUPG tau                     UaMain/SUPGtau.m:24-25,45-77
 *   element R and K blocks       UaMain/uvhNodalLoopSSTREAM.m:39-228 (2025 form)
 *   scatter to R and CSR K       UaMain/uvhMatrixAssembly.m:416-509 (duplicates summed)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double dt, theta, beta0, tau_multiplier, speed_zero, eps_zero, eta_zero, czero, kH, Hh0;
    double g, rhow, alpha;
    int tau_option;   /* 0 taus, 1 taut, 2 tau1, 3 tau2, 4 tau0 */
    int ronly;
    int zero_fields;
    int reset_thickness; /* ResetThicknessToMinThickness && ResetThicknessInNonLinLoop */
    double thick_min;
    int law;             /* CtrlVar.SlidingLaw: 0 Weertman, 1 Budd, 2 Coulomb, 3 Tsai, 4 Cornford,
                            5 Umbi, 6 Joughin */
    double nzero;        /* CtrlVar.Nzero */
    int enforce_b_above; /* CtrlVar.Enforce_bAboveB */
} o2_params;

/* field order in the `f1` array of pointers: h ub vb S B rho C m n AGlen as ab, then the optional
 * q muk V0 and dadh (null when absent); f0: h ub vb as ab */
enum { F1_H, F1_U, F1_V, F1_S, F1_B, F1_RHO, F1_C, F1_M, F1_N, F1_A, F1_AS, F1_AB, F1_Q, F1_MUK, F1_V0, F1_DADH, NF1 };
enum { F0_H, F0_U, F0_V, F0_AS, F0_AB, NF0 };

#define MAXNOD 10

/* Barycentric Lagrange shape functions on Úa's local numbering.
 * l1 = xi (corner 1), l2 = eta (corner "top"), l3 = 1 - xi - eta.
 * 3-node:  N = [l1, l3, l2]
 * 6-node:  corners 1,3,5 at l1,l3,l2; edge nodes 2 (l1-l3), 4 (l3-l2), 6 (l2-l1)
 * 10-node: corners 1,4,7 at l1,l3,l2; edge nodes 2,3 (l1->l3), 5,6 (l3->l2), 8,9 (l2->l1),
 *          node 10 at the centroid.
 * dN/dl (three partials) is returned and chained to (xi,eta) below. */
static void shape_barycentric(int nod, double xi, double eta, double *N, double *Nxi, double *Neta)
{
    double L[3] = {xi, eta, 1.0 - xi - eta};
    /* d l_k / d xi and d l_k / d eta */
    const double dLxi[3] = {1.0, 0.0, -1.0};
    const double dLeta[3] = {0.0, 1.0, -1.0};
    double dNdL[MAXNOD][3];
    memset(dNdL, 0, sizeof(dNdL));
    if (nod == 3) {
        int cidx[3] = {0, 2, 1};  /* node k is corner L[cidx[k]] */
        for (int k = 0; k < 3; ++k) { N[k] = L[cidx[k]]; dNdL[k][cidx[k]] = 1.0; }
    } else if (nod == 6) {
        int corner_node[3] = {0, 4, 2};   /* node index of the corner at L[0], L[1], L[2] */
        for (int c = 0; c < 3; ++c) {
            double l = L[c];
            N[corner_node[c]] = (2.0 * l - 1.0) * l;
            dNdL[corner_node[c]][c] = 4.0 * l - 1.0;
        }
        /* edge node -> (corner a, corner b) as indices into L */
        int ea[3] = {0, 2, 1}, eb[3] = {2, 1, 0}, en[3] = {1, 3, 5};
        for (int k = 0; k < 3; ++k) {
            N[en[k]] = 4.0 * L[ea[k]] * L[eb[k]];
            dNdL[en[k]][ea[k]] = 4.0 * L[eb[k]];
            dNdL[en[k]][eb[k]] = 4.0 * L[ea[k]];
        }
    } else { /* 10 */
        int corner_node[3] = {0, 6, 3};
        for (int c = 0; c < 3; ++c) {
            double l = L[c];
            N[corner_node[c]] = 0.5 * (3.0 * l - 1.0) * (3.0 * l - 2.0) * l;
            dNdL[corner_node[c]][c] = 0.5 * (27.0 * l * l - 18.0 * l + 2.0);
        }
        /* edge node k sits at 2/3 of corner near[k] and 1/3 of corner far[k]:
           N = 9/2 * l_near * l_far * (3 l_near - 1) */
        int nodes[6] = {1, 2, 4, 5, 7, 8};
        int near_[6] = {0, 2, 2, 1, 1, 0};
        int far_[6]  = {2, 0, 1, 2, 0, 1};
        for (int k = 0; k < 6; ++k) {
            double a = L[near_[k]], b = L[far_[k]];
            N[nodes[k]] = 4.5 * a * b * (3.0 * a - 1.0);
            dNdL[nodes[k]][near_[k]] = 4.5 * b * (6.0 * a - 1.0);
            dNdL[nodes[k]][far_[k]] = 4.5 * a * (3.0 * a - 1.0);
        }
        N[9] = 27.0 * L[0] * L[1] * L[2];
        dNdL[9][0] = 27.0 * L[1] * L[2];
        dNdL[9][1] = 27.0 * L[0] * L[2];
        dNdL[9][2] = 27.0 * L[0] * L[1];
    }
    for (int k = 0; k < nod; ++k) {
        Nxi[k] = dNdL[k][0] * dLxi[0] + dNdL[k][1] * dLxi[1] + dNdL[k][2] * dLxi[2];
        Neta[k] = dNdL[k][0] * dLeta[0] + dNdL[k][1] * dLeta[1] + dNdL[k][2] * dLeta[2];
    }
}

void o2_shape(int nod, double xi, double eta, double *N, double *Nxi, double *Neta)
{
    shape_barycentric(nod, xi, eta, N, Nxi, Neta);
}

static inline double soft_step(double k, double x, double x0) { return 0.5 + 0.5 * tanh(k * (x - x0)); }
static inline double soft_delta(double k, double x, double x0)
{
    double c = cosh(k * (x - x0));
    return 0.5 * k / (c * c);
}

/* Floatation pre-pass: nodal lower surface b from h, S, B (Calc_bs_From_hBS). */
static void nodal_b(const o2_params *P, int64_t nn, const double *h, const double *S, const double *B,
                    const double *rho, double *b)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nn; ++k) {
        double hk = h[k];
        if (P->reset_thickness && hk < P->thick_min) hk = P->thick_min;
        double hf = P->rhow * (S[k] - B[k]) / rho[k];
        double GF = soft_step(P->kH, hk - hf, P->Hh0);
        double bfloat = S[k] - rho[k] * hk / P->rhow;
        b[k] = GF * B[k] + (1.0 - GF) * bfloat;
        /* UaMain/Calc_bs_From_hBS.m:82-93: no lower surface below the bed */
        if (P->enforce_b_above && b[k] < B[k]) b[k] = B[k];
    }
}

/* Sliding laws other than Weertman, UaMain/BasalDrag.m:182-257: Budd (:448-498), Coulomb
 * (:365-408), Tsai (the Coulomb values where |tau_C|^2 < |tau_W|^2), Cornford (rpCWN0.m), Umbi
 * (rCWN0.m) and Joughin (rCWV0.m), with the effective pressure N0 (:500-521). Inputs at the
 * point: h, H = S - B, rho, u, v, m, C + Czero, q, muk, V0, and the Weertman quantities (U2,
 * beta2, Dbeta2; He and delta clamped at zero). The outputs hold the Weertman values on entry
 * and are replaced where the law says so. The last three are written as the reference's
 * generated expressions evaluate: tau = f(U) (u, v) with U = |(u, v, u0)|; their delta terms
 * take He = 1/2 and Umbi's dN/dh term carries He once. */
static void other_drag(const o2_params *P, double h, double H, double rho, double u, double v, double m, double Cp,
                       double q, double muk, double V0, double U2, double beta2, double Dbeta2, double He,
                       double delta, double *taux, double *tauy, double *dtxdu, double *dtydv, double *dtxdv,
                       double *dtxdh, double *dtydh)
{
    double hf = P->rhow * H / rho;
    if (hf < DBL_EPSILON) hf = 0.0;
    double Dh = h - hf;
    const int I = Dh < DBL_EPSILON;
    if (I) Dh = 0.0;
    const double N = rho * P->g * Dh + P->nzero;
    const double dNdh = I ? 0.0 : rho * P->g;
    if (P->law == 1) {
        const double qm = q / m, Nqm = pow(N, qm), T = He * Nqm * beta2;
        *taux = T * u; *tauy = T * v;
        *dtxdu = He * Nqm * (beta2 + Dbeta2 * u * u);
        *dtydv = He * Nqm * (beta2 + Dbeta2 * v * v);
        *dtxdv = He * Nqm * Dbeta2 * u * v;
        const double E = delta * T + He * qm * pow(N, qm - 1.0) * dNdh * beta2;
        *dtxdh = E * u; *dtydh = E * v;
        return;
    }
    const double speed = sqrt(U2);
    if (P->law >= 4) {
        double f, fU;   /* tau = f (u, v); fU = df/dU */
        double dTdh;     /* d|tau|/dh (reference form) */
        if (P->law == 4) {
            /* |tau| = tauC tauW / (tauW^m + tauC^m)^(1/m), tauC = muk N, tauW = He C'^(-1/m) U^(1/m) */
            const double tC = muk * N, w = pow(Cp, -1.0 / m) * pow(speed, 1.0 / m);
            const double D = pow(tC, m) + pow(He * w, m);
            const double mag = tC * He * w * pow(D, -1.0 / m);
            f = mag / speed;
            /* d mag / dU = mag (1/(m U)) (1 - (He w)^m / D) */
            fU = mag / (m * speed) * (1.0 - pow(He * w, m) / D) / speed - mag / (speed * speed);
            const double Dh = pow(tC, m) + pow(0.5 * w, m);   /* He = 1/2 */
            const double dmag_dHe = tC * w * pow(Dh, -1.0 / m) * (1.0 - pow(0.5 * w, m) / Dh);
            const double dmag_dN = muk * He * w * pow(D, -1.0 / m) * (1.0 - pow(tC, m) / D);
            dTdh = delta * dmag_dHe + dNdh * dmag_dN;
        } else if (P->law == 5) {
            /* |tau| = tauC tauW / (tauC + tauW) */
            const double tC = muk * N, tW = He * pow(Cp, -1.0 / m) * pow(speed, 1.0 / m);
            const double mag = tC * tW / (tC + tW);
            f = mag / speed;
            fU = (tC * tC / ((tC + tW) * (tC + tW))) * tW / (m * speed) / speed - mag / (speed * speed);
            const double tW2 = 0.5 * pow(Cp, -1.0 / m) * pow(speed, 1.0 / m);   /* He = 1/2 */
            const double dmag_dHe = tC * tC * pow(Cp, -1.0 / m) * pow(speed, 1.0 / m) / ((tC + tW2) * (tC + tW2));
            const double dmag_dN = muk * tW * tW / ((tC + tW) * (tC + tW)) / (He > 0 ? He : 1.0);
            dTdh = delta * dmag_dHe + dNdh * (He > 0 ? dmag_dN : 0.0);
        } else {
            /* |tau| = He C'^(-1/m) (U / (U + V0))^(1/m) */
            const double base = pow(Cp, -1.0 / m) * pow(speed / (speed + V0), 1.0 / m);
            const double mag = He * base;
            f = mag / speed;
            fU = mag * (1.0 / m) * (1.0 / speed - 1.0 / (speed + V0)) / speed - mag / (speed * speed);
            dTdh = delta * base;
        }
        const double g = fU / speed;
        *taux = f * u; *tauy = f * v;
        *dtxdu = f + u * u * g; *dtydv = f + v * v * g; *dtxdv = u * v * g;
        *dtxdh = dTdh * u / speed; *dtydh = dTdh * v / speed;
        return;
    }
    const double Tau = muk * N;
    const double cx = Tau * u / speed, cy = Tau * v / speed;
    if (P->law == 3 && !(cx * cx + cy * cy < *taux * *taux + *tauy * *tauy)) return;
    const double temp = speed * speed * speed;
    *taux = cx; *tauy = cy;
    *dtxdu = Tau * (1.0 / speed - u * u / temp);
    *dtydv = Tau * (1.0 / speed - v * v / temp);
    *dtxdv = -Tau * u * v / temp;
    const double E = I ? 0.0 : muk * dNdh / speed;
    *dtxdh = E * u; *dtydh = E * v;
}

/* One element: R parts (Tx,Fx,Ty,Fy,Th,Fh) x nod and 9 K blocks in the order
   Kxu Kxv Kxh Kyu Kyv Kyh Khu Khv Khh, each nod x nod row-major (i = row node). */
static void element(const o2_params *P, int nod, int nip, const double *pts, const double *wts,
                    const double *xn, const double *yn, const double nv[][MAXNOD], double area,
                    double *Rel, double *Kel)
{
    enum { H, U, V, S, B, RHO, C, M, NN, A, AS1, AB1, H0, U0, V0, AS0, AB0, DADH, BL, QQ, MUK, JV0 };
    const double ca = cos(P->alpha), sa = sin(P->alpha);
    const double g = P->g, rhow = P->rhow, theta = P->theta, dt = P->dt;
    double Tx[MAXNOD] = {0}, Fx[MAXNOD] = {0}, Ty[MAXNOD] = {0}, Fy[MAXNOD] = {0}, Th[MAXNOD] = {0}, Fh[MAXNOD] = {0};
    if (!P->ronly) memset(Kel, 0, sizeof(double) * 9 * nod * nod);
    for (int q = 0; q < nip; ++q) {
        double N[MAXNOD], Nxi[MAXNOD], Neta[MAXNOD], Nx[MAXNOD], Ny[MAXNOD];
        shape_barycentric(nod, pts[2 * q], pts[2 * q + 1], N, Nxi, Neta);
        double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
        for (int i = 0; i < nod; ++i) {
            J11 += Nxi[i] * xn[i]; J12 += Nxi[i] * yn[i];
            J21 += Neta[i] * xn[i]; J22 += Neta[i] * yn[i];
        }
        double det = J11 * J22 - J12 * J21;
        for (int i = 0; i < nod; ++i) {
            Nx[i] = (J22 * Nxi[i] - J12 * Neta[i]) / det;
            Ny[i] = (-J21 * Nxi[i] + J11 * Neta[i]) / det;
        }
        double val[JV0 + 1];
        for (int f = 0; f <= JV0; ++f) {
            double s = 0;
            for (int i = 0; i < nod; ++i) s += N[i] * nv[f][i];
            val[f] = s;
        }
        double h = val[H], u = val[U], v = val[V], rho = val[RHO];
        double h0 = val[H0], u0 = val[U0], v0 = val[V0];
        double a1 = val[AS1] + val[AB1], a0 = val[AS0] + val[AB0], dadh = val[DADH];
        /* gradients */
        double dhdx = 0, dhdy = 0, dh0dx = 0, dh0dy = 0, exx0 = 0, eyy0 = 0, dbdx = 0, dbdy = 0;
        double drhodx = 0, drhody = 0, exx = 0, eyy = 0, exy = 0;
        for (int i = 0; i < nod; ++i) {
            dhdx += Nx[i] * nv[H][i];   dhdy += Ny[i] * nv[H][i];
            dh0dx += Nx[i] * nv[H0][i]; dh0dy += Ny[i] * nv[H0][i];
            exx0 += Nx[i] * nv[U0][i];  eyy0 += Ny[i] * nv[V0][i];
            dbdx += Nx[i] * nv[BL][i];  dbdy += Ny[i] * nv[BL][i];
            drhodx += Nx[i] * nv[RHO][i]; drhody += Ny[i] * nv[RHO][i];
            exx += Nx[i] * nv[U][i];    eyy += Ny[i] * nv[V][i];
            exy += 0.5 * (Nx[i] * nv[V][i] + Ny[i] * nv[U][i]);
        }
        /* floatation */
        double Hth = val[S] - val[B];
        double hf = rhow * Hth / rho;
        double He = soft_step(P->kH, h - hf, P->Hh0);
        double HE = soft_step(P->kH, hf - h, P->Hh0);
        double dl = soft_delta(P->kH, h - hf, P->Hh0);
        double Dl = soft_delta(P->kH, hf - h, P->Hh0);
        double Hpos = soft_step(P->kH, Hth, P->Hh0) * Hth;
        double d = HE * rho * h / rhow + He * Hpos;
        double dd = HE * rho / rhow - Dl * h * rho / rhow + dl * Hpos;
        /* Glen's law */
        double n = val[NN], AG = val[A];
        double e = sqrt(P->eps_zero * P->eps_zero + exx * exx + eyy * eyy + exx * eyy + exy * exy);
        double eta = 0.5 * pow(AG, -1.0 / n) * pow(e, (1.0 - n) / n) + P->eta_zero;
        double E = (1.0 - n) / (4.0 * n) * pow(AG, -1.0 / n) * pow(e, (1.0 - 3.0 * n) / n);
        /* Weertman sliding */
        double m = val[M], Cq = val[C] + P->czero, uz = P->speed_zero;
        double Hec = He < 0 ? 0 : He, dlc = dl < 0 ? 0 : dl;
        double U2 = u * u + v * v + uz * uz;
        double beta2 = pow(Cq, -1.0 / m) * pow(sqrt(U2), 1.0 / m - 1.0);
        double Dbeta2 = (1.0 / m - 1.0) * pow(Cq, -1.0 / m) * pow(U2, (1.0 - 3.0 * m) / (2.0 * m));
        double taux = Hec * beta2 * u, tauy = Hec * beta2 * v;
        double dtxdu = Hec * (beta2 + Dbeta2 * u * u), dtydv = Hec * (beta2 + Dbeta2 * v * v);
        double dtxdv = Hec * Dbeta2 * u * v, dtydu = dtxdv;
        double dtxdh = dlc * beta2 * u, dtydh = dlc * beta2 * v;
        if (P->law != 0) {
            other_drag(P, h, Hth, rho, u, v, m, Cq, val[QQ], val[MUK], val[JV0], U2, beta2, Dbeta2, Hec, dlc,
                       &taux, &tauy, &dtxdu, &dtydv, &dtxdv, &dtxdh, &dtydh);
            dtydu = dtxdv;
        }
        /* SUPG */
        double l = sqrt(2.0 * area);
        double speed0 = sqrt(u0 * u0 + v0 * v0 + uz * uz);
        double taut = dt / 2.0 + DBL_EPSILON;
        double taus = l / (2.0 * speed0 + uz);
        double tau;
        switch (P->tau_option) {
            case 1: tau = taut; break;
            case 2: { double ecn = taut / taus; double K = (ecn < 1e-6) ? ecn / 3.0 : 1.0 / tanh(ecn) - 1.0 / ecn; tau = K * taus; } break;
            case 3: { double ecn = taut / taus; tau = (dt / 2.0) / (1.0 + ecn); } break;
            case 4: tau = 0.0; break;
            default: tau = taus;
        }
        double tau0 = P->beta0 * (P->tau_multiplier * tau);
        double W = det * wts[q];
        /* fluxes */
        double qx1 = rho * exx * h + rho * u * dhdx + drhodx * u * h;
        double qy1 = rho * eyy * h + rho * v * dhdy + drhody * v * h;
        double qx0 = rho * exx0 * h0 + rho * u0 * dh0dx + drhodx * u0 * h0;
        double qy0 = rho * eyy0 * h0 + rho * v0 * dh0dy + drhody * v0 * h0;
        for (int i = 0; i < nod; ++i) {
            double psi = N[i] + 0.5 * tau0 * (u0 * Nx[i] + v0 * Ny[i]);
            if (!P->ronly) {
                for (int j = 0; j < nod; ++j) {
                    double Deu = E * ((2 * exx + eyy) * Nx[j] + exy * Ny[j]);
                    double Dev = E * ((2 * eyy + exx) * Ny[j] + exy * Nx[j]);
                    double E11 = h * (4 * exx + 2 * eyy) * Deu * Nx[i] + 2 * h * exy * Deu * Ny[i];
                    double E12 = h * (4 * exx + 2 * eyy) * Dev * Nx[i] + 2 * h * exy * Dev * Ny[i];
                    double E22 = h * (4 * eyy + 2 * exx) * Dev * Ny[i] + 2 * h * exy * Dev * Nx[i];
                    double E21 = h * (4 * eyy + 2 * exx) * Deu * Ny[i] + 2 * h * exy * Deu * Nx[i];
                    double NN2 = N[i] * N[j];
                    double kxu = 4 * h * eta * Nx[i] * Nx[j] + h * eta * Ny[i] * Ny[j] + E11 + dtxdu * NN2;
                    double kyv = 4 * h * eta * Ny[i] * Ny[j] + h * eta * Nx[i] * Nx[j] + E22 + dtydv * NN2;
                    double kxv = eta * h * (2 * Nx[i] * Ny[j] + Ny[i] * Nx[j]) + E12 + dtxdv * NN2;
                    double kyu = eta * h * (2 * Ny[i] * Nx[j] + Nx[i] * Ny[j]) + E21 + dtydu * NN2;
                    double kxh = eta * (4 * exx + 2 * eyy) * Nx[i] * N[j] + eta * 2 * exy * Ny[i] * N[j]
                               + dtxdh * NN2 + ca * g * (rho - rhow * dd) * dbdx * NN2 - sa * g * rho * NN2
                               - ca * g * (rho * h - rhow * d * dd) * Nx[i] * N[j];
                    double kyh = eta * (4 * eyy + 2 * exx) * Ny[i] * N[j] + eta * 2 * exy * Nx[i] * N[j]
                               + dtydh * NN2 + ca * g * (rho - rhow * dd) * dbdy * NN2
                               - ca * g * (rho * h - rhow * d * dd) * Ny[i] * N[j];
                    double khu = theta * (rho * dhdx * N[j] + drhodx * h * N[j] + rho * h * Nx[j]) * psi * dt;
                    double khv = theta * (rho * dhdy * N[j] + drhody * h * N[j] + rho * h * Ny[j]) * psi * dt;
                    double khh = (rho * N[j] - dt * theta * rho * dadh * N[j]
                                  + dt * theta * (rho * exx * N[j] + drhodx * u * N[j] + rho * u * Nx[j]
                                                  + rho * eyy * N[j] + drhody * v * N[j] + rho * v * Ny[j])) * psi;
                    double vals[9] = {kxu, kxv, kxh, kyu, kyv, kyh, khu, khv, khh};
                    for (int bk = 0; bk < 9; ++bk) Kel[(bk * nod + i) * nod + j] += vals[bk] * W;
                }
            }
            double t1x = -ca * g * (rho * h - rhow * d) * dbdx * N[i] + rho * g * h * sa * N[i];
            double t2x = 0.5 * ca * g * (rho * h * h - rhow * d * d) * Nx[i];
            double t3x = h * eta * (4 * exx + 2 * eyy) * Nx[i];
            double t4x = h * eta * 2 * exy * Ny[i];
            Tx[i] += (t3x + t4x + taux * N[i]) * W;
            Fx[i] += (t1x + t2x) * W;
            double t1y = -ca * g * (rho * h - rhow * d) * dbdy * N[i];
            double t2y = 0.5 * ca * g * (rho * h * h - rhow * d * d) * Ny[i];
            double t3y = h * eta * (4 * eyy + 2 * exx) * Ny[i];
            double t4y = h * eta * 2 * exy * Nx[i];
            Ty[i] += (t3y + t4y + tauy * N[i]) * W;
            Fy[i] += (t1y + t2y) * W;
            double qterm = dt * (theta * qx1 + (1 - theta) * qx0 + theta * qy1 + (1 - theta) * qy0) * psi;
            double dhdt = rho * (h0 - h) * psi;
            double acc = dt * rho * ((1 - theta) * a0 + theta * a1) * psi;
            Th[i] += -dhdt * W;
            Fh[i] += (acc - qterm) * W;
        }
    }
    for (int i = 0; i < nod; ++i) {
        Rel[6 * i + 0] = Tx[i]; Rel[6 * i + 1] = Fx[i];
        Rel[6 * i + 2] = Ty[i]; Rel[6 * i + 3] = Fy[i];
        Rel[6 * i + 4] = Th[i]; Rel[6 * i + 5] = Fh[i];
    }
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Node adjacency (each node's sorted neighbour list, itself included) built
 * straight from the element lists: two passes with a worst-case buffer. */
int64_t o2_node_graph(int nele, int nod, int64_t nn, const int32_t *conn, int64_t *rowptr, int32_t *colind)
{
    /* incidence counts */
    int64_t *cnt = calloc((size_t)nn + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nele; ++e)
        for (int a = 0; a < nod; ++a) cnt[conn[e * nod + a] + 1] += nod;
    for (int64_t k = 0; k < nn; ++k) cnt[k + 1] += cnt[k];
    int32_t *buf = malloc(sizeof(int32_t) * (size_t)cnt[nn]);
    int64_t *fill = calloc((size_t)nn, sizeof(int64_t));
    for (int64_t e = 0; e < nele; ++e)
        for (int a = 0; a < nod; ++a) {
            int32_t r = conn[e * nod + a];
            for (int b = 0; b < nod; ++b) buf[cnt[r] + fill[r]++] = conn[e * nod + b];
        }
    int64_t total = 0;
    for (int64_t k = 0; k < nn; ++k) {
        int32_t *s = buf + cnt[k];
        int64_t len = cnt[k + 1] - cnt[k];
        qsort(s, (size_t)len, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t t = 0; t < len; ++t)
            if (t == 0 || s[t] != s[t - 1]) s[u++] = s[t];
        if (rowptr) rowptr[k] = total;
        if (colind) memcpy(colind + total, s, sizeof(int32_t) * (size_t)u);
        total += u;
    }
    if (rowptr) rowptr[nn] = total;
    free(cnt); free(buf); free(fill);
    return total;
}

static int64_t find_col(const int32_t *cols, int64_t len, int32_t c)
{
    int64_t lo = 0, hi = len - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (cols[mid] == c) return mid;
        if (cols[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/*
 * Full assembly: R (3nn), Tint, Fext (3nn, optional), K values (9*nnzN, optional)
 * on the CSR pattern rows r = blk_row*nn + node, row = [adj(node); adj+nn; adj+2nn].
 * Rel/Kel are caller-provided work arrays (nele*nod*6 and nele*9*nod*nod).
 * Returns 0, or 1 on a NaN in R, 2 on a NaN on the diagonal of K.
 */
int o2_assemble(const o2_params *P, int nele, int nod, int64_t nn, const int32_t *conn,
                const double *x, const double *y, int nip, const double *pts, const double *wts,
                const double *const *f1in, const double *const *f0in,
                const int64_t *rowptr_n, const int32_t *colind_n,
                double *R, double *Tint, double *Fext, double *Kval,
                double *Rel, double *Kel, double *bwork, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double *f1[NF1];
    const double *f0[NF0];
    double *zbuf = NULL;
    for (int k = 0; k < NF1; ++k) f1[k] = f1in[k];
    for (int k = 0; k < NF0; ++k) f0[k] = f0in[k];
    double *zf = NULL;
    if (P->zero_fields) {
        /* u = v = u0 = v0 = 0, h1 = h0, as <- |ab| + 1, ab <- |ab| (F0 and F1) */
        zf = malloc(sizeof(double) * (size_t)nn * 6);
        double *zero = zf, *as1 = zf + nn, *ab1 = zf + 2 * nn, *as0 = zf + 3 * nn, *ab0 = zf + 4 * nn;
        for (int64_t k = 0; k < nn; ++k) {
            zero[k] = 0.0;
            as1[k] = fabs(f1in[F1_AB][k]) + 1.0; ab1[k] = fabs(f1in[F1_AB][k]);
            as0[k] = fabs(f0in[F0_AB][k]) + 1.0; ab0[k] = fabs(f0in[F0_AB][k]);
        }
        f1[F1_U] = zero; f1[F1_V] = zero; f0[F0_U] = zero; f0[F0_V] = zero;
        f1[F1_H] = f0in[F0_H];
        f1[F1_AS] = as1; f1[F1_AB] = ab1; f0[F0_AS] = as0; f0[F0_AB] = ab0;
        f1[F1_DADH] = NULL;   /* no mass-balance feedback term under ZeroFields (uvhMatrixAssembly.m:143) */
    }
    nodal_b(P, nn, f1[F1_H], f1[F1_S], f1[F1_B], f1[F1_RHO], bwork);
    const int corners[3][3] = {{0, 1, 2}, {0, 2, 4}, {0, 3, 6}};
    const int *cn = corners[nod == 3 ? 0 : (nod == 6 ? 1 : 2)];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < nele; ++e) {
        double xn[MAXNOD], yn[MAXNOD], nv[22][MAXNOD];
        for (int a = 0; a < nod; ++a) {
            int32_t k = conn[e * nod + a];
            xn[a] = x[k]; yn[a] = y[k];
            nv[0][a] = f1[F1_H][k]; nv[1][a] = f1[F1_U][k]; nv[2][a] = f1[F1_V][k];
            nv[3][a] = f1[F1_S][k]; nv[4][a] = f1[F1_B][k]; nv[5][a] = f1[F1_RHO][k];
            nv[6][a] = f1[F1_C][k]; nv[7][a] = f1[F1_M][k]; nv[8][a] = f1[F1_N][k];
            nv[9][a] = f1[F1_A][k]; nv[10][a] = f1[F1_AS][k]; nv[11][a] = f1[F1_AB][k];
            nv[12][a] = f0[F0_H][k]; nv[13][a] = f0[F0_U][k]; nv[14][a] = f0[F0_V][k];
            nv[15][a] = f0[F0_AS][k]; nv[16][a] = f0[F0_AB][k];
            nv[17][a] = f1[F1_DADH] ? f1[F1_DADH][k] : 0.0;   /* dadh, uvhMatrixAssembly.m:151-158 */
            nv[18][a] = bwork[k];
            nv[19][a] = f1[F1_Q] ? f1[F1_Q][k] : 0.0;
            nv[20][a] = f1[F1_MUK] ? f1[F1_MUK][k] : 0.0;
            nv[21][a] = f1[F1_V0] ? f1[F1_V0][k] : 0.0;
        }
        double x1 = xn[cn[0]], x2 = xn[cn[1]], x3 = xn[cn[2]];
        double y1 = yn[cn[0]], y2 = yn[cn[1]], y3 = yn[cn[2]];
        double area = 0.5 * fabs((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1));
        element(P, nod, nip, pts, wts, xn, yn, (const double (*)[MAXNOD])nv, area,
                Rel + (size_t)e * nod * 6, P->ronly ? NULL : Kel + (size_t)e * 9 * nod * nod);
    }
    /* node incidence lists (element order) for a deterministic gather */
    int64_t *ip = calloc((size_t)nn + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nele; ++e)
        for (int a = 0; a < nod; ++a) ip[conn[e * nod + a] + 1]++;
    for (int64_t k = 0; k < nn; ++k) ip[k + 1] += ip[k];
    int64_t *inc = malloc(sizeof(int64_t) * (size_t)ip[nn]);
    int64_t *fill = calloc((size_t)nn, sizeof(int64_t));
    for (int64_t e = 0; e < nele; ++e)
        for (int a = 0; a < nod; ++a) {
            int32_t k = conn[e * nod + a];
            inc[ip[k] + fill[k]++] = e * nod + a;
        }
    const int64_t nnzN = rowptr_n ? rowptr_n[nn] : 0;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 512) reduction(| : bad)
    for (int64_t k = 0; k < nn; ++k) {
        double T[3] = {0, 0, 0}, F[3] = {0, 0, 0};
        for (int64_t t = ip[k]; t < ip[k + 1]; ++t) {
            int64_t ea = inc[t];
            const double *re = Rel + ea * 6;
            T[0] += re[0]; F[0] += re[1]; T[1] += re[2]; F[1] += re[3]; T[2] += re[4]; F[2] += re[5];
        }
        for (int bi = 0; bi < 3; ++bi) {
            R[bi * nn + k] = T[bi] - F[bi];
            if (Tint) Tint[bi * nn + k] = T[bi];
            if (Fext) Fext[bi * nn + k] = F[bi];
            if (isnan(R[bi * nn + k])) bad |= 1;
        }
        if (P->ronly || !Kval) continue;
        const int64_t deg = rowptr_n[k + 1] - rowptr_n[k];
        const int32_t *cols = colind_n + rowptr_n[k];
        for (int bi = 0; bi < 3; ++bi) {
            double *row = Kval + (int64_t)bi * 3 * nnzN + 3 * rowptr_n[k];
            memset(row, 0, sizeof(double) * (size_t)(3 * deg));
        }
        for (int64_t t = ip[k]; t < ip[k + 1]; ++t) {
            int64_t e = inc[t] / nod;
            int a = (int)(inc[t] % nod);
            for (int b = 0; b < nod; ++b) {
                int64_t pos = find_col(cols, deg, conn[e * nod + b]);
                for (int bi = 0; bi < 3; ++bi) {
                    double *row = Kval + (int64_t)bi * 3 * nnzN + 3 * rowptr_n[k];
                    for (int bj = 0; bj < 3; ++bj)
                        row[bj * deg + pos] += Kel[((size_t)e * 9 + bi * 3 + bj) * nod * nod + a * nod + b];
                }
            }
        }
        for (int bi = 0; bi < 3; ++bi) {
            const double *row = Kval + (int64_t)bi * 3 * nnzN + 3 * rowptr_n[k];
            int64_t p = find_col(cols, deg, (int32_t)k);
            if (p >= 0 && isnan(row[bi * deg + p])) bad |= 2;   /* p < 0: a node no element references */
        }
    }
    free(ip); free(inc); free(fill); free(zf); free(zbuf);
    if (bad & 2) return 2;
    if (bad & 1) return 1;
    return 0;
}

/*
 * uv (diagnostic SSTREAM) assembly: UaMain/uvMatrixAssemblySSTREAM.m (2025 form of :224-240,
 * uvGroupAssembly = false, no melange physics). Fields f[]: h s ub vb S B rho C m n AGlen.
 * Differences from the uvh element: s is an input and b = s - h at the points (:219), C and
 * AGlen are clipped at Cmin and AGlenmin (:185, :211), the viscosity derivative enters with
 * Dvisk = NRviscosity (:72-76, :418-421), there is no thickness equation and no SUPG.
 * R = Tint - Fext (2nn, [x; y]); K (4*nnzN values) on the CSR rows r = blk*nn + node with
 * row = [adj(node); adj + nn], symmetrised as (K + K')/2 (:545-549).
 */
enum { UV_H, UV_S, UV_U, UV_V, UV_SS, UV_B, UV_RHO, UV_C, UV_M, UV_N, UV_A, UV_Q, UV_MUK, UV_V0, NUV };

static void uv_element(const o2_params *P, double cmin, double aglenmin, double dvisk, int nod, int nip,
                       const double *pts, const double *wts, const double *xn, const double *yn,
                       const double nv[][MAXNOD], double *Rel, double *Kel)
{
    const double ca = cos(P->alpha), sa = sin(P->alpha), g = P->g, rhow = P->rhow;
    double Tx[MAXNOD] = {0}, Fx[MAXNOD] = {0}, Ty[MAXNOD] = {0}, Fy[MAXNOD] = {0};
    if (Kel) memset(Kel, 0, sizeof(double) * 4 * nod * nod);
    for (int q = 0; q < nip; ++q) {
        double N[MAXNOD], Nxi[MAXNOD], Neta[MAXNOD], Nx[MAXNOD], Ny[MAXNOD];
        shape_barycentric(nod, pts[2 * q], pts[2 * q + 1], N, Nxi, Neta);
        double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
        for (int i = 0; i < nod; ++i) {
            J11 += Nxi[i] * xn[i]; J12 += Nxi[i] * yn[i];
            J21 += Neta[i] * xn[i]; J22 += Neta[i] * yn[i];
        }
        const double det = J11 * J22 - J12 * J21;
        for (int i = 0; i < nod; ++i) {
            Nx[i] = (J22 * Nxi[i] - J12 * Neta[i]) / det;
            Ny[i] = (-J21 * Nxi[i] + J11 * Neta[i]) / det;
        }
        double val[NUV];
        for (int f = 0; f < NUV; ++f) {
            double s = 0;
            for (int i = 0; i < nod; ++i) s += N[i] * nv[f][i];
            val[f] = s;
        }
        const double h = val[UV_H], u = val[UV_U], v = val[UV_V], rho = val[UV_RHO];
        double Cq = val[UV_C];
        if (Cq < cmin) Cq = cmin;
        double AG = val[UV_A];
        if (AG < aglenmin) AG = aglenmin;
        const double m = val[UV_M], n = val[UV_N];
        const double Hth = val[UV_SS] - val[UV_B];
        double dsdx = 0, dsdy = 0, dhdx = 0, dhdy = 0, exx = 0, eyy = 0, exy = 0;
        for (int i = 0; i < nod; ++i) {
            dsdx += Nx[i] * nv[UV_S][i]; dsdy += Ny[i] * nv[UV_S][i];
            dhdx += Nx[i] * nv[UV_H][i]; dhdy += Ny[i] * nv[UV_H][i];
            exx += Nx[i] * nv[UV_U][i]; eyy += Ny[i] * nv[UV_V][i];
            exy += 0.5 * (Nx[i] * nv[UV_V][i] + Ny[i] * nv[UV_U][i]);
        }
        const double dbdx = dsdx - dhdx, dbdy = dsdy - dhdy;
        /* floatation at the point, 2025 form */
        const double hf = rhow * Hth / rho;
        const double He = soft_step(P->kH, h - hf, P->Hh0);
        const double HE = soft_step(P->kH, hf - h, P->Hh0);
        const double Hpos = soft_step(P->kH, Hth, P->Hh0) * Hth;
        const double d = HE * rho * h / rhow + He * Hpos;
        /* Weertman drag (BasalDrag.m) and Glen's law (EffectiveViscositySSTREAM.m) */
        const double uz = P->speed_zero, Cz = Cq + P->czero;
        const double Hec = He < 0 ? 0 : He;
        const double U2 = u * u + v * v + uz * uz;
        const double beta2 = pow(Cz, -1.0 / m) * pow(sqrt(U2), 1.0 / m - 1.0);
        const double Dbeta2 = (1.0 / m - 1.0) * pow(Cz, -1.0 / m) * pow(U2, (1.0 - 3.0 * m) / (2.0 * m));
        double taux = Hec * beta2 * u, tauy = Hec * beta2 * v;
        double dtxdu = Hec * (beta2 + Dbeta2 * u * u), dtydv = Hec * (beta2 + Dbeta2 * v * v);
        double dtxdv = Hec * Dbeta2 * u * v, dtydu = dtxdv;
        if (P->law != 0) {   /* the thickness derivatives are not used by the uv system */
            double dxh = 0.0, dyh = 0.0;
            const double dl = soft_delta(P->kH, h - hf, P->Hh0);
            other_drag(P, h, Hth, rho, u, v, m, Cz, val[UV_Q], val[UV_MUK], val[UV_V0], U2, beta2, Dbeta2, Hec,
                       dl < 0 ? 0 : dl, &taux, &tauy, &dtxdu, &dtydv, &dtxdv, &dxh, &dyh);
            dtydu = dtxdv;
        }
        const double e = sqrt(P->eps_zero * P->eps_zero + exx * exx + eyy * eyy + exx * eyy + exy * exy);
        const double eta = 0.5 * pow(AG, -1.0 / n) * pow(e, (1.0 - n) / n) + P->eta_zero;
        const double E = (1.0 - n) / (4.0 * n) * pow(AG, -1.0 / n) * pow(e, (1.0 - 3.0 * n) / n);
        const double W = det * wts[q];
        for (int i = 0; i < nod; ++i) {
            if (Kel) {
                for (int j = 0; j < nod; ++j) {
                    const double NN2 = N[i] * N[j];
                    const double Deu = E * ((2 * exx + eyy) * Nx[j] + exy * Ny[j]);
                    const double Dev = E * ((2 * eyy + exx) * Ny[j] + exy * Nx[j]);
                    const double E11 = h * (4 * exx + 2 * eyy) * Deu * Nx[i] + 2 * h * exy * Deu * Ny[i];
                    const double E12 = h * (4 * exx + 2 * eyy) * Dev * Nx[i] + 2 * h * exy * Dev * Ny[i];
                    const double E22 = h * (4 * eyy + 2 * exx) * Dev * Ny[i] + 2 * h * exy * Dev * Nx[i];
                    const double E21 = h * (4 * eyy + 2 * exx) * Deu * Ny[i] + 2 * h * exy * Deu * Nx[i];
                    const double d11 = 4 * h * eta * Nx[i] * Nx[j] + h * eta * Ny[i] * Ny[j] + dtxdu * NN2 + dvisk * E11;
                    const double d22 = 4 * h * eta * Ny[i] * Ny[j] + h * eta * Nx[i] * Nx[j] + dtydv * NN2 + dvisk * E22;
                    const double d12 = eta * h * (2 * Nx[i] * Ny[j] + Ny[i] * Nx[j]) + dtxdv * NN2 + dvisk * E12;
                    const double d21 = eta * h * (2 * Ny[i] * Nx[j] + Nx[i] * Ny[j]) + dtydu * NN2 + dvisk * E21;
                    Kel[(0 * nod + i) * nod + j] += d11 * W;   /* Kxu */
                    Kel[(1 * nod + i) * nod + j] += d12 * W;   /* Kxv */
                    Kel[(2 * nod + i) * nod + j] += d21 * W;   /* Kyu */
                    Kel[(3 * nod + i) * nod + j] += d22 * W;   /* Kyv */
                }
            }
            const double t1x = -g * (rho * h - rhow * d) * dbdx * N[i] * ca + rho * g * h * sa * N[i];
            const double t2x = 0.5 * g * ca * (rho * h * h - rhow * d * d) * Nx[i];
            const double t3x = h * eta * (4 * exx + 2 * eyy) * Nx[i];
            const double t4x = h * eta * 2 * exy * Ny[i];
            Tx[i] += (t3x + t4x + taux * N[i]) * W;
            Fx[i] += (t1x + t2x) * W;
            const double t1y = -g * (rho * h - rhow * d) * dbdy * N[i] * ca;
            const double t2y = 0.5 * ca * g * (rho * h * h - rhow * d * d) * Ny[i];
            const double t3y = h * eta * (4 * eyy + 2 * exx) * Ny[i];
            const double t4y = h * eta * 2 * exy * Nx[i];
            Ty[i] += (t3y + t4y + tauy * N[i]) * W;
            Fy[i] += (t1y + t2y) * W;
        }
    }
    for (int i = 0; i < nod; ++i) {
        Rel[4 * i + 0] = Tx[i]; Rel[4 * i + 1] = Fx[i];
        Rel[4 * i + 2] = Ty[i]; Rel[4 * i + 3] = Fy[i];
    }
}

int o2_uv_assemble(const o2_params *P, double cmin, double aglenmin, double dvisk, int nele, int nod, int64_t nn,
                   const int32_t *conn, const double *x, const double *y, int nip, const double *pts,
                   const double *wts, const double *const *f, const int64_t *rowptr_n, const int32_t *colind_n,
                   double *R, double *Tint, double *Fext, double *Kval, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int want_k = !P->ronly && Kval;
    double *Rel = malloc(sizeof(double) * (size_t)nele * nod * 4);
    double *Kel = want_k ? malloc(sizeof(double) * (size_t)nele * 4 * nod * nod) : NULL;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < nele; ++e) {
        double xn[MAXNOD], yn[MAXNOD], nv[NUV][MAXNOD];
        for (int a = 0; a < nod; ++a) {
            const int32_t k = conn[e * nod + a];
            xn[a] = x[k]; yn[a] = y[k];
            for (int fi = 0; fi < NUV; ++fi) nv[fi][a] = f[fi] ? f[fi][k] : 0.0;
        }
        uv_element(P, cmin, aglenmin, dvisk, nod, nip, pts, wts, xn, yn, (const double (*)[MAXNOD])nv,
                   Rel + (size_t)e * nod * 4, want_k ? Kel + (size_t)e * 4 * nod * nod : NULL);
    }
    /* R: element contributions summed per node in element order */
    double *T = calloc((size_t)nn * 2, sizeof(double)), *F = calloc((size_t)nn * 2, sizeof(double));
    for (int64_t e = 0; e < nele; ++e)
        for (int a = 0; a < nod; ++a) {
            const int32_t k = conn[e * nod + a];
            const double *re = Rel + ((size_t)e * nod + a) * 4;
            T[k] += re[0]; F[k] += re[1]; T[nn + k] += re[2]; F[nn + k] += re[3];
        }
    int bad = 0;
    for (int64_t r = 0; r < 2 * nn; ++r) {
        R[r] = T[r] - F[r];
        if (Tint) Tint[r] = T[r];
        if (Fext) Fext[r] = F[r];
        if (isnan(R[r])) bad = 1;
    }
    if (want_k) {
        const int64_t nnzN = rowptr_n[nn];
        double *Kraw = calloc((size_t)(4 * nnzN), sizeof(double));
        for (int64_t e = 0; e < nele; ++e)
            for (int a = 0; a < nod; ++a) {
                const int32_t k = conn[e * nod + a];
                const int64_t deg = rowptr_n[k + 1] - rowptr_n[k];
                const int32_t *cols = colind_n + rowptr_n[k];
                for (int b = 0; b < nod; ++b) {
                    const int64_t pos = find_col(cols, deg, conn[e * nod + b]);
                    for (int bi = 0; bi < 2; ++bi)
                        for (int bj = 0; bj < 2; ++bj)
                            Kraw[(int64_t)bi * 2 * nnzN + 2 * rowptr_n[k] + bj * deg + pos] +=
                                Kel[((size_t)e * 4 + bi * 2 + bj) * nod * nod + a * nod + b];
                }
            }
        /* (K + K')/2: entry (bi, k; bj, c) pairs with (bj, c; bi, k) */
#pragma omp parallel for schedule(dynamic, 512)
        for (int64_t k = 0; k < nn; ++k) {
            const int64_t deg = rowptr_n[k + 1] - rowptr_n[k];
            for (int bi = 0; bi < 2; ++bi)
                for (int bj = 0; bj < 2; ++bj)
                    for (int64_t t = 0; t < deg; ++t) {
                        const int32_t c = colind_n[rowptr_n[k] + t];
                        const int64_t degc = rowptr_n[c + 1] - rowptr_n[c];
                        const int64_t tp = find_col(colind_n + rowptr_n[c], degc, (int32_t)k);
                        const double a = Kraw[(int64_t)bi * 2 * nnzN + 2 * rowptr_n[k] + bj * deg + t];
                        const double b = Kraw[(int64_t)bj * 2 * nnzN + 2 * rowptr_n[c] + bi * degc + tp];
                        Kval[(int64_t)bi * 2 * nnzN + 2 * rowptr_n[k] + bj * deg + t] = 0.5 * (a + b);
                    }
        }
        free(Kraw);
    }
    free(Rel); free(Kel); free(T); free(F);
    return bad;
}
