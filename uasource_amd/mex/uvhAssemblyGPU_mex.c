/*
 * uvhAssemblyGPU_mex.c -- MATLAB MEX gateway from Úa's uvhAssembly to libuvh (include/uvh.h).
 *
 *   [R, K, status] = uvhAssemblyGPU_mex(meshId, connectivity, coordinates, points, weights, ...
 *                                       ctrl, fields, flags)
 *
 *   meshId        double scalar that changes whenever MUA changes; the mesh (and the CSC map)
 *                 is uploaded again only when it differs from the last call's
 *   connectivity  MUA.connectivity, Nele x nod double, 1-based (UaMain/CreateMUA.m:31-61)
 *   coordinates   MUA.coordinates, Nn x 2 double
 *   points        MUA.points, nip x 2 double;  weights  MUA.weights, nip x 1 double
 *   ctrl          1 x 19 double: dt theta SUPG.beta0 uvh.SUPG.tauMultiplier SpeedZero EpsZero
 *                 etaZero Czero kH Hh0 F.g F.rhow F.alpha ThickMin tau(0..4) resetThickness
 *                 Enforce_bAboveB slidingLaw(0..6) Nzero   (the fields of uvh_ctrl, in order)
 *   fields        Nn x 21 double: h ub vb S B rho C m n AGlen as ab (F1), h ub vb as ab (F0), dadh,
 *                 q muk V0 (F1; read only by the sliding laws that use them)
 *   flags         double: UVH_RONLY | UVH_ZEROFIELDS | UVH_FP32 bits of include/uvh.h
 *
 *   R      3Nn x 1 double, [x; y; h] rows, R = Tint - Fext (UaMain/uvhMatrixAssembly.m:455)
 *   K      3Nn x 3Nn sparse double (CSC, sorted row indices), [] when Ronly
 *   status 0, or UVH_E_UNSUPPORTED (5): the caller falls back to uvhMatrixAssembly
 *
 * Errors: the NaN checks raise the reference's errors: uvhMatrixAssembly:NaNinR
 * (UaMain/uvhMatrixAssembly.m:581-584) through mexErrMsgIdAndTxt, and ' NaN in K ', which has
 * no identifier there (:575-579), through mexErrMsgTxt.
 * Device state (context, mesh, CSC map) lives in this file's statics; MATLAB calls the gateway
 * on its main thread only, and mexAtExit releases the state. Under spmd each worker process
 * holds its own context.
 *
 * Build (where MATLAB is installed):
 *   mex -R2018a uvhAssemblyGPU_mex.c -I../../include -L.. -luvh
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "uvh.h"

static uvh_ctx *g_ctx = NULL;
static double g_mesh_id = -1.0;
static int64_t g_neq = 0, g_nnz = 0;
static int64_t *g_rowptr = NULL, *g_cscmap = NULL;
static int32_t *g_colind = NULL;
static double *g_kval = NULL;

static void release(void) {
    free(g_rowptr); free(g_colind); free(g_cscmap); free(g_kval);
    g_rowptr = g_cscmap = NULL; g_colind = NULL; g_kval = NULL;
    if (g_ctx) uvh_destroy(g_ctx);
    g_ctx = NULL;
    g_mesh_id = -1.0;
}

static void fail(int st) {
    if (st == UVH_E_NAN_R) mexErrMsgIdAndTxt("uvhMatrixAssembly:NaNinR", " NaN in R ");
    /* the reference raises error(' NaN in K ') without an identifier (uvhMatrixAssembly.m:577) */
    if (st == UVH_E_NAN_K) mexErrMsgTxt(" NaN in K ");
    mexErrMsgIdAndTxt("uvhAssemblyGPU:failed", "%s", g_ctx ? uvh_last_error(g_ctx) : "no device context");
}

static void require(int cond, const char *msg) {
    if (!cond) mexErrMsgIdAndTxt("uvhAssemblyGPU:arguments", "%s", msg);
}

static void set_mesh(double id, const mxArray *connA, const mxArray *coordA, const mxArray *ptsA,
                     const mxArray *wtsA) {
    const mwSize nele = mxGetM(connA), nod = mxGetN(connA), nn = mxGetM(coordA), nip = mxGetM(ptsA);
    const double *conn = mxGetDoubles(connA), *xy = mxGetDoubles(coordA);
    const double *pts = mxGetDoubles(ptsA), *wts = mxGetDoubles(wtsA);
    int32_t *c = (int32_t *)mxMalloc(nele * nod * sizeof(int32_t));
    double *p = (double *)mxMalloc(nip * 2 * sizeof(double));
    mwSize e, a;
    int st;
    /* column-major, 1-based Nele x nod -> element-major, 0-based */
    for (e = 0; e < nele; ++e)
        for (a = 0; a < nod; ++a) c[e * nod + a] = (int32_t)conn[a * nele + e] - 1;
    for (e = 0; e < nip; ++e) { p[2 * e] = pts[e]; p[2 * e + 1] = pts[nip + e]; }
    /* uvh_set_mesh drops the context's old mesh before it checks the new one: forget the old
       mesh and its pattern first, so that a failed call leaves no stale mesh id behind */
    g_mesh_id = -1.0;
    free(g_rowptr); free(g_colind); free(g_cscmap); free(g_kval);
    g_rowptr = g_cscmap = NULL; g_colind = NULL; g_kval = NULL;
    st = uvh_set_mesh(g_ctx, (int32_t)nele, (int32_t)nod, (int32_t)nn, c, xy, xy + nn, (int32_t)nip, p, wts);
    mxFree(c);
    mxFree(p);
    if (st != UVH_OK) fail(st);
    uvh_pattern_info(g_ctx, &g_neq, &g_nnz);
    g_rowptr = (int64_t *)malloc((size_t)(g_neq + 1) * sizeof(int64_t));
    g_colind = (int32_t *)malloc((size_t)g_nnz * sizeof(int32_t));
    g_cscmap = (int64_t *)malloc((size_t)g_nnz * sizeof(int64_t));
    g_kval = (double *)malloc((size_t)g_nnz * sizeof(double));
    require(g_rowptr && g_colind && g_cscmap && g_kval, "out of host memory for the pattern of K");
    if ((st = uvh_get_pattern(g_ctx, g_rowptr, g_colind)) != UVH_OK) fail(st);
    if ((st = uvh_get_csc_map(g_ctx, g_cscmap)) != UVH_OK) fail(st);
    g_mesh_id = id;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    double id;
    const double *cv, *F;
    mwSize nn, k;
    uint32_t flags;
    uvh_ctrl ctrl;
    uvh_fields f;
    double *R;
    int st, ronly;

    require(nrhs == 8, "uvhAssemblyGPU_mex needs 8 arguments");
    require(nlhs <= 3, "uvhAssemblyGPU_mex returns [R, K, status]");
    for (k = 0; k < 8; ++k) require(mxIsDouble(prhs[k]) && !mxIsComplex(prhs[k]) && !mxIsSparse(prhs[k]),
                                    "arguments must be real full double arrays");
    if (!g_ctx) {
        if (uvh_create(&g_ctx, 0) != UVH_OK) mexErrMsgIdAndTxt("uvhAssemblyGPU:device", "no HIP device");
        mexAtExit(release);
    }
    id = mxGetScalar(prhs[0]);
    if (id != g_mesh_id) set_mesh(id, prhs[1], prhs[2], prhs[3], prhs[4]);

    nn = mxGetM(prhs[2]);
    require(mxGetNumberOfElements(prhs[5]) == 19, "ctrl must have 19 entries");
    require(mxGetM(prhs[6]) == nn && mxGetN(prhs[6]) == 21, "fields must be Nn x 21");
    cv = mxGetDoubles(prhs[5]);
    memset(&ctrl, 0, sizeof(ctrl));
    ctrl.dt = cv[0]; ctrl.theta = cv[1]; ctrl.supg_beta0 = cv[2]; ctrl.supg_tau_multiplier = cv[3];
    ctrl.speed_zero = cv[4]; ctrl.eps_zero = cv[5]; ctrl.eta_zero = cv[6]; ctrl.czero = cv[7];
    ctrl.kH = cv[8]; ctrl.Hh0 = cv[9]; ctrl.g = cv[10]; ctrl.rhow = cv[11]; ctrl.alpha = cv[12];
    ctrl.thick_min = cv[13]; ctrl.supg_tau = (int32_t)cv[14]; ctrl.reset_thickness = (int32_t)cv[15];
    ctrl.enforce_b_above_B = (int32_t)cv[16]; ctrl.sliding_law = (int32_t)cv[17]; ctrl.nzero = cv[18];
    F = mxGetDoubles(prhs[6]);
    f.h = F; f.ub = F + nn; f.vb = F + 2 * nn; f.S = F + 3 * nn; f.B = F + 4 * nn; f.rho = F + 5 * nn;
    f.C = F + 6 * nn; f.m = F + 7 * nn; f.n = F + 8 * nn; f.AGlen = F + 9 * nn; f.as = F + 10 * nn;
    f.ab = F + 11 * nn; f.h0 = F + 12 * nn; f.ub0 = F + 13 * nn; f.vb0 = F + 14 * nn; f.as0 = F + 15 * nn;
    f.ab0 = F + 16 * nn; f.dadh = F + 17 * nn; f.q = F + 18 * nn; f.muk = F + 19 * nn;
    f.V0 = F + 20 * nn;
    flags = (uint32_t)mxGetScalar(prhs[7]);
    ronly = (flags & UVH_RONLY) != 0;

    plhs[0] = mxCreateDoubleMatrix(3 * nn, 1, mxREAL);
    R = mxGetDoubles(plhs[0]);
    st = uvh_assemble(g_ctx, &ctrl, &f, flags, R, ronly ? NULL : g_kval, NULL, NULL);
    if (st == UVH_E_UNSUPPORTED) {
        if (nlhs > 1) plhs[1] = mxCreateDoubleMatrix(0, 0, mxREAL);
        if (nlhs > 2) plhs[2] = mxCreateDoubleScalar((double)st);
        return;
    }
    if (st != UVH_OK) fail(st);
    if (nlhs > 1) {
        if (ronly) {
            plhs[1] = mxCreateDoubleMatrix(0, 0, mxREAL);
        } else {
            /* CSC: column pointers and row indices are the CSR arrays (symmetric pattern) */
            mxArray *K = mxCreateSparse((mwSize)g_neq, (mwSize)g_neq, (mwSize)g_nnz, mxREAL);
            mwIndex *jc = mxGetJc(K), *ir = mxGetIr(K);
            double *pr = mxGetDoubles(K);
            int64_t i;
            for (i = 0; i <= g_neq; ++i) jc[i] = (mwIndex)g_rowptr[i];
            for (i = 0; i < g_nnz; ++i) {
                ir[i] = (mwIndex)g_colind[i];
                pr[i] = g_kval[g_cscmap[i]];
            }
            plhs[1] = K;
        }
    }
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(0.0);
}
