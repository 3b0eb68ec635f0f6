/*
 * A test double of the few MATLAB C API functions the MEX gateway
 * (uasource_amd/mex/uvhAssemblyGPU_mex.c) calls, and a C entry point that runs its mexFunction
 * on caller-owned arrays. With it the gateway's argument handling (1-based column-major
 * connectivity, the ctrl and fields layouts, the CSC export of K, the NaN errors and their
 * identifiers, the unsupported-options status) is checked on the GPU against the library's own
 * entry points (tests/test_gpu_mex_gateway.py).
 *
 * mexErrMsgIdAndTxt / mexErrMsgTxt end the call with longjmp back to harness_call, which then
 * returns 1 with the identifier and message, as MATLAB would raise them to the caller.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    size_t m, n;
    double *pr;
    int sparse, owned;
    mwIndex *ir, *jc;
};

static mxArray *make(size_t m, size_t n, int owned) {
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->m = m;
    a->n = n;
    a->owned = owned;
    if (owned) a->pr = (double *)calloc(m * n > 0 ? m * n : 1, sizeof(double));
    return a;
}

static void destroy(mxArray *a) {
    if (!a) return;
    if (a->owned) {
        free(a->pr);
        free(a->ir);
        free(a->jc);
    }
    free(a);
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
    (void)c;
    return make(m, n, 1);
}

mxArray *mxCreateDoubleScalar(double v) {
    mxArray *a = make(1, 1, 1);
    a->pr[0] = v;
    return a;
}

mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c) {
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    (void)c;
    a->m = m;
    a->n = n;
    a->owned = 1;
    a->sparse = 1;
    a->pr = (double *)calloc(nzmax > 0 ? nzmax : 1, sizeof(double));
    a->ir = (mwIndex *)calloc(nzmax > 0 ? nzmax : 1, sizeof(mwIndex));
    a->jc = (mwIndex *)calloc(n + 1, sizeof(mwIndex));
    return a;
}

void mxFree(void *p) { free(p); }
void *mxMalloc(mwSize n) { return malloc(n > 0 ? n : 1); }
mxDouble *mxGetDoubles(const mxArray *a) { return a->pr; }
mwIndex *mxGetIr(const mxArray *a) { return a->ir; }
mwIndex *mxGetJc(const mxArray *a) { return a->jc; }
size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
double mxGetScalar(const mxArray *a) { return a->pr[0]; }
bool mxIsComplex(const mxArray *a) { (void)a; return false; }
bool mxIsDouble(const mxArray *a) { (void)a; return true; }
bool mxIsSparse(const mxArray *a) { return a->sparse != 0; }

static jmp_buf g_jmp;
static char g_id[128], g_msg[512];
static void (*g_atexit)(void) = NULL;

int mexAtExit(void (*f)(void)) {
    g_atexit = f;
    return 0;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    va_list ap;
    snprintf(g_id, sizeof(g_id), "%s", id);
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

void mexErrMsgTxt(const char *msg) {
    g_id[0] = 0;
    snprintf(g_msg, sizeof(g_msg), "%s", msg);
    longjmp(g_jmp, 1);
}

static mxArray *g_out[3];

static void clear_outputs(void) {
    int k;
    for (k = 0; k < 3; ++k) {
        destroy(g_out[k]);
        g_out[k] = NULL;
    }
}

/* [R, K, status] = uvhAssemblyGPU_mex(meshId, connectivity, coordinates, points, weights, ctrl,
 * fields, flags) with the inputs as column-major double arrays. Returns 0, or 1 after an error
 * (harness_error gives its identifier and message). The outputs stay until the next call. */
int harness_call(double mesh_id, const double *conn, int64_t nele, int64_t nod, const double *xy, int64_t nn,
                 const double *pts, int64_t nip, const double *wts, const double *ctrl, int64_t nctrl,
                 const double *fields, int64_t nfields, double flags, int nlhs) {
    mxArray *in[8];
    const double scal[2] = {mesh_id, flags};
    const double *data[8] = {&scal[0], conn, xy, pts, wts, ctrl, fields, &scal[1]};
    const size_t rows[8] = {1, (size_t)nele, (size_t)nn, (size_t)nip, (size_t)nip, 1, (size_t)nn, 1};
    const size_t cols[8] = {1, (size_t)nod, 2, 2, 1, (size_t)nctrl, (size_t)nfields, 1};
    int k, failed = 0;
    clear_outputs();
    g_id[0] = g_msg[0] = 0;
    for (k = 0; k < 8; ++k) {
        in[k] = make(rows[k], cols[k], 0);
        in[k]->pr = (double *)data[k];
    }
    if (setjmp(g_jmp) == 0)
        mexFunction(nlhs, g_out, 8, (const mxArray **)in);
    else
        failed = 1;
    for (k = 0; k < 8; ++k) destroy(in[k]);
    return failed;
}

const char *harness_error_id(void) { return g_id; }
const char *harness_error_msg(void) { return g_msg; }

/* R (3 Nn values) of the last call */
const double *harness_R(void) { return g_out[0] ? g_out[0]->pr : NULL; }

/* K of the last call: -1 when it was not returned, 0 when empty ([]), else its nonzeros;
 * harness_K_copy fills jc (n + 1), ir and pr (nnz) */
int64_t harness_K_nnz(int64_t *n) {
    if (!g_out[1]) return -1;
    *n = (int64_t)g_out[1]->n;
    if (!g_out[1]->sparse) return 0;
    return (int64_t)g_out[1]->jc[g_out[1]->n];
}

void harness_K_copy(int64_t *jc, int64_t *ir, double *pr) {
    const mxArray *K = g_out[1];
    size_t i, nnz = K->jc[K->n];
    for (i = 0; i <= K->n; ++i) jc[i] = (int64_t)K->jc[i];
    for (i = 0; i < nnz; ++i) {
        ir[i] = (int64_t)K->ir[i];
        pr[i] = K->pr[i];
    }
}

double harness_status(void) { return g_out[2] ? g_out[2]->pr[0] : -1.0; }

/* MATLAB's clear mex / exit: the gateway's registered release */
void harness_exit(void) {
    clear_outputs();
    if (g_atexit) g_atexit();
    g_atexit = NULL;
}
