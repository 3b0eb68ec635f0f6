/* Test double of the MATLAB C API (R2018a, interleaved complex): the declarations of the calls
 * that uasource_amd/mex/uvhAssemblyGPU_mex.c makes, implemented by mex_harness.c, so that the
 * gateway can be compiled and run without MATLAB. Not the MATLAB header; written from the
 * documented signatures of these functions. */
#include <stddef.h>
#include <stdbool.h>
typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef double mxDouble;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c);
void mxFree(void *p);
void *mxMalloc(mwSize n);
mxDouble *mxGetDoubles(const mxArray *a);
mwIndex *mxGetIr(const mxArray *a);
mwIndex *mxGetJc(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
bool mxIsComplex(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsSparse(const mxArray *a);
int mexAtExit(void (*f)(void));
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
void mexErrMsgTxt(const char *msg);
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
