/*
 * Test double of the part of MATLAB's C matrix API (mex.h, R2018a
 * interleaved complex) that fmmlib3d_amd/matlab/fmm3d_mi355x_mex.c uses, so
 * that the gateway can be compiled and called without MATLAB
 * (tests/test_gpu_matlab_gateway.py).  Not MATLAB's header: only the
 * functions the gateway calls, with MATLAB's documented meaning.
 */
#ifndef FMM3D_MEX_STANDIN_H
#define FMM3D_MEX_STANDIN_H
#include <stddef.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef struct {
  double real, imag;
} mxComplexDouble;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;

size_t mxGetNumberOfElements(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsEmpty(const mxArray *a);
int mxIsChar(const mxArray *a);
mxComplexDouble *mxGetComplexDoubles(const mxArray *a);
double *mxGetDoubles(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
double mxGetScalar(const mxArray *a);
void *mxMalloc(size_t n);
mxArray *mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
void mxDestroyArray(mxArray *a);
int mxGetString(const mxArray *a, char *buf, size_t len);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#endif
