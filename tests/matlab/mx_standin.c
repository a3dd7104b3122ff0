/*
 * Test double of MATLAB's mx/mex API (declared in tests/matlab/mex.h): real
 * and complex double matrices, char arrays, mxMalloc temporaries released
 * after each gateway call, and mexErrMsgIdAndTxt returning to the caller
 * through longjmp as MATLAB does.  standin_* are the hooks the Python test
 * uses to build arguments, call mexFunction and read the outputs.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
  size_t m, n;
  int complex_, is_char;
  double *data; /* m*n doubles, or 2*m*n interleaved */
  char *str;
};

static void *g_tmp[4096];
static int g_ntmp;
static jmp_buf g_jmp;
static char g_err[512];

size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
int mxIsDouble(const mxArray *a) { return !a->is_char; }
int mxIsComplex(const mxArray *a) { return a->complex_; }
int mxIsEmpty(const mxArray *a) { return a->m * a->n == 0; }
int mxIsChar(const mxArray *a) { return a->is_char; }
mxComplexDouble *mxGetComplexDoubles(const mxArray *a) { return a->complex_ ? (mxComplexDouble *)a->data : NULL; }
double *mxGetDoubles(const mxArray *a) { return a->complex_ || a->is_char ? NULL : a->data; }
size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }
double mxGetScalar(const mxArray *a) { return (a->m > 0 && a->n > 0) ? a->data[0] : 0.0; }

void *mxMalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (g_ntmp < 4096) g_tmp[g_ntmp++] = p;
  return p;
}

mxArray *mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity c) {
  mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
  a->m = m;
  a->n = n;
  a->complex_ = c == mxCOMPLEX;
  a->data = (double *)calloc((a->complex_ ? 2 : 1) * m * n + 1, sizeof(double));
  return a;
}

mxArray *mxCreateDoubleScalar(double v) {
  mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->data[0] = v;
  return a;
}

void mxDestroyArray(mxArray *a) {
  if (!a) return;
  free(a->data);
  free(a->str);
  free(a);
}

int mxGetString(const mxArray *a, char *buf, size_t len) {
  if (!a->is_char || strlen(a->str) + 1 > len) return 1;
  strcpy(buf, a->str);
  return 0;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  int k = snprintf(g_err, sizeof(g_err), "%s: ", id);
  vsnprintf(g_err + k, sizeof(g_err) - (size_t)k, fmt, ap);
  va_end(ap);
  longjmp(g_jmp, 1);
}

/* ---- hooks for the test ---- */

mxArray *standin_real(size_t m, size_t n, const double *v) {
  mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
  if (v) memcpy(a->data, v, sizeof(double) * m * n);
  return a;
}

mxArray *standin_complex(size_t m, size_t n, const double *interleaved) {
  mxArray *a = mxCreateDoubleMatrix(m, n, mxCOMPLEX);
  if (interleaved) memcpy(a->data, interleaved, 2 * sizeof(double) * m * n);
  return a;
}

mxArray *standin_string(const char *s) {
  mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
  a->m = 1;
  a->n = strlen(s);
  a->is_char = 1;
  a->str = strdup(s);
  a->data = (double *)calloc(1, sizeof(double));
  return a;
}

const double *standin_data(const mxArray *a) { return a->data; }
int standin_is_complex(const mxArray *a) { return a->complex_; }

/* mexFunction with MATLAB's error semantics: 0, or 1 with the message in
 * err (outputs not set) */
int standin_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *err, size_t errlen) {
  g_ntmp = 0;
  g_err[0] = 0;
  int rc = 0;
  if (setjmp(g_jmp) == 0) {
    mexFunction(nlhs, plhs, nrhs, prhs);
  } else {
    rc = 1;
  }
  for (int i = 0; i < g_ntmp; i++) free(g_tmp[i]);
  g_ntmp = 0;
  if (err && errlen) snprintf(err, errlen, "%s", g_err);
  return rc;
}
