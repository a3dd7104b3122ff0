/*
 * MATLAB gateway over the C ABI of libfmm3d_mi355x.so (include/fmm3d_mi355x.h).
 * It replaces the reference's mwrap-generated gateway for the particle FMM
 * (matlab/fmm3d.mw:829-960, matlab/fmm3d.c): the same routines, arrays in
 * the same shapes, called through the drop-in entry points.  The .m files
 * beside it (lfmm3dpart.m, hfmm3dpart.m, l3dpartdirect.m, h3dpartdirect.m)
 * keep the reference's MATLAB calling sequences and output struct.
 *
 *   [ier,pot,fld,pottarg,fldtarg] = fmm3d_mi355x_mex('lfmm3dparttarg', iprec,
 *        source, charge, dipstr, dipvec, ifpot, iffld, target, ifpottarg, iffldtarg)
 *   [ier,...] = fmm3d_mi355x_mex('hfmm3dparttarg', iprec, zk, source, ...)
 *   [pot,fld,pottarg,fldtarg] = fmm3d_mi355x_mex('l3dpartdirect', source,
 *        charge, dipstr, dipvec, ifpot, iffld, target, ifpottarg, iffldtarg)
 *   [pot,...] = fmm3d_mi355x_mex('h3dpartdirect', zk, source, ...)
 *   [ier,pot,fld,pottarg,fldtarg] = fmm3d_mi355x_mex('lfmm3dtriatarg', iprec,
 *        triaflat, source, ifcharge, charge, ifdipole, dipstr, ifpot, iffld,
 *        target, ifpottarg, iffldtarg)
 *   [ier,...] = fmm3d_mi355x_mex('hfmm3dtriatarg', iprec, zk, triaflat, ...)
 *   [pot,fld,pottarg,fldtarg] = fmm3d_mi355x_mex('l3dtriadirect', triaflat,
 *        source, ifcharge, charge, ifdipole, dipstr, ifpot, iffld, target,
 *        ifpottarg, iffldtarg)
 *   [pot,...] = fmm3d_mi355x_mex('h3dtriadirect', nqtri, zk, triaflat, ...)
 *
 * (the triangle FMMs, matlab/fmm3d.mw of the reference: triaflat is 9 x n --
 * the three vertices of each triangle -- and the dipole layer is along the
 * normal of the vertex order; ifcharge / ifdipole 2 give the Helmholtz
 * difference kernel)
 *
 * source/target: real 3 x n; charge/dipstr: complex (real accepted) vectors
 * of n; dipvec: real 3 x n; an empty charge/dipstr turns that term off and an
 * empty target means no targets.  Outputs whose flag is 0 come back empty.
 * Build (MATLAB R2018a or later, interleaved complex):
 *   mex -R2018a -I../../include fmm3d_mi355x_mex.c -L../lib -lfmm3d_mi355x
 * (build_mex.m does this).
 */
#include <string.h>

#include "mex.h"
#include "fmm3d_mi355x.h"

/* complex vector of n values as interleaved doubles; a real input is copied
 * into a temporary with zero imaginary parts (freed by MATLAB at exit of the
 * gateway) */
static const double *complex_in(const mxArray *a, mwSize n, const char *name) {
  if (mxGetNumberOfElements(a) != n) mexErrMsgIdAndTxt("fmm3d:shape", "%s must have %d elements", name, (int)n);
  if (!mxIsDouble(a)) mexErrMsgIdAndTxt("fmm3d:type", "%s must be double", name);
  if (mxIsComplex(a)) return (const double *)mxGetComplexDoubles(a);
  double *t = (double *)mxMalloc(2 * n * sizeof(double));
  const double *re = mxGetDoubles(a);
  for (mwSize i = 0; i < n; i++) {
    t[2 * i] = re[i];
    t[2 * i + 1] = 0.0;
  }
  return t;
}

static const double *points_in(const mxArray *a, const char *name, mwSize *n) {
  if (mxIsEmpty(a)) {
    *n = 0;
    return NULL;
  }
  if (!mxIsDouble(a) || mxIsComplex(a) || mxGetM(a) != 3)
    mexErrMsgIdAndTxt("fmm3d:shape", "%s must be a real 3 x n array", name);
  *n = mxGetN(a);
  return mxGetDoubles(a);
}

static int flag(const mxArray *a) { return (int)(mxGetScalar(a) == 1.0); }

static mxArray *complex_out(mwSize m, mwSize n, int on, double **p) {
  if (!on) {
    *p = NULL;
    return mxCreateDoubleMatrix(0, 0, mxCOMPLEX);
  }
  mxArray *a = mxCreateDoubleMatrix(m, n, mxCOMPLEX);
  *p = (double *)mxGetComplexDoubles(a);
  return a;
}

/* arguments after the routine name (and iprec / zk): source, charge, dipstr,
 * dipvec, ifpot, iffld, target, ifpottarg, iffldtarg */
static void particle_call(int kind, int nlhs, mxArray *plhs[], const mxArray *prhs[], int iprec,
                          const double *zk) {
  mwSize ns, nt;
  const double *src = points_in(prhs[0], "source", &ns);
  if (ns < 1) mexErrMsgIdAndTxt("fmm3d:shape", "source must hold at least one point");
  const int ifcharge = !mxIsEmpty(prhs[1]);
  const int ifdipole = !mxIsEmpty(prhs[2]);
  const double *q = ifcharge ? complex_in(prhs[1], ns, "charge") : NULL;
  const double *ds = ifdipole ? complex_in(prhs[2], ns, "dipstr") : NULL;
  const double *dv = NULL;
  if (ifdipole) {
    if (mxGetM(prhs[3]) != 3 || mxGetN(prhs[3]) != ns || mxIsComplex(prhs[3]))
      mexErrMsgIdAndTxt("fmm3d:shape", "dipvec must be a real 3 x nsource array");
    dv = mxGetDoubles(prhs[3]);
  }
  const int ifpot = flag(prhs[4]), iffld = flag(prhs[5]);
  const double *trg = points_in(prhs[6], "target", &nt);
  const int ifpottarg = nt > 0 && flag(prhs[7]), iffldtarg = nt > 0 && flag(prhs[8]);
  double *pot, *fld, *pt, *ft;
  const int direct = kind >= 2, off = direct ? 0 : 1;
  mxArray *out[5];
  out[off + 0] = complex_out(1, ns, ifpot, &pot);
  out[off + 1] = complex_out(3, ns, iffld, &fld);
  out[off + 2] = complex_out(1, nt, ifpottarg, &pt);
  out[off + 3] = complex_out(3, nt, iffldtarg, &ft);
  int ier = 0;
  switch (kind) {
    case 0:
      ier = fmm3d_lfmm3dparttarg(iprec, (int)ns, src, ifcharge, q, ifdipole, ds, dv, ifpot, pot, iffld, fld,
                                 (int)nt, trg, ifpottarg, pt, iffldtarg, ft);
      break;
    case 1:
      ier = fmm3d_hfmm3dparttarg(iprec, zk, (int)ns, src, ifcharge, q, ifdipole, ds, dv, ifpot, pot, iffld, fld,
                                 (int)nt, trg, ifpottarg, pt, iffldtarg, ft);
      break;
    case 2:
      fmm3d_l3dpartdirect((int)ns, src, ifcharge, q, ifdipole, ds, dv, ifpot, pot, iffld, fld, (int)nt, trg,
                          ifpottarg, pt, iffldtarg, ft);
      break;
    default:
      fmm3d_h3dpartdirect(zk, (int)ns, src, ifcharge, q, ifdipole, ds, dv, ifpot, pot, iffld, fld, (int)nt, trg,
                          ifpottarg, pt, iffldtarg, ft);
      break;
  }
  if (!direct) out[0] = mxCreateDoubleScalar((double)ier);
  const int nout = direct ? 4 : 5;
  for (int k = 0; k < nout; k++) {
    if (k < (nlhs > 0 ? nlhs : 1)) plhs[k] = out[k];
    else mxDestroyArray(out[k]);
  }
}

/* arguments after the routine name (and iprec / zk / nqtri): triaflat,
 * source, ifcharge, charge, ifdipole, dipstr, ifpot, iffld, target,
 * ifpottarg, iffldtarg */
static void triangle_call(int kind, int nlhs, mxArray *plhs[], const mxArray *prhs[], int iprec,
                          const double *zk) {
  mwSize ns, nt;
  if (!mxIsDouble(prhs[0]) || mxIsComplex(prhs[0]) || mxGetM(prhs[0]) != 9)
    mexErrMsgIdAndTxt("fmm3d:shape", "triaflat must be a real 9 x n array");
  ns = mxGetN(prhs[0]);
  if (ns < 1) mexErrMsgIdAndTxt("fmm3d:shape", "triaflat must hold at least one triangle");
  const double *tri = mxGetDoubles(prhs[0]);
  mwSize nsrc;
  const double *src = points_in(prhs[1], "source", &nsrc);
  if (nsrc != ns) mexErrMsgIdAndTxt("fmm3d:shape", "source must be 3 x (number of triangles)");
  const int ifcharge = (int)mxGetScalar(prhs[2]), ifdipole = (int)mxGetScalar(prhs[4]);
  const double *q = ifcharge ? complex_in(prhs[3], ns, "charge") : NULL;
  const double *ds = ifdipole ? complex_in(prhs[5], ns, "dipstr") : NULL;
  /* the unit normals of the vertex order (the reference's trianorm and dipvec) */
  double *nrm = (double *)mxMalloc(3 * ns * sizeof(double));
  for (mwSize i = 0; i < ns; i++) {
    const double *v = tri + 9 * i;
    const double u[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, w[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    double l = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    l = l > 0 ? 1.0 / __builtin_sqrt(l) : 0.0;
    for (int k = 0; k < 3; k++) nrm[3 * i + k] = n[k] * l;
  }
  const int ifpot = flag(prhs[6]), iffld = flag(prhs[7]);
  const double *trg = points_in(prhs[8], "target", &nt);
  const int ifpottarg = nt > 0 && flag(prhs[9]), iffldtarg = nt > 0 && flag(prhs[10]);
  double *pot, *fld, *pt, *ft;
  const int direct = kind >= 2, off = direct ? 0 : 1;
  mxArray *out[5];
  out[off + 0] = complex_out(1, ns, ifpot, &pot);
  out[off + 1] = complex_out(3, ns, iffld, &fld);
  out[off + 2] = complex_out(1, nt, ifpottarg, &pt);
  out[off + 3] = complex_out(3, nt, iffldtarg, &ft);
  int ier = 0;
  switch (kind) {
    case 0:
      ier = fmm3d_lfmm3dtriatarg(iprec, (int)ns, tri, nrm, src, ifcharge, q, ifdipole, ds, nrm, ifpot, pot, iffld,
                                 fld, (int)nt, trg, ifpottarg, pt, iffldtarg, ft);
      break;
    case 1:
      ier = fmm3d_hfmm3dtriatarg(iprec, zk, (int)ns, tri, nrm, src, ifcharge, q, ifdipole, ds, nrm, ifpot, pot,
                                 iffld, fld, (int)nt, trg, ifpottarg, pt, iffldtarg, ft);
      break;
    case 2:
      fmm3d_l3dtriadirect((int)ns, tri, nrm, src, ifcharge, q, ifdipole, ds, nrm, ifpot, pot, iffld, fld, (int)nt,
                          trg, ifpottarg, pt, iffldtarg, ft);
      break;
    default: /* iprec carries nqtri */
      fmm3d_h3dtriadirect(iprec, zk, (int)ns, tri, nrm, src, ifcharge, q, ifdipole, ds, nrm, ifpot, pot, iffld,
                          fld, (int)nt, trg, ifpottarg, pt, iffldtarg, ft);
      break;
  }
  if (!direct) out[0] = mxCreateDoubleScalar((double)ier);
  const int nout = direct ? 4 : 5;
  for (int k = 0; k < nout; k++) {
    if (k < (nlhs > 0 ? nlhs : 1)) plhs[k] = out[k];
    else mxDestroyArray(out[k]);
  }
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  char name[32];
  if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], name, sizeof(name)) != 0)
    mexErrMsgIdAndTxt("fmm3d:usage", "first argument: routine name");
  double zk[2] = {0.0, 0.0};
  if (!strcmp(name, "lfmm3dparttarg")) {
    if (nrhs != 11 || nlhs > 5) mexErrMsgIdAndTxt("fmm3d:usage", "lfmm3dparttarg: 10 arguments after the name");
    particle_call(0, nlhs, plhs, prhs + 2, (int)mxGetScalar(prhs[1]), NULL);
  } else if (!strcmp(name, "hfmm3dparttarg")) {
    if (nrhs != 12 || nlhs > 5) mexErrMsgIdAndTxt("fmm3d:usage", "hfmm3dparttarg: 11 arguments after the name");
    zk[0] = mxGetScalar(prhs[2]);
    if (mxIsComplex(prhs[2])) zk[1] = ((const double *)mxGetComplexDoubles(prhs[2]))[1];
    particle_call(1, nlhs, plhs, prhs + 3, (int)mxGetScalar(prhs[1]), zk);
  } else if (!strcmp(name, "l3dpartdirect")) {
    if (nrhs != 10 || nlhs > 4) mexErrMsgIdAndTxt("fmm3d:usage", "l3dpartdirect: 9 arguments after the name");
    particle_call(2, nlhs, plhs, prhs + 1, 0, NULL);
  } else if (!strcmp(name, "h3dpartdirect")) {
    if (nrhs != 11 || nlhs > 4) mexErrMsgIdAndTxt("fmm3d:usage", "h3dpartdirect: 10 arguments after the name");
    zk[0] = mxGetScalar(prhs[1]);
    if (mxIsComplex(prhs[1])) zk[1] = ((const double *)mxGetComplexDoubles(prhs[1]))[1];
    particle_call(3, nlhs, plhs, prhs + 2, 0, zk);
  } else if (!strcmp(name, "lfmm3dtriatarg")) {
    if (nrhs != 13 || nlhs > 5) mexErrMsgIdAndTxt("fmm3d:usage", "lfmm3dtriatarg: 12 arguments after the name");
    triangle_call(0, nlhs, plhs, prhs + 2, (int)mxGetScalar(prhs[1]), NULL);
  } else if (!strcmp(name, "hfmm3dtriatarg")) {
    if (nrhs != 14 || nlhs > 5) mexErrMsgIdAndTxt("fmm3d:usage", "hfmm3dtriatarg: 13 arguments after the name");
    zk[0] = mxGetScalar(prhs[2]);
    if (mxIsComplex(prhs[2])) zk[1] = ((const double *)mxGetComplexDoubles(prhs[2]))[1];
    triangle_call(1, nlhs, plhs, prhs + 3, (int)mxGetScalar(prhs[1]), zk);
  } else if (!strcmp(name, "l3dtriadirect")) {
    if (nrhs != 12 || nlhs > 4) mexErrMsgIdAndTxt("fmm3d:usage", "l3dtriadirect: 11 arguments after the name");
    triangle_call(2, nlhs, plhs, prhs + 1, 0, NULL);
  } else if (!strcmp(name, "h3dtriadirect")) {
    if (nrhs != 14 || nlhs > 4) mexErrMsgIdAndTxt("fmm3d:usage", "h3dtriadirect: 13 arguments after the name");
    zk[0] = mxGetScalar(prhs[2]);
    if (mxIsComplex(prhs[2])) zk[1] = ((const double *)mxGetComplexDoubles(prhs[2]))[1];
    triangle_call(3, nlhs, plhs, prhs + 3, (int)mxGetScalar(prhs[1]), zk);
  } else {
    mexErrMsgIdAndTxt("fmm3d:usage", "unknown routine %s", name);
  }
}
