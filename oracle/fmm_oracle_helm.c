/*
 * fmm_oracle_helm.c -- CPU oracle pieces for the Helmholtz particle path
 * (src/hfmm3dpart.f, src/helmrouts3d.f).  TEST INFRASTRUCTURE ONLY.
 */
#include "fmm_oracle.h"
#include <math.h>
#include <stdlib.h>

/* helmrouts3d.f:1335 hpotfld3d and :1844 hpotfld3d_dp for one pair:
 * charge  pot = q e^{ikr}/r,            fld = q (1-ikr) e^{ikr} x / r^3
 * dipole  pot = ds (v.x)(1-ikr)e^{ikr}/r^3 (the dipole direction derivative
 *         of the charge kernel), fld = -grad_target pot. */
static inline void helm_pair(zdouble zk, const double *s, int ifcharge, zdouble q, int ifdipole,
                             zdouble ds, const double *dv, const double *t, zdouble *pot, zdouble *fld) {
  double dx = t[0] - s[0], dy = t[1] - s[1], dz = t[2] - s[2];
  double dd = dx * dx + dy * dy + dz * dz, d = sqrt(dd), dinv = 1.0 / d;
  zdouble ikr = I * zk * d;
  zdouble e = cexp(ikr) * dinv;                         /* e^{ikr}/r */
  if (ifcharge) {
    zdouble pq = q * e;
    *pot += pq;
    zdouble cd = (1.0 - ikr) * pq / dd;
    fld[0] += cd * dx; fld[1] += cd * dy; fld[2] += cd * dz;
  }
  if (ifdipole) {
    double dot = dx * dv[0] + dy * dv[1] + dz * dv[2];
    zdouble g = (1.0 - ikr) * e / dd;                   /* (1-ikr)e^{ikr}/r^3 */
    *pot += ds * dot * g;
    zdouble h = e * (zk * zk * dd + 3.0 * ikr - 3.0) / (dd * dd);  /* e^{ikr}(k^2r^2+3ikr-3)/r^5 */
    fld[0] -= ds * (dv[0] * g + dot * dx * h);
    fld[1] -= ds * (dv[1] * g + dot * dy * h);
    fld[2] -= ds * (dv[2] * g + dot * dz * h);
  }
}

/* hfmm3dpart.f:1268-1401 h3dpartdirect */
void orc_h3dpartdirect(zdouble zk, int ns, const double *src, int ifcharge, const zdouble *q,
                       int ifdipole, const zdouble *ds, const double *dv,
                       int ifpot, zdouble *pot, int iffld, zdouble *fld,
                       int nt, const double *trg, int ifpottarg, zdouble *pottarg,
                       int iffldtarg, zdouble *fldtarg) {
  if (ifpot || iffld) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ns; j++) {
      zdouble p = 0, f[3] = {0, 0, 0};
      for (int i = 0; i < ns; i++) {
        if (i == j) continue;
        helm_pair(zk, src + 3 * i, ifcharge, ifcharge ? q[i] : 0, ifdipole, ifdipole ? ds[i] : 0,
                  ifdipole ? dv + 3 * i : NULL, src + 3 * j, &p, f);
      }
      if (ifpot) pot[j] = p;
      if (iffld) for (int k = 0; k < 3; k++) fld[3 * j + k] = f[k];
    }
  }
  if (ifpottarg || iffldtarg) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nt; j++) {
      zdouble p = 0, f[3] = {0, 0, 0};
      for (int i = 0; i < ns; i++)
        helm_pair(zk, src + 3 * i, ifcharge, ifcharge ? q[i] : 0, ifdipole, ifdipole ? ds[i] : 0,
                  ifdipole ? dv + 3 * i : NULL, trg + 3 * j, &p, f);
      if (ifpottarg) pottarg[j] = p;
      if (iffldtarg) for (int k = 0; k < 3; k++) fldtarg[3 * j + k] = f[k];
    }
  }
}

void orc_h3ddirect_sample(zdouble zk, int ns, const double *src, int ifcharge, const zdouble *q,
                          int ifdipole, const zdouble *ds, const double *dv,
                          int nt, const double *trg, const int *self_index,
                          zdouble *pot, zdouble *fld) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < nt; j++) {
    zdouble p = 0, f[3] = {0, 0, 0};
    int skip = self_index ? self_index[j] : -1;
    for (int i = 0; i < ns; i++) {
      if (i == skip) continue;
      helm_pair(zk, src + 3 * i, ifcharge, ifcharge ? q[i] : 0, ifdipole, ifdipole ? ds[i] : 0,
                ifdipole ? dv + 3 * i : NULL, trg + 3 * j, &p, f);
    }
    pot[j] = p;
    for (int k = 0; k < 3; k++) fld[3 * j + k] = f[k];
  }
}
