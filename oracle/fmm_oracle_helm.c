/*
 * fmm_oracle_helm.c -- CPU oracle for the Helmholtz particle path
 * (src/hfmm3dpart.f, src/hfmm3drouts.f, src/helmrouts3d.f, src/h3dterms.f,
 * src/cdjseval3d.f).  TEST INFRASTRUCTURE ONLY.
 *
 * Expansion convention (the reference's, helmrouts3d.f:3119 h3dformmp_trunc0,
 * :3438 h3dformta_trunc0, :2405 h3dmpeval_trunc0, :2732 h3dtaeval_trunc0):
 *   Y_n^m      = sqrt(2n+1) Yhat_n^m   (ylgndr, yrecursion.f:189, Yhat as in
 *                fmm_oracle.h),
 *   multipole  M_n^m = ik sum_j q_j j_n(k r_j)/rscale^n Y_n^{-m}(y_j - c),
 *   local      L_n^m = ik sum_j q_j h_n(k r_j) rscale^n Y_n^{-m}(y_j - c),
 *   evaluated  sum M_n^m h_n(kr) rscale^n Y_n^m  and  sum L_n^m j_n(kr)/rscale^n Y_n^m.
 * Gradients (fields, dipole sources) use the exact derivative identities of
 * f_n(kr) Yhat_n^m (f = j or h, U = d/dx + i d/dy, D = d/dx - i d/dy):
 *   dz f_n Yhat_n^m = k/(2n+1) [sqrt((n+m)(n-m)) f_{n-1} Yhat_{n-1}^m
 *                               - sqrt((n+1+m)(n+1-m)) f_{n+1} Yhat_{n+1}^m]
 *   U  f_n Yhat_n^m = k/(2n+1) [bR(n,m) f_{n-1} Yhat_{n-1}^{m+1} + bS(n,m) f_{n+1} Yhat_{n+1}^{m+1}]
 *   D  f_n Yhat_n^m = k/(2n+1) [cR(n,m) f_{n-1} Yhat_{n-1}^{m-1} + cS(n,m) f_{n+1} Yhat_{n+1}^{m-1}]
 * (the reference differentiates in spherical coordinates instead; the
 * result is the same function).
 *
 * Translations.  The reference shifts expansions by evaluating them on a
 * sphere and projecting with Gauss-Legendre quadrature (h3dtrans.f:211
 * h3dmpmpquad0, :1229 h3dmplocquad_trunc0, :1998 h3dloclocquad0).  This
 * oracle applies the same operators (rotate so the shift is along +z,
 * coaxial shift, rotate back; same truncation orders) with the coaxial
 * coefficients in closed form,
 *   (R|R)_{ln}^m(d) = sum_s 4 pi i^{l+s-n} j_s(kd) sqrt((2s+1)/(4pi)) G(n,s,l;m),
 *   (S|R)_{ln}^m(d) = the same with h_s(kd),
 *   G(n,s,l;m) = int Ynorm_n^m Ynorm_s^0 conj(Ynorm_l^m) dOmega (orthonormal
 *   harmonics), computed exactly by Gauss-Legendre quadrature,
 * so the two differ by the reference's quadrature (aliasing) error, which
 * is far below the FMM tolerance.
 */
#include "fmm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(n, m) ((n) * (n) + (n) + (m))
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* direct sums                                                               */
/* ------------------------------------------------------------------------ */

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

/* ------------------------------------------------------------------------ */
/* spherical Bessel and Hankel functions                                     */
/* ------------------------------------------------------------------------ */

/* helmrouts3d.f:666 h3d01 + :690 h3dall: hs[i] = h_i(z) sc^i, i = 0..n,
 * by the upward recurrence (stable for h); zero for |z| < 1e-15. */
void orc_h3dall(int n, zdouble z, double sc, zdouble *hs) {
  if (cabs(z) < 1.0e-15) {
    for (int i = 0; i <= n; i++) hs[i] = 0;
    return;
  }
  zdouble h0 = cexp(I * z) / (I * z);
  hs[0] = h0;
  if (n >= 1) hs[1] = h0 * (1.0 / z - I) * sc;
  double sc2 = sc * sc;
  zdouble zinv = sc / z;
  for (int i = 1; i < n; i++) {
    zdouble zt = zinv * (2.0 * i + 1.0);
    hs[i + 1] = zt * hs[i] - sc2 * hs[i - 1];
  }
}

/* cdjseval3d.f:15 jfuns3d: fj[i] = j_i(z)/sc^i, i = 0..n, by Miller's
 * backward recurrence started where the forward recurrence passes 1e20 in
 * modulus (search window n..n+1000), rescaled by 1e-40 whenever the
 * backward sweep would overflow, normalised with j_0 or j_1, whichever is
 * larger.  Returns 8 when the window is too short (as the reference). */
int orc_jfuns3d(int n, zdouble z, double sc, zdouble *fj) {
  const double up2 = 1.0e40, up2inv = 1.0e-40;
  if (cabs(z) < 1.0e-200) {
    fj[0] = 1.0;
    for (int i = 1; i <= n; i++) fj[i] = 0;
    return 0;
  }
  int nn = n < 1 ? 1 : n, lw = nn + 1000;
  zdouble *f = malloc(sizeof(zdouble) * (lw + 3));
  int *isc = calloc(lw + 3, sizeof(int));
  zdouble zinv = 1.0 / z;
  f[nn] = 1.0;
  f[nn - 1] = 0.0;
  int ntop = 0;
  for (int i = nn; i <= lw; i++) {
    zdouble zt = (2.0 * i + 1.0) * zinv * f[i] - f[i - 1];
    f[i + 1] = zt;
    if (creal(zt) * creal(zt) + cimag(zt) * cimag(zt) > up2) {
      ntop = i + 1;
      break;
    }
  }
  if (ntop == 0) {
    free(f);
    free(isc);
    return 8;
  }
  f[ntop] = 0.0;
  f[ntop - 1] = 1.0;
  for (int i = ntop - 1; i >= 1; i--) {
    zdouble zt = (2.0 * i + 1.0) * zinv * f[i] - f[i + 1];
    f[i - 1] = zt;
    if (creal(zt) * creal(zt) + cimag(zt) * cimag(zt) > up2) {
      f[i] *= up2inv;
      f[i - 1] *= up2inv;
      isc[i] = 1;
    }
  }
  double scinv = 1.0 / sc, sctot = 1.0;
  for (int i = 1; i <= ntop; i++) {
    sctot *= scinv;
    if (isc[i - 1]) sctot *= up2inv;
    f[i] *= sctot;
  }
  zdouble fj0 = csin(z) * zinv, fj1 = fj0 * zinv - ccos(z) * zinv;
  zdouble zs = (cabs(fj1) > cabs(fj0)) ? fj1 / (f[1] * sc) : fj0 / f[0];
  for (int i = 0; i <= n; i++) fj[i] = f[i] * zs;
  free(f);
  free(isc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* expansion orders (h3dterms.f)                                             */
/* ------------------------------------------------------------------------ */

/* Smallest j >= 2 with |j_j(z2) h_j(z1)| + |j_{j-1} h_{j-1}| < eps (|j_0 h_0| +
 * |j_1 h_1|), returning j+1 (1000 and ier = 13 when none up to 1000). */
static int h_terms_decay(zdouble z1, zdouble z2, double sc, double eps, int *ier) {
  const int ntmax = 1000;
  zdouble *hf = malloc(sizeof(zdouble) * (ntmax + 1)), *jf = malloc(sizeof(zdouble) * (ntmax + 1));
  orc_h3dall(ntmax, z1, sc, hf);
  int r = 1000;
  if (orc_jfuns3d(ntmax, z2, sc, jf) != 0) {
    *ier = 11;
    free(hf);
    free(jf);
    return 1;
  }
  double x0 = cabs(jf[0] * hf[0]) + cabs(jf[1] * hf[1]);
  *ier = 13;
  for (int j = 2; j <= ntmax; j++) {
    double x = cabs(jf[j] * hf[j]) + cabs(jf[j - 1] * hf[j - 1]);
    if (x < eps * x0) {
      r = j + 1;
      *ier = 0;
      break;
    }
  }
  free(hf);
  free(jf);
  return r;
}

/* h3dterms.f:28 h3dterms */
int orc_h3dterms(double size, zdouble zk, double eps, int *ier) {
  double sc = cabs(zk * size) < 1.0 ? cabs(zk * size) : 1.0;
  return h_terms_decay(zk * size * 1.5, zk * size * sqrt(3.0) / 2.0, sc, eps, ier);
}

/* h3dterms.f:192 h3dterms_list2 (same offsets and symmetrisation as the
 * Laplace table, l3dterms.f:136) */
void orc_h3dterms_list2(double size, zdouble zk, double eps, int *itable, int *ier) {
  int nt[2][4][4];
  double sc = cabs(zk * size) < 1.0 ? cabs(zk * size) : 1.0;
  *ier = 0;
  for (int ii = 2; ii <= 3; ii++)
    for (int jj = 0; jj <= 3; jj++)
      for (int kk = 0; kk <= 3; kk++) {
        double dx = ii, dy = jj, dz = kk;
        if (dx > 0) dx -= 0.5;
        if (dy > 0) dy -= 0.5;
        if (dz > 0) dz -= 0.5;
        double rr = sqrt(dx * dx + dy * dy + dz * dz);
        int e;
        nt[ii - 2][jj][kk] = h_terms_decay(zk * size * rr, zk * size * sqrt(3.0) / 2.0, sc, eps, &e);
        if (e) *ier = e;
      }
  for (int i = -3; i <= 3; i++)
    for (int j = -3; j <= 3; j++)
      for (int k = -3; k <= 3; k++) {
        int v = 0, ai = abs(i), aj = abs(j), ak = abs(k);
        if (ai > 1) v = nt[ai - 2][aj][ak];
        else if (aj > 1) v = nt[aj - 2][ai][ak];
        if (ai <= 1 && aj <= 1 && ak > 1) v = (ai >= aj) ? nt[ak - 2][ai][aj] : nt[ak - 2][aj][ai];
        itable[((i + 3) * 7 + (j + 3)) * 7 + (k + 3)] = v;
      }
}

/* ------------------------------------------------------------------------ */
/* angular functions                                                         */
/* ------------------------------------------------------------------------ */

/* yrecursion.f:189 ylgndr: y[n*(p+1)+m] = sqrt(2n+1) (-1)^m
 * sqrt((n-m)!/(n+m)!) P_n^m(x), 0 <= m <= n <= p. */
static void ylgndr(int p, double x, double *y) {
  int P = p + 1;
  double u = -sqrt((1.0 - x) * (1.0 + x));
  y[0] = 1.0;
  for (int m = 0; m <= p; m++) {
    if (m > 0) y[m * P + m] = y[(m - 1) * P + m - 1] * u * sqrt((2.0 * m - 1.0) / (2.0 * m));
    if (m < p) y[(m + 1) * P + m] = x * y[m * P + m] * sqrt(2.0 * m + 1.0);
    for (int n = m + 2; n <= p; n++)
      y[n * P + m] = ((2.0 * n - 1.0) * x * y[(n - 1) * P + m] -
                      sqrt((n + m - 1.0) * (n - m - 1.0)) * y[(n - 2) * P + m]) /
                     sqrt((double)(n - m) * (n + m));
  }
  for (int n = 0; n <= p; n++)
    for (int m = 0; m <= n; m++) y[n * P + m] *= sqrt(2.0 * n + 1.0);
}

/* helmrouts3d.f:627 cart2polar */
static void h_cart2polar(const double *v, double *r, double *theta, double *phi) {
  *r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  *theta = atan2(sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
  *phi = (v[0] == 0.0 && v[1] == 0.0) ? 0.0 : atan2(v[1], v[0]);
}

/* Yhat_n^m(direction of v) for n <= p, all m, packed, plus |v|. */
static double yhat_packed(int p, const double *v, zdouble *Y) {
  double r, th, ph;
  h_cart2polar(v, &r, &th, &ph);
  int P = p + 1;
  double *y = malloc(sizeof(double) * P * P);
  ylgndr(p, cos(th), y);
  zdouble e1 = cos(ph) + I * sin(ph), em = 1.0;
  for (int m = 0; m <= p; m++) {
    if (m > 0) em *= e1;
    for (int n = m; n <= p; n++) {
      double a = y[n * P + m] / sqrt(2.0 * n + 1.0);
      Y[IDX(n, m)] = a * em;
      if (m) Y[IDX(n, -m)] = a * conj(em);
    }
  }
  free(y);
  return r;
}

static inline double dA1(int n, int m) { return sqrt((double)(n + m) * (n - m)); }
static inline double dA2(int n, int m) { return sqrt((double)(n + 1 + m) * (n + 1 - m)); }
static inline double dbR(int n, int m) { return sqrt((double)(n - m) * (n - m - 1)) * (m >= 0 ? 1.0 : -1.0); }
static inline double dcR(int n, int m) { return sqrt((double)(n + m) * (n + m - 1)) * (m >= 1 ? -1.0 : 1.0); }
static inline double dbS(int n, int m) { return sqrt((double)(n + m + 1) * (n + m + 2)) * (m >= 0 ? 1.0 : -1.0); }
static inline double dcS(int n, int m) { return sqrt((double)(n - m + 1) * (n - m + 2)) * (m >= 1 ? -1.0 : 1.0); }

/* F(n,m) = f[n] Y[n,m] when (n,m) is a valid index, else 0 */
static inline zdouble fy(int pmax, const zdouble *f, const zdouble *Y, int n, int m) {
  if (n < 0 || n > pmax || m > n || m < -n) return 0;
  return f[n] * Y[IDX(n, m)];
}

/* Directional derivative w.(dz, U, D) of f_n Yhat_n^mu with the identities in
 * the header; lo/hi multiply the f_{n-1} / f_{n+1} terms (scale factors). */
static zdouble dir_deriv(zdouble zk, int pmax, const zdouble *f, const zdouble *Y, int n, int mu,
                         zdouble wz, zdouble wU, zdouble wD, double lo, double hi) {
  zdouble a = wz * (dA1(n, mu) * lo * fy(pmax, f, Y, n - 1, mu) - dA2(n, mu) * hi * fy(pmax, f, Y, n + 1, mu));
  zdouble b = wU * ((n >= 1 ? dbR(n, mu) : 0.0) * lo * fy(pmax, f, Y, n - 1, mu + 1) +
                    dbS(n, mu) * hi * fy(pmax, f, Y, n + 1, mu + 1));
  zdouble c = wD * ((n >= 1 ? dcR(n, mu) : 0.0) * lo * fy(pmax, f, Y, n - 1, mu - 1) +
                    dcS(n, mu) * hi * fy(pmax, f, Y, n + 1, mu - 1));
  return zk / (2.0 * n + 1.0) * (a + b + c);
}

/* ------------------------------------------------------------------------ */
/* formation and evaluation                                                  */
/* ------------------------------------------------------------------------ */

/* helmrouts3d.f:2972 h3dformmp_add_trunc (charges) and :3580
 * h3dformmp_dp_add_trunc (dipoles); IRREG selects h3dformta_add_trunc
 * (:3299) and h3dformta_dp_add_trunc (:3978).  Adds into `out`. */
static void h_form(int irreg, zdouble zk, double sc, const double *src, const zdouble *q,
                   const zdouble *ds, const double *dv, int ns, const double *c, int p, zdouble *out) {
  int K = (p + 1) * (p + 1), K1 = (p + 2) * (p + 2);
  zdouble *acc = calloc(K, sizeof(zdouble)), *Y = malloc(sizeof(zdouble) * K1);
  zdouble *f = malloc(sizeof(zdouble) * (p + 3));
  for (int i = 0; i < ns; i++) {
    double v[3] = {src[3 * i] - c[0], src[3 * i + 1] - c[1], src[3 * i + 2] - c[2]};
    double r = yhat_packed(p + 1, v, Y);
    zdouble z = zk * r;
    if (irreg) {
      if (cabs(z) < 1.0e-15) continue;       /* h3dformta_trunc0 returns ier=4 */
      orc_h3dall(p + 1, z, sc, f);
    } else {
      orc_jfuns3d(p + 1, z, sc, f);
    }
    for (int n = 0; n <= p; n++) {
      double w = sqrt(2.0 * n + 1.0);
      for (int m = -n; m <= n; m++) {
        zdouble t = 0;
        if (q) t += q[i] * f[n] * Y[IDX(n, -m)];
        if (ds) {
          const double *u = dv + 3 * i;
          zdouble wU = 0.5 * (u[0] - I * u[1]), wD = 0.5 * (u[0] + I * u[1]);
          /* scaled f: regular j_n/sc^n -> f_{n-1} term * 1/sc, f_{n+1} * sc;
           * irregular h_n sc^n -> the reverse */
          double lo = irreg ? sc : 1.0 / sc, hi = irreg ? 1.0 / sc : sc;
          t += ds[i] * dir_deriv(zk, p + 1, f, Y, n, -m, u[2], wU, wD, lo, hi);
        }
        acc[IDX(n, m)] += w * t;
      }
    }
  }
  for (int k = 0; k < K; k++) out[k] += acc[k] * I * zk;
  free(acc);
  free(Y);
  free(f);
}

void orc_h3dformmp(zdouble zk, double sc, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *mp) {
  h_form(0, zk, sc, src, q, NULL, NULL, ns, center, p, mp);
}
void orc_h3dformmp_dp(zdouble zk, double sc, const double *src, const zdouble *ds, const double *dv,
                      int ns, const double *center, int p, zdouble *mp) {
  h_form(0, zk, sc, src, NULL, ds, dv, ns, center, p, mp);
}
void orc_h3dformta(zdouble zk, double sc, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *loc) {
  h_form(1, zk, sc, src, q, NULL, NULL, ns, center, p, loc);
}
void orc_h3dformta_dp(zdouble zk, double sc, const double *src, const zdouble *ds, const double *dv,
                      int ns, const double *center, int p, zdouble *loc) {
  h_form(1, zk, sc, src, NULL, ds, dv, ns, center, p, loc);
}

/* helmrouts3d.f:2211 h3dmpevalall_trunc (IRREG=1: multipole, h_n) and
 * :2550 h3dtaevalall_trunc (local, j_n).  Adds pot and fld = -grad pot. */
static void h_eval(int irreg, zdouble zk, double sc, const double *c, const zdouble *ex, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld) {
  int K1 = (p + 2) * (p + 2);
  zdouble *Y = malloc(sizeof(zdouble) * K1), *f = malloc(sizeof(zdouble) * (p + 3));
  for (int i = 0; i < nt; i++) {
    double v[3] = {trg[3 * i] - c[0], trg[3 * i + 1] - c[1], trg[3 * i + 2] - c[2]};
    double r = yhat_packed(p + 1, v, Y);
    zdouble z = zk * r;
    if (irreg) {
      if (cabs(z) < 1.0e-15) continue;       /* h3dmpeval_trunc0 returns ier=8 */
      orc_h3dall(p + 1, z, sc, f);
    } else {
      orc_jfuns3d(p + 1, z, sc, f);
    }
    zdouble ph = 0, gz = 0, gU = 0, gD = 0;
    double lo = irreg ? sc : 1.0 / sc, hi = irreg ? 1.0 / sc : sc;
    for (int n = 0; n <= p; n++) {
      double w = sqrt(2.0 * n + 1.0);
      for (int m = -n; m <= n; m++) {
        zdouble cf = ex[IDX(n, m)] * w;
        ph += cf * f[n] * Y[IDX(n, m)];
        if (iffld) {
          gz += cf * dir_deriv(zk, p + 1, f, Y, n, m, 1.0, 0.0, 0.0, lo, hi);
          gU += cf * dir_deriv(zk, p + 1, f, Y, n, m, 0.0, 1.0, 0.0, lo, hi);
          gD += cf * dir_deriv(zk, p + 1, f, Y, n, m, 0.0, 0.0, 1.0, lo, hi);
        }
      }
    }
    if (ifpot) pot[i] += ph;
    if (iffld) {
      fld[3 * i + 0] -= 0.5 * (gU + gD);
      fld[3 * i + 1] -= (gU - gD) / (2.0 * I);
      fld[3 * i + 2] -= gz;
    }
  }
  free(Y);
  free(f);
}

void orc_h3dmpeval(zdouble zk, double sc, const double *center, const zdouble *mp, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld) {
  h_eval(1, zk, sc, center, mp, p, trg, nt, ifpot, pot, iffld, fld);
}
void orc_h3dtaeval(zdouble zk, double sc, const double *center, const zdouble *loc, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld) {
  h_eval(0, zk, sc, center, loc, p, trg, nt, ifpot, pot, iffld, fld);
}

/* ------------------------------------------------------------------------ */
/* coaxial translations                                                      */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre nodes and weights on [-1,1] (Newton on P_q). */
static void gauleg(int q, double *x, double *w) {
  for (int i = 0; i < q; i++) {
    double t = cos(M_PI * (i + 0.75) / (q + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= q; k++) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (q == 1) { p1 = t; p0 = 1.0; }
      dp = q * (t * p1 - p0) / (t * t - 1.0);
      double dt = p1 / dp;
      t -= dt;
      if (fabs(dt) < 1e-16) break;
    }
    x[i] = t;
    w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

/* Gaunt integrals G(n,s,l;m), 0 <= m <= min(n,l) <= max(n,l) <= P,
 * s = |n-l|, |n-l|+2, ..., n+l, stored compactly. */
typedef struct {
  int P;
  long *off;      /* [(m*(P+1)+n)*(P+1)+l] -> start of the s run */
  double *g;
} gaunt_t;

static gaunt_t *gaunt_build(int P) {
  gaunt_t *G = malloc(sizeof(gaunt_t));
  G->P = P;
  int P1 = P + 1, S = 2 * P + 1, Q = 2 * P + 2;
  G->off = malloc(sizeof(long) * P1 * P1 * P1);
  long tot = 0;
  for (int m = 0; m <= P; m++)
    for (int n = 0; n <= P; n++)
      for (int l = 0; l <= P; l++) {
        G->off[(m * P1 + n) * P1 + l] = tot;
        if (n >= m && l >= m) tot += (n < l ? n : l) + 1;
      }
  G->g = calloc(tot > 0 ? tot : 1, sizeof(double));
  double *x = malloc(sizeof(double) * Q), *w = malloc(sizeof(double) * Q);
  gauleg(Q, x, w);
  /* Theta[q][n][m] = ylgndr / sqrt(4 pi): orthonormal up to the phase e^{im phi} */
  double *th = malloc(sizeof(double) * Q * (S + 1) * (S + 1));
  for (int q = 0; q < Q; q++) {
    ylgndr(S, x[q], th + (size_t)q * (S + 1) * (S + 1));
    for (int k = 0; k < (S + 1) * (S + 1); k++) th[(size_t)q * (S + 1) * (S + 1) + k] /= sqrt(4.0 * M_PI);
  }
#define TH(q, n, m) th[(size_t)(q) * (S + 1) * (S + 1) + (n) * (S + 1) + (m)]
#pragma omp parallel for schedule(dynamic, 1)
  for (int m = 0; m <= P; m++) {
    double *prod = malloc(sizeof(double) * Q);
    for (int n = m; n <= P; n++)
      for (int l = m; l <= P; l++) {
        for (int q = 0; q < Q; q++) prod[q] = 2.0 * M_PI * w[q] * TH(q, n, m) * TH(q, l, m);
        long o = G->off[(m * P1 + n) * P1 + l];
        int s0 = abs(n - l);
        for (int s = s0, k = 0; s <= n + l; s += 2, k++) {
          double a = 0;
          for (int q = 0; q < Q; q++) a += prod[q] * TH(q, s, 0);
          G->g[o + k] = a;
        }
      }
    free(prod);
  }
#undef TH
  free(th);
  free(x);
  free(w);
  return G;
}

static void gaunt_free(gaunt_t *G) {
  free(G->off);
  free(G->g);
  free(G);
}

/* One coaxial operator along +z by distance d: T[m] is (pout+1-m) x (pin+1-m),
 * row l, column n, acting on reference-convention coefficients.
 * mode 0 = M2M (R|R), 1 = M2L (S|R), 2 = L2L (R|R). */
typedef struct {
  int pin, pout;
  zdouble **T;
} coax_t;

static coax_t *coax_build(const gaunt_t *G, int mode, zdouble zk, double d, double sc_in,
                          double sc_out, int pin, int pout) {
  coax_t *C = malloc(sizeof(coax_t));
  C->pin = pin;
  C->pout = pout;
  int S = pin + pout, P1 = G->P + 1;
  C->T = malloc(sizeof(zdouble *) * (pin < pout ? pout + 1 : pin + 1));
  zdouble *f = malloc(sizeof(zdouble) * (S + 2));
  /* scaled functions: M2L uses h_s sc^s (sc = sc_in = sc_out), M2M j_s/sc_out^s,
   * L2L j_s/sc_in^s; the remaining powers are non-negative (no overflow). */
  double base = (mode == 1) ? sc_in : (mode == 0 ? sc_out : sc_in);
  if (mode == 1) orc_h3dall(S, zk * d, base, f);
  else orc_jfuns3d(S, zk * d, base, f);
  int mtop = pin < pout ? pin : pout;
  for (int m = 0; m <= mtop; m++) {
    int R = pout + 1 - m, Cn = pin + 1 - m;
    zdouble *t = calloc((size_t)R * Cn, sizeof(zdouble));
    for (int l = m; l <= pout; l++)
      for (int n = m; n <= pin; n++) {
        long o = G->off[(m * P1 + n) * P1 + l];
        zdouble acc = 0;
        for (int s = abs(n - l), k = 0; s <= n + l; s += 2, k++) {
          double sg = (((l + s - n) / 2) & 1) ? -1.0 : 1.0;     /* i^{l+s-n} */
          double c = 4.0 * M_PI * sg * sqrt((2.0 * s + 1.0) / (4.0 * M_PI)) * G->g[o + k];
          double pw;
          if (mode == 1) pw = pow(base, n + l - s);
          else if (mode == 0) pw = pow(base, s - l + n) * pow(sc_in / sc_out, n);
          else pw = pow(base, s - n + l) * pow(sc_out / sc_in, l);
          acc += c * pw * f[s];
        }
        t[(size_t)(l - m) * Cn + (n - m)] = acc;
      }
    C->T[m] = t;
  }
  for (int m = mtop + 1; m <= (pin < pout ? pout : pin); m++) C->T[m] = NULL;
  free(f);
  return C;
}

static void coax_free(coax_t *C) {
  int mtop = C->pin < C->pout ? C->pin : C->pout;
  for (int m = 0; m <= mtop; m++) free(C->T[m]);
  free(C->T);
  free(C);
}

/* Point-and-shoot with a coaxial operator: rotate the shift vector (to -
 * from) onto +z, apply C (truncated to pin/pout <= its own orders), rotate
 * back, add into out. */
static void h_shift(const double *from, const double *to, const coax_t *C, int pin,
                    const zdouble *in, int pout, zdouble *out) {
  double v[3] = {to[0] - from[0], to[1] - from[1], to[2] - from[2]}, rho, th, ph;
  h_cart2polar(v, &rho, &th, &ph);
  int pm = pin > pout ? pin : pout, K = (pm + 1) * (pm + 1);
  zdouble *a = calloc(K, sizeof(zdouble)), *b = calloc(K, sizeof(zdouble));
  for (int n = 0; n <= pin; n++)
    for (int m = -n; m <= n; m++) a[IDX(n, m)] = in[IDX(n, m)] * cexp(I * m * ph);
  orc_rotate_y(pin, th, a, b);
  memset(a, 0, sizeof(zdouble) * K);
  int Cn = C->pin + 1;
  for (int m = -pout; m <= pout; m++) {
    int am = abs(m);
    if (am > pin) continue;
    const zdouble *t = C->T[am];
    for (int l = am; l <= pout; l++) {
      zdouble acc = 0;
      for (int n = am; n <= pin; n++) acc += t[(size_t)(l - am) * (Cn - am) + (n - am)] * b[IDX(n, m)];
      a[IDX(l, m)] = acc;
    }
  }
  orc_rotate_y(pout, -th, a, b);
  for (int n = 0; n <= pout; n++)
    for (int m = -n; m <= n; m++) out[IDX(n, m)] += b[IDX(n, m)] * cexp(-I * m * ph);
  free(a);
  free(b);
}

static void h_translate(int mode, zdouble zk, double sc1, const double *c1, const zdouble *in, int p1,
                        double sc2, const double *c2, zdouble *out, int p2) {
  int P = p1 > p2 ? p1 : p2;
  gaunt_t *G = gaunt_build(P);
  double v[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
  double d = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  coax_t *C = coax_build(G, mode, zk, d, sc1, sc2, p1, p2);
  h_shift(c1, c2, C, p1, in, p2, out);
  coax_free(C);
  gaunt_free(G);
}

/* h3dtrans.f:120 h3dmpmpquadu_add: multipole at c1 (order p1, scale sc1)
 * shifted to c2 (order p2, scale sc2), added. */
void orc_h3dmpmp(zdouble zk, double sc1, const double *c1, const zdouble *mp1, int p1,
                 double sc2, const double *c2, zdouble *mp2, int p2) {
  h_translate(0, zk, sc1, c1, mp1, p1, sc2, c2, mp2, p2);
}

/* h3dtrans.f:1151 h3dmplocquadu_add_trunc: multipole at c1 to a local at c2,
 * input and output truncated to ptrunc (hfmm3drouts.f:417-450). */
void orc_h3dmploc(zdouble zk, double sc1, const double *c1, const zdouble *mp, int p1, int ptrunc,
                  double sc2, const double *c2, zdouble *loc, int p2) {
  (void)p1; (void)p2;
  h_translate(1, zk, sc1, c1, mp, ptrunc, sc2, c2, loc, ptrunc);
}

/* h3dtrans.f:1911 h3dloclocquadu_add */
void orc_h3dlocloc(zdouble zk, double sc1, const double *c1, const zdouble *loc1, int p1,
                   double sc2, const double *c2, zdouble *loc2, int p2) {
  h_translate(2, zk, sc1, c1, loc1, p1, sc2, c2, loc2, p2);
}

/* ------------------------------------------------------------------------ */
/* FMM driver                                                                */
/* ------------------------------------------------------------------------ */

/* hfmm3dpart.f:1092-1262 hfmm3dpart_direct_self/_targ */
static void h_p2p_block(zdouble zk, const double *src, int ifcharge, const zdouble *q, int ifdipole,
                        const zdouble *ds, const double *dv, int s0, int sn, int t0, int tn,
                        int ifpot, zdouble *pot, int iffld, zdouble *fld, const double *trg,
                        int u0, int un, int ifpottarg, zdouble *pottarg, int iffldtarg,
                        zdouble *fldtarg) {
  for (int pass = 0; pass < 2; pass++) {
    int b0 = pass ? u0 : t0, bn = pass ? un : tn;
    int wp = pass ? ifpottarg : ifpot, wf = pass ? iffldtarg : iffld;
    if (!wp && !wf) continue;
    const double *Pp = pass ? trg : src;
    zdouble *po = pass ? pottarg : pot, *fo = pass ? fldtarg : fld;
    for (int j = b0; j < b0 + bn; j++) {
      zdouble ph = 0, f[3] = {0, 0, 0};
      for (int i = s0; i < s0 + sn; i++) {
        if (!pass && i == j) continue;
        helm_pair(zk, src + 3 * i, ifcharge, ifcharge ? q[i] : 0, ifdipole, ifdipole ? ds[i] : 0,
                  ifdipole ? dv + 3 * i : NULL, Pp + 3 * j, &ph, f);
      }
      if (wp) po[j] += ph;
      if (wf) for (int k = 0; k < 3; k++) fo[3 * j + k] += f[k];
    }
  }
}

int orc_iprec_params(int iprec, int ns, int nt, double *eps, int *nbox);

/* hfmm3dpart.f:141-539 hfmm3dparttarg with hfmm3dparttargmain (:545-1086)
 * and hfmm3d_list2 (hfmm3drouts.f:153-564). */
int orc_hfmm3dparttarg(int iprec, zdouble zk, int ns, const double *src_in,
                       int ifcharge, const zdouble *q_in, int ifdipole, const zdouble *ds_in,
                       const double *dv_in, int ifpot, zdouble *pot_out, int iffld, zdouble *fld_out,
                       int nt, const double *trg_in, int ifpottarg, zdouble *pottarg_out,
                       int iffldtarg, zdouble *fldtarg_out) {
  double eps;
  int nbox;
  if (!orc_iprec_params(iprec, ns, nt, &eps, &nbox)) return 32;
  if (nt < 0) nt = 0;
  orc_tree t;
  if (ns < 1 || orc_tree_build(src_in, ns, trg_in, nt, nbox, &t) != 0) return 4;
  const int nb = t.nboxes, nlev = t.nlev;
  const orc_box *B = t.boxes;

  double *src = malloc(sizeof(double) * 3 * ns), *trg = malloc(sizeof(double) * 3 * (nt + 1));
  zdouble *q = calloc(ns, sizeof(zdouble)), *ds = calloc(ns, sizeof(zdouble));
  double *dv = calloc(3 * ns, sizeof(double));
  for (int i = 0; i < ns; i++) {
    int o = t.isrc[i];
    memcpy(src + 3 * i, src_in + 3 * o, sizeof(double) * 3);
    if (ifcharge) q[i] = q_in[o];
    if (ifdipole) { ds[i] = ds_in[o]; memcpy(dv + 3 * i, dv_in + 3 * o, sizeof(double) * 3); }
  }
  for (int i = 0; i < nt; i++) memcpy(trg + 3 * i, trg_in + 3 * t.itrg[i], sizeof(double) * 3);
  zdouble *pot = calloc(ns, sizeof(zdouble)), *fld = calloc(3 * ns, sizeof(zdouble));
  zdouble *pottarg = calloc(nt + 1, sizeof(zdouble)), *fldtarg = calloc(3 * (nt + 1), sizeof(zdouble));

  /* per-level scale and orders (hfmm3dpart.f:337-341, :401-406) */
  double *scale = malloc(sizeof(double) * (nlev + 1)), *bsize = malloc(sizeof(double) * (nlev + 1));
  int *pl = malloc(sizeof(int) * (nlev + 1));
  int pmax = 0;
  for (int l = 0; l <= nlev; l++) {
    bsize[l] = t.size / pow(2.0, l);
    double bs = cabs(bsize[l] * zk);
    scale[l] = bs < 1.0 ? bs : 1.0;
    int e;
    pl[l] = (l >= 2) ? orc_h3dterms(bsize[l], zk, eps, &e) : 0;
    if (pl[l] > pmax) pmax = pl[l];
  }
  long *eoff = malloc(sizeof(long) * (nb + 1));
  eoff[0] = 0;
  for (int b = 0; b < nb; b++) eoff[b + 1] = eoff[b] + (long)(pl[B[b].level] + 1) * (pl[B[b].level] + 1);
  zdouble *mp = calloc(eoff[nb] + 1, sizeof(zdouble)), *lo = calloc(eoff[nb] + 1, sizeof(zdouble));

  /* STEP 1: P2M at leaves of level >= 2 with sources (hfmm3dpart.f:681-757) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    int l = B[b].level;
    if (l < 2 || B[b].nchild || B[b].nsrc == 0) continue;
    if (ifcharge) orc_h3dformmp(zk, scale[l], src + 3 * B[b].isrc, q + B[b].isrc, B[b].nsrc, B[b].center, pl[l], mp + eoff[b]);
    if (ifdipole) orc_h3dformmp_dp(zk, scale[l], src + 3 * B[b].isrc, ds + B[b].isrc, dv + 3 * B[b].isrc, B[b].nsrc, B[b].center, pl[l], mp + eoff[b]);
  }
  /* STEP 2: P2L, sources of a box into the locals of its list-3 boxes
   * (hfmm3dpart.f:765-830), looped here over the receiving box (list 4) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    int l = B[b].level;
    for (int i = t.lst_off[3][b]; i < t.lst_off[3][b + 1]; i++) {
      int j = t.lst[3][i];
      if (B[j].nsrc == 0) continue;
      if (ifcharge) orc_h3dformta(zk, scale[l], src + 3 * B[j].isrc, q + B[j].isrc, B[j].nsrc, B[b].center, pl[l], lo + eoff[b]);
      if (ifdipole) orc_h3dformta_dp(zk, scale[l], src + 3 * B[j].isrc, ds + B[j].isrc, dv + 3 * B[j].isrc, B[j].nsrc, B[b].center, pl[l], lo + eoff[b]);
    }
  }
  gaunt_t *G = gaunt_build(pmax > 0 ? pmax : 1);
  /* STEP 3: M2M into parents of level >= 2 (hfmm3drouts.f:193-266) */
  for (int lev = nlev - 1; lev >= 2; lev--) {
    double d = bsize[lev + 1] * sqrt(3.0) / 2.0;
    coax_t *C = coax_build(G, 0, zk, d, scale[lev + 1], scale[lev], pl[lev + 1], pl[lev]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      if (B[b].nsrc == 0 || B[b].nchild == 0) continue;
      memset(mp + eoff[b], 0, sizeof(zdouble) * (eoff[b + 1] - eoff[b]));
      for (int k = 0; k < B[b].nchild; k++) {
        int c = B[b].child[k];
        h_shift(B[c].center, B[b].center, C, pl[lev + 1], mp + eoff[c], pl[lev], mp + eoff[b]);
      }
    }
    coax_free(C);
  }
  /* STEP 4: M2L over list 2, levels 2..nlev (hfmm3drouts.f:330-477) */
  int prune = !(ifpot || iffld);
  for (int lev = 2; lev <= nlev; lev++) {
    int itable[343], e;
    orc_h3dterms_list2(bsize[lev], zk, eps, itable, &e);
    /* one coaxial operator per distinct offset length, at the largest order
     * any offset of that length needs (smaller orders use a sub-block) */
    coax_t *Cd[28] = {0};
    int need[28] = {0};
    for (int o = 0; o < 343; o++) {
      int di = o / 49 - 3, dj = (o / 7) % 7 - 3, dk = o % 7 - 3, d2 = di * di + dj * dj + dk * dk;
      int pt = itable[o] < pl[lev] ? itable[o] : pl[lev];
      if (d2 >= 4 && pt > need[d2]) need[d2] = pt;
    }
    for (int d2 = 4; d2 < 28; d2++)
      if (need[d2] > 0) Cd[d2] = coax_build(G, 1, zk, bsize[lev] * sqrt((double)d2), scale[lev], scale[lev], need[d2], need[d2]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      for (int i = t.lst_off[1][b]; i < t.lst_off[1][b + 1]; i++) {
        int j = t.lst[1][i];
        if (B[j].nsrc == 0) continue;
        if (B[b].ntrg == 0 && prune) continue;
        int di = B[j].ijk[0] - B[b].ijk[0], dj = B[j].ijk[1] - B[b].ijk[1], dk = B[j].ijk[2] - B[b].ijk[2];
        int pt = itable[((di + 3) * 7 + (dj + 3)) * 7 + (dk + 3)];
        if (pt > pl[lev]) pt = pl[lev];
        int d2 = di * di + dj * dj + dk * dk;
        h_shift(B[j].center, B[b].center, Cd[d2], pt, mp + eoff[j], pt, lo + eoff[b]);
      }
    }
    for (int d2 = 0; d2 < 28; d2++)
      if (Cd[d2]) coax_free(Cd[d2]);
  }
  /* STEP 5: L2L, parents at levels 2..nlev-1 (hfmm3drouts.f:483-556) */
  for (int lev = 2; lev <= nlev - 1; lev++) {
    double d = bsize[lev + 1] * sqrt(3.0) / 2.0;
    coax_t *C = coax_build(G, 2, zk, d, scale[lev], scale[lev + 1], pl[lev], pl[lev + 1]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      for (int k = 0; k < B[b].nchild; k++) {
        int c = B[b].child[k];
        h_shift(B[b].center, B[c].center, C, pl[lev], lo + eoff[b], pl[lev + 1], lo + eoff[c]);
      }
    }
    coax_free(C);
  }
  gaunt_free(G);
  /* STEP 6: M2P, the multipole of a box at the points of its list-4 boxes
   * (hfmm3dpart.f:851-921), looped here over the receiving box (list 3) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    for (int i = t.lst_off[2][b]; i < t.lst_off[2][b + 1]; i++) {
      int j = t.lst[2][i], l = B[j].level;
      if (B[j].nsrc == 0) continue;
      orc_h3dmpeval(zk, scale[l], B[j].center, mp + eoff[j], pl[l], src + 3 * B[b].isrc, B[b].nsrc,
                    ifpot, pot + B[b].isrc, iffld, fld + 3 * B[b].isrc);
      orc_h3dmpeval(zk, scale[l], B[j].center, mp + eoff[j], pl[l], trg + 3 * B[b].itrg, B[b].ntrg,
                    ifpottarg, pottarg + B[b].itrg, iffldtarg, fldtarg + 3 * B[b].itrg);
    }
  }
  /* STEP 7: L2P at leaves of level >= 2 (hfmm3dpart.f:926-976) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    int l = B[b].level;
    if (B[b].nchild || l < 2) continue;
    orc_h3dtaeval(zk, scale[l], B[b].center, lo + eoff[b], pl[l], src + 3 * B[b].isrc, B[b].nsrc,
                  ifpot, pot + B[b].isrc, iffld, fld + 3 * B[b].isrc);
    orc_h3dtaeval(zk, scale[l], B[b].center, lo + eoff[b], pl[l], trg + 3 * B[b].itrg, B[b].ntrg,
                  ifpottarg, pottarg + B[b].itrg, iffldtarg, fldtarg + 3 * B[b].itrg);
  }
  /* STEP 8: P2P, self + list 1 (hfmm3dpart.f:981-1038) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    if (B[b].nchild) continue;
    if (B[b].nsrc > 0)
      h_p2p_block(zk, src, ifcharge, q, ifdipole, ds, dv, B[b].isrc, B[b].nsrc, B[b].isrc, B[b].nsrc,
                  ifpot, pot, iffld, fld, trg, B[b].itrg, B[b].ntrg, ifpottarg, pottarg, iffldtarg, fldtarg);
    for (int i = t.lst_off[0][b]; i < t.lst_off[0][b + 1]; i++) {
      int j = t.lst[0][i];
      if (B[j].nsrc == 0) continue;
      h_p2p_block(zk, src, ifcharge, q, ifdipole, ds, dv, B[j].isrc, B[j].nsrc, B[b].isrc, B[b].nsrc,
                  ifpot, pot, iffld, fld, trg, B[b].itrg, B[b].ntrg, ifpottarg, pottarg, iffldtarg, fldtarg);
    }
  }
  /* h3dpsort / h3dfsort (hfmm3drouts.f:627-698) */
  for (int i = 0; i < ns; i++) {
    int o = t.isrc[i];
    if (ifpot) pot_out[o] = pot[i];
    if (iffld) for (int k = 0; k < 3; k++) fld_out[3 * o + k] = fld[3 * i + k];
  }
  for (int i = 0; i < nt; i++) {
    int o = t.itrg[i];
    if (ifpottarg) pottarg_out[o] = pottarg[i];
    if (iffldtarg) for (int k = 0; k < 3; k++) fldtarg_out[3 * o + k] = fldtarg[3 * i + k];
  }
  free(src); free(trg); free(q); free(ds); free(dv); free(pot); free(fld); free(pottarg); free(fldtarg);
  free(mp); free(lo); free(eoff); free(scale); free(bsize); free(pl);
  orc_tree_free(&t);
  return 0;
}
