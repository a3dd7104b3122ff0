/*
 * fmm_oracle_qp.c -- CPU oracle for Laplace quadrupole sources (TEST
 * INFRASTRUCTURE ONLY; see fmm_oracle.h).
 *
 * A quadrupole at y with complex strength quadstr and real components
 * quadvec = (Qxx, Qyy, Qzz, Qxy, Qxz, Qyz) has the potential
 *   pot(x) = quadstr * sum_k quadvec_k V_k(x - y),
 *   V_xx = d2/dx2 (1/r) = -1/r^3 + 3 dx^2/r^5,  V_xy = d2/dxdy (1/r) = 3 dx dy/r^5, ...
 * and fld = -grad pot, as the reference defines it (lpotfld3dhess_qp,
 * src/laprouts3d.f:8691; quadvec(4..6) weigh the mixed derivatives once).
 *
 * Its multipole and local expansions (the reference's l3dformmp_qp_trunc and
 * l3dformta_qp_trunc, src/laprouts3d.f:7914, :8261) follow from those of a
 * charge by differentiating with respect to the source position.  With
 * d+ = d/dx + i d/dy, d- = d/dx - i d/dy, the operator is
 *   D_Q = c_{+2} d+^2 + c_{-2} d-^2 + c_0 dz^2 + c_{+1} dz d+ + c_{-1} dz d-,
 *   c_{+-2} = (Qxx - Qyy)/4 -+ i Qxy/4,  c_0 = Qzz - (Qxx + Qyy)/2,
 *   c_{+-1} = Qxz/2 -+ i Qyz/2           (d+ d- = -dz^2 on harmonic functions),
 * and on the solid harmonics of this library's convention (fmm_oracle.h)
 *   dz R_n^m = a(n,m) R_{n-1}^m,        a = sqrt((n-m)(n+m)),
 *   d+ R_n^m = b(n,m) R_{n-1}^{m+1},    b = s(m) sqrt((n-m)(n-m-1)),
 *   d- R_n^m = g(n,m) R_{n-1}^{m-1},    g = s'(m) sqrt((n+m)(n+m-1)),
 *   dz S_n^m = -sqrt((n+1-m)(n+1+m)) S_{n+1}^m,
 *   d+ S_n^m = s(m) sqrt((n+m+1)(n+m+2)) S_{n+1}^{m+1},
 *   d- S_n^m = s'(m) sqrt((n-m+1)(n-m+2)) S_{n+1}^{m-1},
 * s(m) = 1 for m >= 0 and -1 below, s'(m) = 1 for m <= 0 and -1 above.
 * So D_Q R_n^m = sum_j c_j L_j(n,m) R_{n-2}^{m+j}, D_Q S_n^m = sum_j c_j
 * L'_j(n,m) S_{n+2}^{m+j}, and with charges q_j = quadstr conj(c_j):
 *   M_n^m += rscale^n  sum_j L_j(n,m)  q_j conj(R_{n-2}^{m+j}(y - c)),
 *   L_n^m += rscale^-n sum_j L'_j(n,m) q_j conj(S_{n+2}^{m+j}(y - c)).
 * (The ladder constants are checked here by central differences of
 * orc_regular / orc_irregular, and the expansions against the reference's
 * own l3dformmp_qp_trunc / l3dformta_qp_trunc.)
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_oracle.h"

#define QIDX(n, m) ((n) * (n) + (n) + (m))

static double sgn_ge0(int m) { return m >= 0 ? 1.0 : -1.0; }
static double sgn_le0(int m) { return m <= 0 ? 1.0 : -1.0; }
static double ra(int n, int m) { return (abs(m) <= n) ? sqrt((double)(n - m) * (n + m)) : 0.0; }
static double rb(int n, int m) {
  return (n - m >= 1) ? sgn_ge0(m) * sqrt((double)(n - m) * (n - m - 1)) : 0.0;
}
static double rg(int n, int m) {
  return (n + m >= 1) ? sgn_le0(m) * sqrt((double)(n + m) * (n + m - 1)) : 0.0;
}
static double sa(int n, int m) { return -sqrt((double)(n + 1 - m) * (n + 1 + m)); }
static double sb(int n, int m) { return sgn_ge0(m) * sqrt((double)(n + m + 1) * (n + m + 2)); }
static double sg(int n, int m) { return sgn_le0(m) * sqrt((double)(n - m + 1) * (n - m + 2)); }

/* L_j(n,m), j = -2..2 (index j+2): D_Q R_n^m = sum_j c_j L_j R_{n-2}^{m+j} */
void orc_qp_ladder_regular(int n, int m, double *Lj) {
  Lj[4] = rb(n, m) * rb(n - 1, m + 1);    /* d+^2 */
  Lj[0] = rg(n, m) * rg(n - 1, m - 1);    /* d-^2 */
  Lj[2] = ra(n, m) * ra(n - 1, m);        /* dz^2 */
  Lj[3] = rb(n, m) * ra(n - 1, m + 1);    /* dz d+ */
  Lj[1] = rg(n, m) * ra(n - 1, m - 1);    /* dz d- */
}

/* L'_j(n,m): D_Q S_n^m = sum_j c_j L'_j S_{n+2}^{m+j} */
void orc_qp_ladder_irregular(int n, int m, double *Lj) {
  Lj[4] = sb(n, m) * sb(n + 1, m + 1);
  Lj[0] = sg(n, m) * sg(n + 1, m - 1);
  Lj[2] = sa(n, m) * sa(n + 1, m);
  Lj[3] = sb(n, m) * sa(n + 1, m + 1);
  Lj[1] = sg(n, m) * sa(n + 1, m - 1);
}

/* the five channel charges q_j = quadstr conj(c_j), j = -2..2 */
void orc_qp_channels(zdouble quadstr, const double *qv, zdouble *qj) {
  const double xx = qv[0], yy = qv[1], zz = qv[2], xy = qv[3], xz = qv[4], yz = qv[5];
  const zdouble c[5] = {(xx - yy) / 4 + I * xy / 4, xz / 2 + I * yz / 2, zz - (xx + yy) / 2,
                        xz / 2 - I * yz / 2, (xx - yy) / 4 - I * xy / 4};
  for (int j = 0; j < 5; j++) qj[j] = quadstr * conj(c[j]);
}

/* multipole (IRREG = 0, order p, packed n^2+n+m) or local (IRREG = 1) about
 * center with scaling rscale, added into out, of ns quadrupoles */
void orc_l3dform_qp(int irreg, double rscale, const double *src, const zdouble *quadstr, const double *quadvec,
                    int ns, const double *center, int p, zdouble *out) {
  const int pe = p + 2;
  zdouble *H = (zdouble *)malloc(sizeof(zdouble) * (size_t)(pe + 1) * (pe + 1));
  for (int s = 0; s < ns; s++) {
    const double d[3] = {src[3 * s] - center[0], src[3 * s + 1] - center[1], src[3 * s + 2] - center[2]};
    if (irreg) orc_irregular(pe, d, H);
    else orc_regular(pe, d, H);
    zdouble qj[5];
    orc_qp_channels(quadstr[s], quadvec + 6 * (size_t)s, qj);
    for (int n = 0; n <= p; n++) {
      const double sc = irreg ? pow(rscale, -n) : pow(rscale, n);
      for (int m = -n; m <= n; m++) {
        double Lj[5];
        if (irreg) orc_qp_ladder_irregular(n, m, Lj);
        else orc_qp_ladder_regular(n, m, Lj);
        const int k = irreg ? n + 2 : n - 2;
        if (k < 0) continue;
        zdouble acc = 0;
        for (int j = -2; j <= 2; j++) {
          const int mm = m + j;
          if (mm < -k || mm > k || Lj[j + 2] == 0.0) continue;
          acc += Lj[j + 2] * qj[j + 2] * conj(H[QIDX(k, mm)]);
        }
        out[QIDX(n, m)] += sc * acc;
      }
    }
  }
  free(H);
}

/* pot and fld at npts points of ns quadrupoles (lpotfld3dhess_qp summed);
 * self_index[j] >= 0 skips that source */
void orc_l3ddirect_qp_sample(int ns, const double *src, const zdouble *quadstr, const double *quadvec, int npts,
                             const double *pts, const int *self_index, zdouble *pot, zdouble *fld) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < npts; j++) {
    zdouble p = 0, f[3] = {0, 0, 0};
    for (int s = 0; s < ns; s++) {
      if (self_index && self_index[j] == s) continue;
      const double rx = pts[3 * j] - src[3 * s], ry = pts[3 * j + 1] - src[3 * s + 1],
                   rz = pts[3 * j + 2] - src[3 * s + 2];
      const double r2 = rx * rx + ry * ry + rz * rz, r = sqrt(r2);
      const double d3 = 1.0 / (r * r2), d5 = d3 / r2, d7 = d5 / r2;
      const double *q = quadvec + 6 * (size_t)s;
      const double v11 = -d3 + 3 * d5 * rx * rx, v22 = -d3 + 3 * d5 * ry * ry, v33 = -d3 + 3 * d5 * rz * rz;
      const double v12 = 3 * d5 * rx * ry, v13 = 3 * d5 * rx * rz, v23 = 3 * d5 * ry * rz;
      const double pv = v11 * q[0] + v22 * q[1] + v33 * q[2] + v12 * q[3] + v13 * q[4] + v23 * q[5];
      /* third derivatives of 1/r: v_ijk */
      const double v111 = 9 * rx * d5 - 15 * d7 * rx * rx * rx, v222 = 9 * ry * d5 - 15 * d7 * ry * ry * ry,
                   v333 = 9 * rz * d5 - 15 * d7 * rz * rz * rz;
      const double v112 = 3 * ry * d5 - 15 * d7 * ry * rx * rx, v113 = 3 * rz * d5 - 15 * d7 * rz * rx * rx;
      const double v122 = 3 * rx * d5 - 15 * d7 * rx * ry * ry, v133 = 3 * rx * d5 - 15 * d7 * rx * rz * rz;
      const double v223 = 3 * rz * d5 - 15 * d7 * rz * ry * ry, v233 = 3 * ry * d5 - 15 * d7 * ry * rz * rz;
      const double v123 = -15 * d7 * rx * ry * rz;
      const double gx = v111 * q[0] + v122 * q[1] + v133 * q[2] + v112 * q[3] + v113 * q[4] + v123 * q[5];
      const double gy = v112 * q[0] + v222 * q[1] + v233 * q[2] + v122 * q[3] + v123 * q[4] + v223 * q[5];
      const double gz = v113 * q[0] + v223 * q[1] + v333 * q[2] + v123 * q[3] + v133 * q[4] + v233 * q[5];
      p += quadstr[s] * pv;
      f[0] -= quadstr[s] * gx;
      f[1] -= quadstr[s] * gy;
      f[2] -= quadstr[s] * gz;
    }
    pot[j] = p;
    for (int k = 0; k < 3; k++) fld[3 * (size_t)j + k] = f[k];
  }
}
