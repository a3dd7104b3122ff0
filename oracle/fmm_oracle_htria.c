/*
 * fmm_oracle_htria.c -- CPU oracle for fmmlib3d's Helmholtz triangle layer
 * potentials (TEST INFRASTRUCTURE ONLY; see fmm_oracle.h).
 *
 * Piecewise-constant single layer (density `charge`) and dipole layer
 * (density `dipstr` along the triangle normal n) on flat triangles, Green's
 * function e^{ikr}/r without 1/(4 pi), field = -gradient, as the reference
 * evaluates them (direct3dtritarghelms3 / direct3dtritarghelmd3 for a point
 * off the triangle and direct3dtrihelms2 / direct3dtrihelmd2 at a triangle's
 * own centroid, src/triquadflatlib.f:819-1358 -> triquadhelm2 :1516,
 * triquadhelmd2 :1870, triquadselfhelm :2060, triquadselfhelmd :2236; the
 * all-pairs sum h3dtriadirect, src/hfmm3dtria.f:1456):
 *
 *   e^{ikr}/r = 1/r + G(r),   G(r) = (e^{ikr} - 1)/r   (bounded, smooth in r)
 *
 * The 1/r part is the Laplace closed form (orc_tri_unit); G and its
 * derivatives are integrated with a D3-symmetric rule of the requested
 * degree on the triangle.  With d = x - y (target minus node), r = |d|,
 * u = i k r, T(u) = e^u (u - 1) + 1:
 *   Phi = T / r^3,   Psi = (-k^2 e^u - 3 T / r^2) / r^3
 *   single layer  pot += sum w G,            fld -= sum w d Phi
 *   dipole layer  pot -= sum w (n.d) Phi,    fld += sum w (n Phi + (n.d) d Psi)
 * and, where |k| r < 0.01, G and T/r^2 by their Taylor series (four and
 * five terms, as the reference does against cancellation).
 * At a triangle's own centroid (in the plane) the leading -k^2/(2r) term of
 * Phi is integrated analytically:
 *   single layer  fld = fld_Laplace - (k^2/2) int (y-x)/|y-x| dA
 *                       - sum w d (Phi + k^2/(2r))
 *   dipole layer  pot = 0,  fld = fld_Laplace + n (-(k^2/2) int 1/r dA
 *                       + sum w (Phi + k^2/(2r)))
 * with int (y-x)/|y-x| dA = sum over edges m_k int_edge |y-x| dl by the
 * divergence theorem in the plane (the reference's triahquad types 8, 9).
 * Flag 2 for a density (the reference's ifl = 1) drops its Laplace part:
 * the difference kernel e^{ikr}/r - 1/r.
 *
 * Symmetric rules: the reference uses the D3-symmetric rules of Xiao and
 * Gimbutas (triasymq, src/triasymq.f:18).  For degrees 1..6 they have 1, 3,
 * 6, 6, 7 and 12 nodes (degree 3 is the degree-4 rule) made of orbits of the
 * centroid (S3), of (a, a, 1-2a) (S21) and of (a, b, 1-a-b) (S111); the
 * orbit parameters are the solution of the moment equations
 * sum w x^i y^j = i! j! / (i+j+2)!, i + j <= degree, found here by
 * Gauss-Newton from 2-digit starting values.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_oracle.h"

/* ---- symmetric rules ---- */

enum { ORB_S3 = 3, ORB_S21 = 21, ORB_S111 = 111 };

static int orbit_nodes(const double *par, const int *orb, int norb, double *xi, double *eta, double *w) {
  int k = 0, m = 0;
  for (int o = 0; o < norb; o++) {
    if (orb[o] == ORB_S3) {
      xi[m] = eta[m] = 1.0 / 3.0;
      w[m++] = par[k++];
    } else if (orb[o] == ORB_S21) {
      const double a = par[k], ww = par[k + 1];
      const double l[3][3] = {{a, a, 1 - 2 * a}, {a, 1 - 2 * a, a}, {1 - 2 * a, a, a}};
      for (int j = 0; j < 3; j++) {
        xi[m] = l[j][1];
        eta[m] = l[j][2];
        w[m++] = ww;
      }
      k += 2;
    } else {
      const double a = par[k], b = par[k + 1], c = 1 - a - b, ww = par[k + 2];
      const double l[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
      for (int j = 0; j < 6; j++) {
        xi[m] = l[j][1];
        eta[m] = l[j][2];
        w[m++] = ww;
      }
      k += 3;
    }
  }
  return m;
}

static double fact(int n) {
  double f = 1;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

static int moment_resid(const double *par, const int *orb, int norb, int deg, double *r) {
  double xi[16], eta[16], w[16];
  const int m = orbit_nodes(par, orb, norb, xi, eta, w);
  int e = 0;
  for (int i = 0; i <= deg; i++)
    for (int j = 0; i + j <= deg; j++) {
      double s = 0;
      for (int q = 0; q < m; q++) s += w[q] * pow(xi[q], i) * pow(eta[q], j);
      r[e++] = s - fact(i) * fact(j) / fact(i + j + 2);
    }
  return e;
}

/* n x n system A x = b by Gaussian elimination with partial pivoting */
static void solve_dense(int n, double *A, double *b) {
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(A[r * n + c]) > fabs(A[p * n + c])) p = r;
    if (p != c) {
      for (int k = 0; k < n; k++) {
        const double t = A[c * n + k];
        A[c * n + k] = A[p * n + k];
        A[p * n + k] = t;
      }
      const double t = b[c];
      b[c] = b[p];
      b[p] = t;
    }
    for (int r = c + 1; r < n; r++) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int c = n - 1; c >= 0; c--) {
    double s = b[c];
    for (int k = c + 1; k < n; k++) s -= A[c * n + k] * b[k];
    b[c] = s / A[c * n + c];
  }
}

/* Nodes (xi, eta) and weights (summing to 1/2, the reference triangle's
 * area) of the symmetric rule of degree 1..6; returns the node count, 0 for
 * an unsupported degree.  At most 12 nodes. */
int orc_tri_symrule(int degree, double *xi, double *eta, double *w) {
  int orb[3], norb, np;
  double par[7];
  switch (degree) {
    case 1:
      xi[0] = eta[0] = 1.0 / 3.0;
      w[0] = 0.5;
      return 1;
    case 2:
      orb[0] = ORB_S21;
      norb = 1;
      par[0] = 1.0 / 6.0;
      par[1] = 1.0 / 6.0;
      return orbit_nodes(par, orb, norb, xi, eta, w);
    case 3:
    case 4: {
      orb[0] = orb[1] = ORB_S21;
      norb = 2;
      const double s[4] = {0.09, 0.05, 0.45, 0.11};
      memcpy(par, s, sizeof(s));
      np = 4;
      degree = 4;
      break;
    }
    case 5: {
      orb[0] = ORB_S3;
      orb[1] = orb[2] = ORB_S21;
      norb = 3;
      const double s[5] = {0.11, 0.10, 0.06, 0.47, 0.07};
      memcpy(par, s, sizeof(s));
      np = 5;
      break;
    }
    case 6: {
      orb[0] = orb[1] = ORB_S21;
      orb[2] = ORB_S111;
      norb = 3;
      const double s[7] = {0.48, 0.04, 0.22, 0.09, 0.02, 0.14, 0.02};
      memcpy(par, s, sizeof(s));
      np = 7;
      break;
    }
    default:
      return 0;
  }
  /* Gauss-Newton on the (consistent, overdetermined) moment equations,
   * Jacobian by central differences */
  double r[64], rp[64], rm[64], J[64 * 7];
  for (int it = 0; it < 40; it++) {
    const int ne = moment_resid(par, orb, norb, degree, r);
    for (int k = 0; k < np; k++) {
      double pp[7], pm[7];
      memcpy(pp, par, sizeof(par));
      memcpy(pm, par, sizeof(par));
      pp[k] += 1e-7;
      pm[k] -= 1e-7;
      moment_resid(pp, orb, norb, degree, rp);
      moment_resid(pm, orb, norb, degree, rm);
      for (int e = 0; e < ne; e++) J[e * np + k] = (rp[e] - rm[e]) / 2e-7;
    }
    double A[49], b[7];
    for (int i = 0; i < np; i++) {
      b[i] = 0;
      for (int e = 0; e < ne; e++) b[i] -= J[e * np + i] * r[e];
      for (int j = 0; j < np; j++) {
        double s = 0;
        for (int e = 0; e < ne; e++) s += J[e * np + i] * J[e * np + j];
        A[i * np + j] = s;
      }
    }
    solve_dense(np, A, b);
    double mx = 0;
    for (int i = 0; i < np; i++) {
      par[i] += b[i];
      if (fabs(b[i]) > mx) mx = fabs(b[i]);
    }
    if (mx < 1e-17) break;
  }
  return orbit_nodes(par, orb, norb, xi, eta, w);
}

/* ---- unit integrals ---- */

/* int_T (y - x)/|y - x| dA for x in the triangle's plane (x's normal
 * component is ignored), in global coordinates */
static void tri_rvec(const double *tri, const double *x, double *out) {
  const double *a = tri, *b = tri + 3, *c = tri + 6;
  double u[3], v[3], n[3], e1[3], e2[3];
  for (int k = 0; k < 3; k++) {
    u[k] = b[k] - a[k];
    v[k] = c[k] - a[k];
  }
  n[0] = u[1] * v[2] - u[2] * v[1];
  n[1] = u[2] * v[0] - u[0] * v[2];
  n[2] = u[0] * v[1] - u[1] * v[0];
  const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; k++) {
    e1[k] = u[k] / lu;
    n[k] /= ln;
  }
  e2[0] = n[1] * e1[2] - n[2] * e1[1];
  e2[1] = n[2] * e1[0] - n[0] * e1[2];
  e2[2] = n[0] * e1[1] - n[1] * e1[0];
  double V[3][2];
  for (int j = 0; j < 3; j++) {
    const double *p = tri + 3 * j;
    const double r[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    V[j][0] = r[0] * e1[0] + r[1] * e1[1] + r[2] * e1[2];
    V[j][1] = r[0] * e2[0] + r[1] * e2[1] + r[2] * e2[2];
  }
  const double r[3] = {x[0] - a[0], x[1] - a[1], x[2] - a[2]};
  const double px = r[0] * e1[0] + r[1] * e1[1] + r[2] * e1[2];
  const double py = r[0] * e2[0] + r[1] * e2[1] + r[2] * e2[2];
  double ix = 0, iy = 0;
  for (int k = 0; k < 3; k++) {
    const double *p = V[k], *q = V[(k + 1) % 3];
    const double ex = q[0] - p[0], ey = q[1] - p[1];
    const double L = sqrt(ex * ex + ey * ey);
    const double tx = ex / L, ty = ey / L, mx = ty, my = -tx;
    const double ax = p[0] - px, ay = p[1] - py, bx = q[0] - px, by = q[1] - py;
    const double sa = ax * tx + ay * ty, sb = bx * tx + by * ty, d = ax * mx + ay * my;
    const double Ra = sqrt(ax * ax + ay * ay), Rb = sqrt(bx * bx + by * by);
    /* ln((Rb + sb)/(Ra + sa)) without cancellation for s < 0 */
    double f;
    if (sa >= 0) f = log((Rb + sb) / (Ra + sa));
    else if (sb <= 0) f = log((Ra - sa) / (Rb - sb));
    else f = log(Rb + sb) - (log(d * d) - log(Ra - sa));
    const double e = 0.5 * (sb * Rb - sa * Ra) + 0.5 * d * d * f;
    ix += mx * e;
    iy += my * e;
  }
  for (int k = 0; k < 3; k++) out[k] = ix * e1[k] + iy * e2[k];
}

/* Unit-density integrals of triangle `tri` at x for wavenumber zk with the
 * symmetric rule (nq nodes xi, eta, w on the reference triangle); self != 0:
 * x is the triangle's own centroid.  lap_c / lap_d: include the Laplace
 * part of the single / dipole layer (0 for the difference kernel).
 * out: SLP pot, SLP fld[3], DLP pot, DLP fld[3] (complex). */
void orc_htri_unit(const double *tri, const double *x, int self, zdouble zk, int nq, const double *rxi,
                   const double *reta, const double *rw, int lap_c, int lap_d, zdouble *out) {
  double lap[8];
  orc_tri_unit(tri, x, self, lap);
  const double *a = tri, *b = tri + 3, *c = tri + 6;
  double u[3], v[3], n[3];
  for (int k = 0; k < 3; k++) {
    u[k] = b[k] - a[k];
    v[k] = c[k] - a[k];
  }
  n[0] = u[1] * v[2] - u[2] * v[1];
  n[1] = u[2] * v[0] - u[0] * v[2];
  n[2] = u[0] * v[1] - u[1] * v[0];
  const double area2 = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; k++) n[k] /= area2;
  const zdouble zz = I * zk, zk2 = zk * zk;
  const double ak = cabs(zk);
  zdouble sp = 0, sf[3] = {0, 0, 0}, dp = 0, df[3] = {0, 0, 0};
  for (int q = 0; q < nq; q++) {
    const double w = rw[q] * area2;
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = x[k] - (a[k] + rxi[q] * u[k] + reta[q] * v[k]);
    const double dn0 = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
    if (self)
      for (int k = 0; k < 3; k++) d[k] -= dn0 * n[k];
    const double dn = self ? 0.0 : dn0;
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2), r3 = r * r2;
    const zdouble zarg = zz * r;
    const zdouble ze = cexp(zarg);
    zdouble G, Phi, Psi, Phis; /* Phis = Phi + k^2/(2r) (self) */
    if (ak * r < 0.01) {
      G = zz * (1.0 + zarg / 2.0 + zarg * zarg / 6.0 + zarg * zarg * zarg / 24.0 + zarg * zarg * zarg * zarg / 120.0);
      const zdouble t1 = zz * zz / 2.0 + r * zz * zz * zz / 3.0 + r2 * zz * zz * zz * zz / 8.0 +
                         r3 * zz * zz * zz * zz * zz / 30.0; /* T / r^2 */
      Phi = r > 0 ? t1 / r : 0;
      Psi = r > 0 ? (zz * zz * ze - 3.0 * t1) / r3 : 0;
      Phis = self ? zz * zz * zz / 3.0 + r * zz * zz * zz * zz / 8.0 + r2 * zz * zz * zz * zz * zz / 30.0 +
                        r3 * zz * zz * zz * zz * zz * zz / 144.0
                  : 0;
    } else {
      const zdouble T = ze * (zarg - 1.0) + 1.0;
      G = (ze - 1.0) / r;
      Phi = T / r3;
      Psi = (zz * zz * ze - 3.0 * T / r2) / r3;
      Phis = Phi + zk2 / (2.0 * r);
    }
    sp += w * G;
    if (self) {
      /* single layer: the quadrature of d (Phi + k^2/(2r)); its series
       * lacks the r^3 term (the reference's 3-term series for the field) */
      const zdouble gs = ak * r < 0.01
                             ? zz * zz * zz / 3.0 + r * zz * zz * zz * zz / 8.0 + r2 * zz * zz * zz * zz * zz / 30.0
                             : Phis;
      for (int k = 0; k < 3; k++) sf[k] -= w * d[k] * gs;
      for (int k = 0; k < 3; k++) df[k] += w * n[k] * Phis;
    } else {
      for (int k = 0; k < 3; k++) sf[k] -= w * d[k] * Phi;
      dp -= w * dn * Phi;
      for (int k = 0; k < 3; k++) df[k] += w * (n[k] * Phi + dn * d[k] * Psi);
    }
  }
  if (self) {
    double iv[3];
    tri_rvec(tri, x, iv);
    for (int k = 0; k < 3; k++) sf[k] -= 0.5 * zk2 * iv[k];
    for (int k = 0; k < 3; k++) df[k] -= 0.5 * zk2 * lap[0] * n[k];
  }
  out[0] = sp + (lap_c ? lap[0] : 0.0);
  for (int k = 0; k < 3; k++) out[1 + k] = sf[k] + (lap_c ? lap[1 + k] : 0.0);
  out[4] = dp + (lap_d ? lap[4] : 0.0);
  for (int k = 0; k < 3; k++) out[5 + k] = df[k] + (lap_d ? lap[5 + k] : 0.0);
}

/* h3dtriadirect (src/hfmm3dtria.f:1456) at npts points with the rule of
 * degree nqtri; self_index[j] >= 0 names the triangle whose centroid point j
 * is.  ifcharge / ifdipole: 0 off, 1 on, 2 on with the Laplace part
 * subtracted (the difference kernel). */
void orc_h3dtriadirect_sample(int ntri, const double *tri, zdouble zk, int nqtri, int ifcharge,
                              const zdouble *charge, int ifdipole, const zdouble *dipstr, int npts,
                              const double *pts, const int *self_index, zdouble *pot, zdouble *fld) {
  double rxi[16], reta[16], rw[16];
  const int nq = orc_tri_symrule(nqtri, rxi, reta, rw);
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < npts; j++) {
    zdouble p = 0, f[3] = {0, 0, 0};
    const int me = self_index ? self_index[j] : -1;
    for (int i = 0; i < ntri; i++) {
      zdouble o[8];
      orc_htri_unit(tri + 9 * (size_t)i, pts + 3 * (size_t)j, i == me, zk, nq, rxi, reta, rw, ifcharge == 1,
                    ifdipole == 1, o);
      if (ifcharge) {
        p += charge[i] * o[0];
        for (int k = 0; k < 3; k++) f[k] += charge[i] * o[1 + k];
      }
      if (ifdipole) {
        p += dipstr[i] * o[4];
        for (int k = 0; k < 3; k++) f[k] += dipstr[i] * o[5 + k];
      }
    }
    pot[j] = p;
    for (int k = 0; k < 3; k++) fld[3 * (size_t)j + k] = f[k];
  }
}
