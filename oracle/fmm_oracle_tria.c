/*
 * fmm_oracle_tria.c -- CPU oracle for fmmlib3d's Laplace triangle layer
 * potentials (TEST INFRASTRUCTURE ONLY; see fmm_oracle.h).
 *
 * Piecewise-constant single layer (SLP, density `charge`) and dipole layer
 * (DLP, density `dipstr`, oriented along the triangle normal) on flat
 * triangles, Green's function 1/r without 1/(4 pi), field = -gradient:
 *   SLP  pot(x) = int_T 1/|x-y| dA,            fld = int_T (x-y)/|x-y|^3 dA
 *   DLP  pot(x) = int_T n.(x-y)/|x-y|^3 dA,    fld = -grad pot
 * as the reference evaluates them with its analytic quadratures
 * (direct3dtritarglaps2 / direct3dtritarglapd2 for a point off the
 * triangle, direct3dtrilaps2 / direct3dtrilapd2 for a triangle's own
 * centroid, src/triquadflatlib.f:261-818 -> triahquad, src/triahquad.f:63),
 * and the all-pairs sum l3dtriadirect (src/lfmm3dtria.f:1385).
 *
 * The integrals are the classical closed forms, written here in the frame
 * of the triangle (tri_ini, src/trilib.f:527: vertex 1 at the origin, edge
 * 1->2 along +x, normal n = (v2-v1) x (v3-v1) / |.| along +z), for a point
 * p = (px, py, h).  Per edge k from a to b (counter-clockwise in the frame)
 * with unit tangent t, outward normal m = (t_y, -t_x), s_a = (a-p).t,
 * s_b = (b-p).t, d = (a-p).m, R_a = |a-p| (3-D), R0^2 = d^2 + h^2:
 *   f_k    = int_edge dl / R                = ln((R_b+s_b)/(R_a+s_a))
 *   beta_k = atan(d s_b/(R0^2+|h|R_b)) - atan(d s_a/(R0^2+|h|R_a))
 *   SLP pot   = sum d_k f_k - |h| sum beta_k
 *   SLP fld   = (sum m_k f_k, sign(h) sum beta_k)        (divergence theorem
 *               in the plane for the in-plane part; solid angle normally)
 *   DLP pot   = sign(h) sum beta_k   (= h int 1/R^3, the signed solid angle)
 *   DLP fld   = (-h sum m_k [1/(R_b(R_b+s_b)) - 1/(R_a(R_a+s_a))],
 *                sum (d_k/R0_k^2)(s_b/R_b - s_a/R_a))     (= -d/dh of the SLP
 *               field; the normal part by the divergence theorem applied to
 *               (y-p)/R^3 in the plane)
 * At a triangle's own centroid (h = 0, principal values as the reference's
 * iquad = 0 branch): SLP pot = sum d f, SLP fld = (sum m f, 0), DLP pot = 0,
 * DLP fld = (0, 0, sum (1/d)(s_b/R_b - s_a/R_a)) -- the finite part.
 * Every log and 1/(R(R+s)) is formed without cancellation for s < 0 through
 * (R+s)(R-s) = R0^2.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_oracle.h"

typedef struct {
  double o[3], e1[3], e2[3], e3[3];  /* origin (vertex 1) and frame axes */
  double v[3][2];                     /* vertices in the frame */
} TriFrame;

static void tri_frame(const double *tri, TriFrame *F) {
  const double *a = tri, *b = tri + 3, *c = tri + 6;
  double u[3], w[3], n[3];
  for (int k = 0; k < 3; k++) {
    F->o[k] = a[k];
    u[k] = b[k] - a[k];
    w[k] = c[k] - a[k];
  }
  n[0] = u[1] * w[2] - u[2] * w[1];
  n[1] = u[2] * w[0] - u[0] * w[2];
  n[2] = u[0] * w[1] - u[1] * w[0];
  const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; k++) {
    F->e1[k] = u[k] / lu;
    F->e3[k] = n[k] / ln;
  }
  F->e2[0] = F->e3[1] * F->e1[2] - F->e3[2] * F->e1[1];
  F->e2[1] = F->e3[2] * F->e1[0] - F->e3[0] * F->e1[2];
  F->e2[2] = F->e3[0] * F->e1[1] - F->e3[1] * F->e1[0];
  for (int j = 0; j < 3; j++) {
    const double *p = tri + 3 * j;
    double r[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    F->v[j][0] = r[0] * F->e1[0] + r[1] * F->e1[1] + r[2] * F->e1[2];
    F->v[j][1] = r[0] * F->e2[0] + r[1] * F->e2[1] + r[2] * F->e2[2];
  }
}

/* ln(R + s) for an edge end, with R + s = R0^2/(R - s) when s < 0 */
static double log_rs(double s, double R, double R02) {
  return s >= 0 ? log(R + s) : log(R02) - log(R - s);
}

/* 1/(R (R + s)) without cancellation */
static double inv_rrs(double s, double R, double R02) {
  return s >= 0 ? 1.0 / (R * (R + s)) : (R - s) / (R * R02);
}

/* Unit-density integrals of triangle `tri` (3 vertices, xyz each) at x.
 * self != 0: x is the triangle's own centroid (evaluated in the plane).
 * out[0] SLP pot, out[1..3] SLP fld, out[4] DLP pot, out[5..7] DLP fld,
 * fields in global coordinates. */
void orc_tri_unit(const double *tri, const double *x, int self, double *out) {
  TriFrame F;
  tri_frame(tri, &F);
  const double r[3] = {x[0] - F.o[0], x[1] - F.o[1], x[2] - F.o[2]};
  const double px = r[0] * F.e1[0] + r[1] * F.e1[1] + r[2] * F.e1[2];
  const double py = r[0] * F.e2[0] + r[1] * F.e2[1] + r[2] * F.e2[2];
  const double h = self ? 0.0 : r[0] * F.e3[0] + r[1] * F.e3[1] + r[2] * F.e3[2];
  const double ah = fabs(h), sg = h > 0 ? 1.0 : (h < 0 ? -1.0 : 0.0);
  double pot = 0, fx = 0, fy = 0, beta = 0, dfx = 0, dfy = 0, dfz = 0;
  double Rv[3];
  for (int j = 0; j < 3; j++) {
    const double dx = F.v[j][0] - px, dy = F.v[j][1] - py;
    Rv[j] = sqrt(dx * dx + dy * dy + h * h);
  }
  for (int k = 0; k < 3; k++) {
    const double *a = F.v[k], *b = F.v[(k + 1) % 3];
    const double ex = b[0] - a[0], ey = b[1] - a[1];
    const double L = sqrt(ex * ex + ey * ey);
    const double tx = ex / L, ty = ey / L, mx = ty, my = -tx;
    const double ax = a[0] - px, ay = a[1] - py, bx = b[0] - px, by = b[1] - py;
    const double sa = ax * tx + ay * ty, sb = bx * tx + by * ty;
    const double d = ax * mx + ay * my;
    const double Ra = Rv[k], Rb = Rv[(k + 1) % 3];
    const double R02 = d * d + h * h;
    double f;
    if (sa >= 0) f = log((Rb + sb) / (Ra + sa));
    else if (sb <= 0) f = log((Ra - sa) / (Rb - sb));
    else f = log_rs(sb, Rb, R02) - log_rs(sa, Ra, R02);
    pot += d * f;
    fx += mx * f;
    fy += my * f;
    beta += atan2(d * sb, R02 + ah * Rb) - atan2(d * sa, R02 + ah * Ra);
    if (h != 0) {
      const double g = inv_rrs(sb, Rb, R02) - inv_rrs(sa, Ra, R02);
      dfx -= h * mx * g;
      dfy -= h * my * g;
    }
    if (R02 > 0) dfz += (d / R02) * (sb / Rb - sa / Ra);
  }
  pot -= ah * beta;
  const double fz = sg * beta, dpot = sg * beta;
  double loc[2][3] = {{fx, fy, fz}, {dfx, dfy, dfz}};
  out[0] = pot;
  out[4] = dpot;
  for (int q = 0; q < 2; q++)
    for (int k = 0; k < 3; k++)
      out[1 + 4 * q + k] = loc[q][0] * F.e1[k] + loc[q][1] * F.e2[k] + loc[q][2] * F.e3[k];
}

/* l3dtriadirect (src/lfmm3dtria.f:1385): every triangle on every centroid
 * (its own by the self formulas) and on every target. */
void orc_l3dtriadirect(int ntri, const double *tri, const double *cent, int ifcharge,
                       const zdouble *charge, int ifdipole, const zdouble *dipstr, int ifpot,
                       zdouble *pot, int iffld, zdouble *fld, int ntarget, const double *target,
                       int ifpottarg, zdouble *pottarg, int iffldtarg, zdouble *fldtarg) {
  if (ifpot || iffld) {
    int *self = (int *)malloc(sizeof(int) * (size_t)(ntri > 0 ? ntri : 1));
    for (int j = 0; j < ntri; j++) self[j] = j;
    zdouble *p = (zdouble *)calloc((size_t)ntri + 1, sizeof(zdouble));
    zdouble *f = (zdouble *)calloc(3 * (size_t)ntri + 3, sizeof(zdouble));
    orc_l3dtriadirect_sample(ntri, tri, ifcharge, charge, ifdipole, dipstr, ntri, cent, self, p, f);
    for (int j = 0; j < ntri; j++) {
      if (ifpot) pot[j] = p[j];
      if (iffld)
        for (int k = 0; k < 3; k++) fld[3 * j + k] = f[3 * j + k];
    }
    free(self);
    free(p);
    free(f);
  }
  if (ntarget > 0 && (ifpottarg || iffldtarg)) {
    zdouble *p = (zdouble *)calloc((size_t)ntarget, sizeof(zdouble));
    zdouble *f = (zdouble *)calloc(3 * (size_t)ntarget, sizeof(zdouble));
    orc_l3dtriadirect_sample(ntri, tri, ifcharge, charge, ifdipole, dipstr, ntarget, target, NULL, p, f);
    for (int j = 0; j < ntarget; j++) {
      if (ifpottarg) pottarg[j] = p[j];
      if (iffldtarg)
        for (int k = 0; k < 3; k++) fldtarg[3 * j + k] = f[3 * j + k];
    }
    free(p);
    free(f);
  }
}

/* The direct sum at npts points; self_index[j] >= 0 names the triangle whose
 * centroid point j is (evaluated by the self formulas). */
void orc_l3dtriadirect_sample(int ntri, const double *tri, int ifcharge, const zdouble *charge,
                              int ifdipole, const zdouble *dipstr, int npts, const double *pts,
                              const int *self_index, zdouble *pot, zdouble *fld) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < npts; j++) {
    zdouble p = 0, f[3] = {0, 0, 0};
    const int me = self_index ? self_index[j] : -1;
    for (int i = 0; i < ntri; i++) {
      double u[8];
      orc_tri_unit(tri + 9 * i, pts + 3 * j, i == me, u);
      if (ifcharge) {
        p += charge[i] * u[0];
        for (int k = 0; k < 3; k++) f[k] += charge[i] * u[1 + k];
      }
      if (ifdipole) {
        p += dipstr[i] * u[4];
        for (int k = 0; k < 3; k++) f[k] += dipstr[i] * u[5 + k];
      }
    }
    pot[j] = p;
    for (int k = 0; k < 3; k++) fld[3 * j + k] = f[k];
  }
}
