// Laplace layer potentials of flat triangles with constant densities on the
// device: single layer (density q, 1/r) and dipole layer (density mu along
// the triangle normal), pot and fld = -grad pot, as fmmlib3d's triangle FMM
// defines them (lfmm3dtriatarg, src/lfmm3dtria.f:121; the analytic
// quadratures direct3dtritarglaps2 / direct3dtritarglapd2 and, at a
// triangle's own centroid, direct3dtrilaps2 / direct3dtrilapd2,
// src/triquadflatlib.f:147-818).
//
// Geometry per triangle (12 doubles, sorted order): the frame of tri_ini
// (src/trilib.f:527): origin at vertex 1, e1 along edge 1->2, e3 the unit
// normal (v2-v1)x(v3-v1)/|.|, and the in-plane vertices (0,0), (L,0),
// (x3,y3), y3 > 0.  For a point p = (px, py, h) in that frame, per edge k
// from a to b with unit tangent t, outward normal m = (t_y, -t_x),
// s = (.-p).t at both ends, d = (a-p).m, R the 3-D distances,
// R0^2 = d^2 + h^2:
//   f_k = ln((R_b+s_b)/(R_a+s_a)),  beta_k = atan(d s_b/(R0^2+|h|R_b)) - atan(d s_a/(R0^2+|h|R_a))
//   single layer: pot = sum d f - |h| sum beta,  fld = (sum m f, sign(h) sum beta)
//   dipole layer: pot = sign(h) sum beta,
//                 fld = (-h sum m [1/(R_b(R_b+s_b)) - 1/(R_a(R_a+s_a))], sum (d/R0^2)(s_b/R_b - s_a/R_a))
// At a triangle's own centroid h = 0 (principal values: single-layer normal
// field 0, dipole-layer potential 0, its normal field the finite part).
// Far field: each triangle becomes the nq weighted nodes of a rule on it
// (symmetric up to degree 6, a collapsed-square product Gauss rule beyond)
// with the particle strengths
// layout (charge, dipole vector = mu w 2A n), and the particle P2M/P2L
// kernels form the expansions of the leaf boxes from those points.
#pragma once
#include <hip/hip_runtime.h>

#include "kern_common.hpp"
#include "lap_kernels.hpp"

namespace fmm3d {

constexpr int kTriGeo = 12;   // doubles per triangle frame

// frames, strengths (q, mu: 4 doubles) and quadrature points (4 doubles of
// position, 8 of strengths each) of the triangles in sorted order
static __global__ __launch_bounds__(256) void k_tri_prep(int n, const int* __restrict__ perm, const double* __restrict__ tri,
                                                  const double* __restrict__ q, const double* __restrict__ mu, int ifc,
                                                  int ifd, int nq, const double* __restrict__ rule,
                                                  double* __restrict__ geo, double* __restrict__ tstr,
                                                  double* __restrict__ npos, double* __restrict__ nstr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  const double* v = tri + 9 * (size_t)o;
  const double ux = v[3] - v[0], uy = v[4] - v[1], uz = v[5] - v[2];
  const double wx = v[6] - v[0], wy = v[7] - v[1], wz = v[8] - v[2];
  double nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
  const double area2 = sqrt(nx * nx + ny * ny + nz * nz);  // twice the area
  const double L = sqrt(ux * ux + uy * uy + uz * uz);
  nx /= area2; ny /= area2; nz /= area2;
  const double e1x = ux / L, e1y = uy / L, e1z = uz / L;
  const double e2x = ny * e1z - nz * e1y, e2y = nz * e1x - nx * e1z, e2z = nx * e1y - ny * e1x;
  double* g = geo + kTriGeo * (size_t)i;
  g[0] = v[0]; g[1] = v[1]; g[2] = v[2];
  g[3] = e1x; g[4] = e1y; g[5] = e1z;
  g[6] = nx; g[7] = ny; g[8] = nz;
  g[9] = L;
  g[10] = wx * e1x + wy * e1y + wz * e1z;
  g[11] = wx * e2x + wy * e2y + wz * e2z;
  const double qr = ifc ? q[2 * (size_t)o] : 0.0, qi = ifc ? q[2 * (size_t)o + 1] : 0.0;
  const double mr = ifd ? mu[2 * (size_t)o] : 0.0, mi = ifd ? mu[2 * (size_t)o + 1] : 0.0;
  double* ts = tstr + 4 * (size_t)i;
  ts[0] = qr; ts[1] = qi; ts[2] = mr; ts[3] = mi;
  for (int k = 0; k < nq; k++) {
    const double xi = rule[3 * k], eta = rule[3 * k + 1], w = rule[3 * k + 2] * area2;
    double* pp = npos + 4 * ((size_t)i * nq + k);
    pp[0] = v[0] + xi * ux + eta * wx;
    pp[1] = v[1] + xi * uy + eta * wy;
    pp[2] = v[2] + xi * uz + eta * wz;
    pp[3] = 0.0;
    double* ps = nstr + 8 * ((size_t)i * nq + k);
    ps[0] = w * qr; ps[1] = w * qi;
    ps[2] = w * mr * nx; ps[3] = w * mi * nx;
    ps[4] = w * mr * ny; ps[5] = w * mi * ny;
    ps[6] = w * mr * nz; ps[7] = w * mi * nz;
  }
}

// each box's source range counted in quadrature points (nq per triangle)
static __global__ void k_scale_ranges(int nb, const int* __restrict__ isrc, const int* __restrict__ nsrc, int nq,
                               int* __restrict__ isrc_q, int* __restrict__ nsrc_q) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  isrc_q[b] = isrc[b] * nq;
  nsrc_q[b] = nsrc[b] * nq;
}

// ln(R + s) with R + s = R0^2/(R - s) when s < 0
__device__ __forceinline__ double tri_log_rs(double s, double R, double R02) {
  return s >= 0.0 ? log(R + s) : log(R02) - log(R - s);
}
// 1/(R (R + s)) without cancellation
__device__ __forceinline__ double tri_inv_rrs(double s, double R, double R02) {
  return s >= 0.0 ? 1.0 / (R * (R + s)) : (R - s) / (R * R02);
}

// The point (x, y, z) in the frame of triangle g: in-plane coordinates
// (px, py), height h along the normal (0 at the triangle's own centroid),
// and the frame's second axis e2 = n x e1.
struct TriPoint {
  double px, py, h, e2x, e2y, e2z;
};

__device__ __forceinline__ TriPoint tri_point(const double* __restrict__ g, double x, double y, double z, bool self) {
  const double rx = x - g[0], ry = y - g[1], rz = z - g[2];
  const double e1x = g[3], e1y = g[4], e1z = g[5], nx = g[6], ny = g[7], nz = g[8];
  TriPoint t;
  t.e2x = ny * e1z - nz * e1y;
  t.e2y = nz * e1x - nx * e1z;
  t.e2z = nx * e1y - ny * e1x;
  t.px = rx * e1x + ry * e1y + rz * e1z;
  t.py = rx * t.e2x + ry * t.e2y + rz * t.e2z;
  t.h = self ? 0.0 : rx * nx + ry * ny + rz * nz;
  return t;
}

// Unit-density Laplace integrals of triangle g at the point t: u[0] single-
// layer pot, u[1..3] its fld, u[4] dipole-layer pot, u[5..7] its fld, the
// fields in global coordinates (CHG / DIP: which of the two are formed).
template <bool CHG, bool DIP>
__device__ __forceinline__ void tri_unit(const double* __restrict__ g, const TriPoint& t, double* u) {
  const double px = t.px, py = t.py, h = t.h;
  const double ah = fabs(h), h2 = h * h;
  const double vx[3] = {0.0, g[9], g[10]}, vy[3] = {0.0, 0.0, g[11]};
  double R[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double dx = vx[j] - px, dy = vy[j] - py;
    R[j] = sqrt(fma(dx, dx, fma(dy, dy, h2)));
  }
  double spot = 0.0, sfx = 0.0, sfy = 0.0, beta = 0.0, dfx = 0.0, dfy = 0.0, dfz = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int kb = k == 2 ? 0 : k + 1;
    const double ex = vx[kb] - vx[k], ey = vy[kb] - vy[k];
    const double il = 1.0 / sqrt(ex * ex + ey * ey);
    const double tx = ex * il, ty = ey * il, mx = ty, my = -tx;
    const double ax = vx[k] - px, ay = vy[k] - py;
    const double sa = ax * tx + ay * ty, sb = sa + (ex * tx + ey * ty);
    const double d = ax * mx + ay * my;
    const double Ra = R[k], Rb = R[kb];
    const double R02 = fma(d, d, h2);
    double f;
    if (sa >= 0.0) f = log((Rb + sb) / (Ra + sa));
    else if (sb <= 0.0) f = log((Ra - sa) / (Rb - sb));
    else f = tri_log_rs(sb, Rb, R02) - tri_log_rs(sa, Ra, R02);
    // beta: difference of two angles in (-pi/2, pi/2), as one atan2
    const double ya = d * sa, xa = fma(ah, Ra, R02), yb = d * sb, xb = fma(ah, Rb, R02);
    beta += atan2(yb * xa - xb * ya, xb * xa + yb * ya);
    if (CHG) {
      spot = fma(d, f, spot);
      sfx = fma(mx, f, sfx);
      sfy = fma(my, f, sfy);
    }
    if (DIP) {
      if (h != 0.0) {
        const double gk = tri_inv_rrs(sb, Rb, R02) - tri_inv_rrs(sa, Ra, R02);
        dfx = fma(-h * mx, gk, dfx);
        dfy = fma(-h * my, gk, dfy);
      }
      if (R02 > 0.0) dfz = fma(d / R02, sb / Rb - sa / Ra, dfz);
    }
  }
  const double sg = h > 0.0 ? 1.0 : (h < 0.0 ? -1.0 : 0.0);
  const double e1x = g[3], e1y = g[4], e1z = g[5], nx = g[6], ny = g[7], nz = g[8];
  if (CHG) {
    const double sfz = sg * beta;
    u[0] = fma(-ah, beta, spot);
    u[1] = sfx * e1x + sfy * t.e2x + sfz * nx;
    u[2] = sfx * e1y + sfy * t.e2y + sfz * ny;
    u[3] = sfx * e1z + sfy * t.e2z + sfz * nz;
  }
  if (DIP) {
    u[4] = sg * beta;
    u[5] = dfx * e1x + dfy * t.e2x + dfz * nx;
    u[6] = dfx * e1y + dfy * t.e2y + dfz * ny;
    u[7] = dfx * e1z + dfy * t.e2z + dfz * nz;
  }
}

// Unit-density single- and dipole-layer pot and fld of one triangle at
// (x, y, z); self: the point is the triangle's own centroid.  Adds
// cq * (single layer) + cm * (dipole layer) into acc[8] (pot, fld; complex).
template <bool CHG, bool DIP>
__device__ __forceinline__ void tri_pair(const double* __restrict__ g, const double* __restrict__ ts, double x,
                                         double y, double z, bool self, double* acc) {
  double u[8];
  tri_unit<CHG, DIP>(g, tri_point(g, x, y, z, self), u);
  if (CHG) {
    acc[0] = fma(ts[0], u[0], acc[0]); acc[1] = fma(ts[1], u[0], acc[1]);
    acc[2] = fma(ts[0], u[1], acc[2]); acc[3] = fma(ts[1], u[1], acc[3]);
    acc[4] = fma(ts[0], u[2], acc[4]); acc[5] = fma(ts[1], u[2], acc[5]);
    acc[6] = fma(ts[0], u[3], acc[6]); acc[7] = fma(ts[1], u[3], acc[7]);
  }
  if (DIP) {
    acc[0] = fma(ts[2], u[4], acc[0]); acc[1] = fma(ts[3], u[4], acc[1]);
    acc[2] = fma(ts[2], u[5], acc[2]); acc[3] = fma(ts[3], u[5], acc[3]);
    acc[4] = fma(ts[2], u[6], acc[4]); acc[5] = fma(ts[3], u[6], acc[5]);
    acc[6] = fma(ts[2], u[7], acc[6]); acc[7] = fma(ts[3], u[7], acc[7]);
  }
}

// Lanes of a near-field work item of n <= 64 points split into 64 / npad
// groups of npad lanes (npad the next power of two >= n): lane `pt` of each
// group evaluates point pt, group g takes sources g, g + G, ... of every box.
// The triangle FMMs' leaves hold a few points (10 at iprec 2), which one
// point per lane would leave most lanes of the wave idle.  The groups' sums
// are added across lanes (tri_group_sum) before the point's lane writes.
struct TriLanes {
  int npad, G, g, pt;
};
__device__ __forceinline__ TriLanes tri_lanes(int n) {
  int npad = 1;
  while (npad < n) npad <<= 1;
  const int lane = (int)(threadIdx.x & 63);
  return {npad, 64 / npad, lane / npad, lane & (npad - 1)};
}
// every lane of the wave takes part (no lane may have left before)
__device__ __forceinline__ void tri_group_sum(const TriLanes& L, double (&acc)[8]) {
  for (int off = L.npad; off < 64; off <<= 1)
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] += __shfl_xor(acc[k], off);
}

// Near field of the triangle FMM over the P2P work items: the triangles of
// the item's leaf (its own one by the centroid formulas) and of its list 1,
// lanes split over points and triangles as tri_lanes.
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_tri_p2p(DevPlan P, const double* __restrict__ geo,
                                                 const double* __restrict__ tstr, const double* __restrict__ cpos,
                                                 const double* __restrict__ tpos, int kind, double* __restrict__ pot,
                                                 double* __restrict__ fld) {
  const int w = __builtin_amdgcn_readfirstlane((int)(xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= P.n_p2p) return;
  const int b = P.p2p_items[3 * w], start = P.p2p_items[3 * w + 2];
  if (b < 0) return;  // padding item
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const TriLanes L = tri_lanes(min(64, nn - start));
  const int i = start + L.pt;
  const bool active = i < nn && L.g == 0;
  const int me = n0 + (i < nn ? i : 0);
  const double* mp = (kind ? tpos : cpos) + 4 * (size_t)me;
  const double x = mp[0], y = mp[1], z = mp[2];
  const int self = kind ? -1 : me;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int jj = P.l1_off[b] - 1; jj < P.l1_off[b + 1]; jj++) {
    const int j = jj < P.l1_off[b] ? b : P.l1_idx[jj];
    const int s0 = P.isrc[j], s1 = s0 + P.nsrc[j];
    for (int s = s0 + L.g; s < s1; s += L.G)
      tri_pair<CHG, DIP>(geo + kTriGeo * (size_t)s, tstr + 4 * (size_t)s, x, y, z, s == self, acc);
  }
  tri_group_sum(L, acc);
  if (!active) return;
  pot[2 * (size_t)me] += acc[0];
  pot[2 * (size_t)me + 1] += acc[1];
  if (fld)
    for (int k = 0; k < 6; k++) fld[6 * (size_t)me + k] += acc[2 + k];
}

// l3dtriadirect (src/lfmm3dtria.f:1385) on the device: every triangle on
// every point, in the caller's order; pts_are_centroids: point i is the
// centroid of triangle i (its own triangle by the centroid formulas).
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_tri_direct(int n, const double* __restrict__ geo,
                                                    const double* __restrict__ tstr, int npts,
                                                    const double* __restrict__ pts, int pts_are_centroids,
                                                    double* __restrict__ pot, double* __restrict__ fld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < npts;
  const int me = active ? i : 0;
  const double x = pts[3 * (size_t)me], y = pts[3 * (size_t)me + 1], z = pts[3 * (size_t)me + 2];
  const int self = pts_are_centroids ? me : -1;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = 0; s < n; s++)
    tri_pair<CHG, DIP>(geo + kTriGeo * (size_t)s, tstr + 4 * (size_t)s, x, y, z, s == self, acc);
  if (!active) return;
  pot[2 * (size_t)i] = acc[0];
  pot[2 * (size_t)i + 1] = acc[1];
  for (int k = 0; k < 6; k++) fld[6 * (size_t)i + k] = acc[2 + k];
}

}  // namespace fmm3d
