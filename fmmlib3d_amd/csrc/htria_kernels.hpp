// Helmholtz layer potentials of flat triangles with constant densities on
// the device: single layer (density q) and dipole layer (density mu along
// the triangle normal n), Green's function e^{ikr}/r without 1/(4 pi),
// pot and fld = -grad pot, as fmmlib3d's Helmholtz triangle FMM defines them
// (hfmm3dtriatarg, src/hfmm3dtria.f:133; the quadratures of h3dtriadirect,
// src/hfmm3dtria.f:1456 -> triquadhelm2, triquadhelmd2, triquadselfhelm,
// triquadselfhelmd, src/triquadflatlib.f:1516-2420).
//
// e^{ikr}/r = 1/r + G(r), G(r) = (e^{ikr} - 1)/r bounded: the 1/r part by
// the Laplace closed forms (tri_unit), G and its derivatives by a symmetric
// rule on the triangle.  In the triangle's frame, with d = (px - a, py - b, h)
// from the node (a, b) to the point, r = |d|, u = i k r, T = e^u (u - 1) + 1,
// Phi = T / r^3 and Psi = (-k^2 e^u - 3 T / r^2) / r^3:
//   single layer  pot += sum w G,            fld -= sum w d Phi
//   dipole layer  pot -= sum w h Phi,        fld += sum w (e3 Phi + h d Psi)
// where |k| r < 0.01, G and T / r^2 by their Taylor series in u (as the
// reference, against cancellation).  At a triangle's own centroid (h = 0)
// the leading term -k^2/(2r) of Phi is integrated in closed form:
//   single layer  fld -= (k^2/2) int (y-x)/|y-x| dA + sum w d (Phi + k^2/(2r))
//   dipole layer  pot = 0,  fld += e3 (-(k^2/2) int 1/r dA + sum w (Phi + k^2/(2r)))
// with int (y-x)/|y-x| dA = sum over edges m_k int_edge |y-x| dl in the
// plane.  lapc / lapd = 0 (a density flag of 2 in the reference) leave out
// that density's Laplace part: the difference kernel e^{ikr}/r - 1/r.
#pragma once
#include <hip/hip_runtime.h>

#include "tria_kernels.hpp"

namespace fmm3d {

constexpr int kSymRuleMax = 12;

// wavenumber, |k|, whether each density's Laplace part is included, and
// the symmetric rule on the reference triangle: nr nodes (xi, eta, w) in
// device memory, weights summing to 1/2
struct HtriParams {
  double kre, kim, ak;
  int lapc, lapd;
  int nr;
  const double* rule;
};

struct Zd {
  double r, i;
};
__device__ __forceinline__ Zd zmul(Zd a, Zd b) { return {a.r * b.r - a.i * b.i, fma(a.r, b.i, a.i * b.r)}; }
__device__ __forceinline__ Zd zadd(Zd a, Zd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ Zd zscale(Zd a, double s) { return {a.r * s, a.i * s}; }
// c0 + u (c1 + u (c2 + ...)) for real coefficients
__device__ __forceinline__ Zd zaff(Zd u, Zd t, double c) { Zd m = zmul(u, t); return {m.r + c, m.i}; }

// int_T (y - x)/|y - x| dA in the frame for a point (px, py) in the plane
__device__ __forceinline__ void tri_rvec_local(const double* __restrict__ g, double px, double py, double& ix,
                                               double& iy) {
  const double vx[3] = {0.0, g[9], g[10]}, vy[3] = {0.0, 0.0, g[11]};
  ix = 0.0;
  iy = 0.0;
  for (int k = 0; k < 3; k++) {
    const int kb = k == 2 ? 0 : k + 1;
    const double ex = vx[kb] - vx[k], ey = vy[kb] - vy[k];
    const double il = 1.0 / sqrt(ex * ex + ey * ey);
    const double tx = ex * il, ty = ey * il, mx = ty, my = -tx;
    const double ax = vx[k] - px, ay = vy[k] - py, bx = vx[kb] - px, by = vy[kb] - py;
    const double sa = ax * tx + ay * ty, sb = bx * tx + by * ty, d = ax * mx + ay * my;
    const double Ra = sqrt(ax * ax + ay * ay), Rb = sqrt(bx * bx + by * by), R02 = d * d;
    double f;
    if (sa >= 0.0) f = log((Rb + sb) / (Ra + sa));
    else if (sb <= 0.0) f = log((Ra - sa) / (Rb - sb));
    else f = tri_log_rs(sb, Rb, R02) - tri_log_rs(sa, Ra, R02);
    const double e = 0.5 * (sb * Rb - sa * Ra) + 0.5 * R02 * f;
    ix = fma(mx, e, ix);
    iy = fma(my, e, iy);
  }
}

// Adds cq * (single layer) + cm * (dipole layer) of triangle g at (x, y, z)
// into acc[8] (pot, fld; complex); self: the point is the triangle's own
// centroid.
template <bool CHG, bool DIP>
__device__ __forceinline__ void htri_pair(const double* __restrict__ g, const double* __restrict__ ts, double x,
                                          double y, double z, bool self, const HtriParams& hp, double* acc) {
  const TriPoint t = tri_point(g, x, y, z, self);
  double u[8];
  tri_unit<true, DIP>(g, t, u);
  const double L = g[9], x3 = g[10], y3 = g[11], area2 = L * y3;
  const double h = t.h, h2 = h * h;
  const Zd zz = {-hp.kim, hp.kre};                       // i k
  const Zd zz2 = zmul(zz, zz), zz3 = zmul(zz2, zz);     // -k^2, -i k^3
  Zd sp = {0, 0}, sfx = {0, 0}, sfy = {0, 0}, sfz = {0, 0};
  Zd dp = {0, 0}, dfx = {0, 0}, dfy = {0, 0}, dfz = {0, 0};
#pragma unroll 1
  for (int q = 0; q < hp.nr; q++) {
    const double xi = hp.rule[3 * q], eta = hp.rule[3 * q + 1];
    const double a = fma(xi, L, eta * x3), b = eta * y3, w = hp.rule[3 * q + 2] * area2;
    const double dx = t.px - a, dy = t.py - b;
    const double r2 = fma(dx, dx, fma(dy, dy, h2)), r = sqrt(r2);
    const Zd uu = zscale(zz, r);
    const double ex = hp.kim == 0.0 ? 1.0 : exp(uu.r);   // real k: |e^u| = 1
    double sn, cs;
    sincos(uu.i, &sn, &cs);
    const Zd e = {ex * cs, ex * sn};
    const bool taylor = hp.ak * r < 0.01;
    Zd G;
    if (taylor) {
      G = zmul(zz, zaff(uu, zaff(uu, zaff(uu, zaff(uu, {1.0 / 120.0, 0.0}, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0));
    } else {
      const double ir = 1.0 / r;
      G = {(e.r - 1.0) * ir, e.i * ir};
    }
    sp = zadd(sp, zscale(G, w));
    if (self) {
      // Phi + k^2/(2r) for the single-layer field (3-term series) and the
      // dipole-layer normal field (4-term series)
      Zd gs, gd;
      if (taylor) {
        gs = zmul(zz3, zaff(uu, zaff(uu, {1.0 / 30.0, 0.0}, 1.0 / 8.0), 1.0 / 3.0));
        gd = zmul(zz3, zaff(uu, zaff(uu, zaff(uu, {1.0 / 144.0, 0.0}, 1.0 / 30.0), 1.0 / 8.0), 1.0 / 3.0));
      } else {
        const double ir = 1.0 / r, ir3 = ir * ir * ir;
        const Zd um1 = {uu.r - 1.0, uu.i};
        const Zd T = zaff(e, um1, 1.0);
        gs = {fma(T.r, ir3, -0.5 * zz2.r * ir), fma(T.i, ir3, -0.5 * zz2.i * ir)};
        gd = gs;
      }
      if (CHG) {
        sfx = zadd(sfx, zscale(gs, -w * dx));
        sfy = zadd(sfy, zscale(gs, -w * dy));
      }
      if (DIP) dfz = zadd(dfz, zscale(gd, w));
    } else {
      Zd Phi, Psi;
      const double ir = 1.0 / r, ir3 = ir * ir * ir;
      if (taylor) {
        const Zd t1 = zmul(zz2, zaff(uu, zaff(uu, zaff(uu, {1.0 / 30.0, 0.0}, 1.0 / 8.0), 1.0 / 3.0), 0.5));
        Phi = zscale(t1, ir);
        const Zd ze = zmul(zz2, e);
        Psi = {(ze.r - 3.0 * t1.r) * ir3, (ze.i - 3.0 * t1.i) * ir3};
      } else {
        const Zd um1 = {uu.r - 1.0, uu.i};
        const Zd T = zaff(e, um1, 1.0);
        Phi = zscale(T, ir3);
        const Zd ze = zmul(zz2, e);
        const double ir2 = ir * ir;
        Psi = {(ze.r - 3.0 * T.r * ir2) * ir3, (ze.i - 3.0 * T.i * ir2) * ir3};
      }
      if (CHG) {
        sfx = zadd(sfx, zscale(Phi, -w * dx));
        sfy = zadd(sfy, zscale(Phi, -w * dy));
        sfz = zadd(sfz, zscale(Phi, -w * h));
      }
      if (DIP) {
        dp = zadd(dp, zscale(Phi, -w * h));
        const Zd whpsi = zscale(Psi, w * h);
        dfx = zadd(dfx, zscale(whpsi, dx));
        dfy = zadd(dfy, zscale(whpsi, dy));
        dfz = zadd(dfz, zadd(zscale(Phi, w), zscale(whpsi, h)));
      }
    }
  }
  if (self) {
    const Zd k2h = {-0.5 * zz2.r, -0.5 * zz2.i};  // k^2 / 2
    if (CHG) {
      double ix, iy;
      tri_rvec_local(g, t.px, t.py, ix, iy);
      sfx = zadd(sfx, zscale(k2h, -ix));
      sfy = zadd(sfy, zscale(k2h, -iy));
    }
    if (DIP) dfz = zadd(dfz, zscale(k2h, -u[0]));
  }
  const double e1x = g[3], e1y = g[4], e1z = g[5], nx = g[6], ny = g[7], nz = g[8];
  if (CHG) {
    const double lc = hp.lapc ? 1.0 : 0.0;
    const Zd p = {fma(lc, u[0], sp.r), sp.i};
    const Zd fx = {fma(lc, u[1], sfx.r * e1x + sfy.r * t.e2x + sfz.r * nx), sfx.i * e1x + sfy.i * t.e2x + sfz.i * nx};
    const Zd fy = {fma(lc, u[2], sfx.r * e1y + sfy.r * t.e2y + sfz.r * ny), sfx.i * e1y + sfy.i * t.e2y + sfz.i * ny};
    const Zd fz = {fma(lc, u[3], sfx.r * e1z + sfy.r * t.e2z + sfz.r * nz), sfx.i * e1z + sfy.i * t.e2z + sfz.i * nz};
    const Zd c = {ts[0], ts[1]};
    const Zd a0 = zmul(c, p), a1 = zmul(c, fx), a2 = zmul(c, fy), a3 = zmul(c, fz);
    acc[0] += a0.r; acc[1] += a0.i; acc[2] += a1.r; acc[3] += a1.i;
    acc[4] += a2.r; acc[5] += a2.i; acc[6] += a3.r; acc[7] += a3.i;
  }
  if (DIP) {
    const double ld = hp.lapd ? 1.0 : 0.0;
    const Zd p = {fma(ld, u[4], dp.r), dp.i};
    const Zd fx = {fma(ld, u[5], dfx.r * e1x + dfy.r * t.e2x + dfz.r * nx), dfx.i * e1x + dfy.i * t.e2x + dfz.i * nx};
    const Zd fy = {fma(ld, u[6], dfx.r * e1y + dfy.r * t.e2y + dfz.r * ny), dfx.i * e1y + dfy.i * t.e2y + dfz.i * ny};
    const Zd fz = {fma(ld, u[7], dfx.r * e1z + dfy.r * t.e2z + dfz.r * nz), dfx.i * e1z + dfy.i * t.e2z + dfz.i * nz};
    const Zd c = {ts[2], ts[3]};
    const Zd a0 = zmul(c, p), a1 = zmul(c, fx), a2 = zmul(c, fy), a3 = zmul(c, fz);
    acc[0] += a0.r; acc[1] += a0.i; acc[2] += a1.r; acc[3] += a1.i;
    acc[4] += a2.r; acc[5] += a2.i; acc[6] += a3.r; acc[7] += a3.i;
  }
}

// Near field of the Helmholtz triangle FMM over the P2P work items: the
// triangles of the item's leaf (its own one by the centroid formulas) and of
// its list 1, lanes split over points and triangles as tri_lanes.  Three waves
// per SIMD (168 VGPRs, a few values spilled to scratch) beat two without
// spills: the pair's FP64 sqrt / division / sincos chains are latency bound.
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_htri_p2p(DevPlan P, const double* __restrict__ geo,
                                                  const double* __restrict__ tstr, const double* __restrict__ cpos,
                                                  const double* __restrict__ tpos, int kind, HtriParams hp,
                                                  double* __restrict__ pot, double* __restrict__ fld) {
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
      htri_pair<CHG, DIP>(geo + kTriGeo * (size_t)s, tstr + 4 * (size_t)s, x, y, z, s == self, hp, acc);
  }
  tri_group_sum(L, acc);
  if (!active) return;
  pot[2 * (size_t)me] += acc[0];
  pot[2 * (size_t)me + 1] += acc[1];
  if (fld)
    for (int k = 0; k < 6; k++) fld[6 * (size_t)me + k] += acc[2 + k];
}

// h3dtriadirect (src/hfmm3dtria.f:1456) on the device: every triangle on
// every point, in the caller's order; pts_are_centroids: point i is the
// centroid of triangle i (its own triangle by the centroid formulas).
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_htri_direct(int n, const double* __restrict__ geo,
                                                     const double* __restrict__ tstr, int npts,
                                                     const double* __restrict__ pts, int pts_are_centroids,
                                                     HtriParams hp, double* __restrict__ pot,
                                                     double* __restrict__ fld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < npts;
  const int me = active ? i : 0;
  const double x = pts[3 * (size_t)me], y = pts[3 * (size_t)me + 1], z = pts[3 * (size_t)me + 2];
  const int self = pts_are_centroids ? me : -1;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = 0; s < n; s++)
    htri_pair<CHG, DIP>(geo + kTriGeo * (size_t)s, tstr + 4 * (size_t)s, x, y, z, s == self, hp, acc);
  if (!active) return;
  pot[2 * (size_t)i] = acc[0];
  pot[2 * (size_t)i + 1] = acc[1];
  for (int k = 0; k < 6; k++) fld[6 * (size_t)i + k] = acc[2 + k];
}

// Launchers (htria_near.hip): k_htri_p2p over n P2P work items of P (no
// launch when n is 0 or neither density is on), and k_htri_direct.
void launch_htri_p2p(const DevPlan& P, const int* items, int n, int ifcharge, int ifdipole, const double* geo,
                     const double* tstr, const double* cpos, const double* tpos, int kind, const HtriParams& hp,
                     double* pot, double* fld, hipStream_t s);
void launch_htri_direct(int n, const double* geo, const double* tstr, int npts, const double* pts, int centroids,
                        const HtriParams& hp, int ifcharge, int ifdipole, double* pot, double* fld, hipStream_t s);

}  // namespace fmm3d
