// Hessians of the Laplace potential, hess = (pot_xx, pot_yy, pot_zz, pot_xy,
// pot_xz, pot_yz) per point (6 complex), as the reference's library routines
// define them: lpotfld3dhess (charges, src/laprouts3d.f:5442-5527) and
// lpotfld3dhess_dp (dipoles, :5619-5733) for the near field; the second
// derivatives of local (l3dtaevalhess, :5297) and multipole (l3dmpevalhess,
// :5229) expansions for the far field.  The reference's particle FMM itself
// returns pot and fld only; these kernels add the Hessian to the same
// schedule in FP64.
//
// Far field: the derivative operators Z = d/dz, U = d/dx + i d/dy,
// D = d/dx - i d/dy map an expansion to an expansion of the derivative
// (the factors AR/BR/CR for local, AS/BS/CS for multipole expansions, the
// ones the gradient uses); applied twice they give ZZ, ZU, ZD, UU, UD, DD
// and
//   xx = (UU + 2 UD + DD)/4,  yy = -(UU - 2 UD + DD)/4,  zz = ZZ,
//   xy = (UU - DD)/(4i),      xz = (ZU + ZD)/2,          yz = (ZU - ZD)/(2i),
// each scaled by 1/h^2 (local) or 1/h^3 (multipole, whose potential carries
// 1/h) of the expansion's box.  The factor table is the one of order p + 1
// (degrees up to p + 2, which the multipole's second derivatives reach).
#pragma once
#include <hip/hip_runtime.h>

#include "kern_common.hpp"
#include "lap_kernels.hpp"

namespace fmm3d {

// ---- near field: the self box and list 1 of a P2P work item, one point per
// lane, sources in order (self pair excluded by index)
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_p2p_hess(DevPlan P, const double* __restrict__ spos,
                                                  const double* __restrict__ str, const double* __restrict__ tpos,
                                                  int kind, double* __restrict__ hess) {
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= P.n_p2p) return;
  const int b = P.p2p_items[3 * w], start = P.p2p_items[3 * w + 2];
  if (b < 0) return;  // padding item
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + (int)(threadIdx.x & 63);
  const bool active = i < nn;
  const int me = n0 + (active ? i : 0);
  const double* mp = (kind ? tpos : spos) + 4 * (size_t)me;
  const double x = mp[0], y = mp[1], z = mp[2];
  double h[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto add = [&](int s, bool self) {
    const double* sp = spos + 4 * (size_t)s;
    const double* st = str + 8 * (size_t)s;
    const double dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir = self ? 0.0 : rsqrt_t<double>(r2);
    const double ir2 = ir * ir, ir5 = ir2 * ir2 * ir;
    const double d[3] = {dx, dy, dz};
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    if (CHG) {
      // q (3 d_a d_b - r^2 delta_ab) / r^5
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double g = (3.0 * d[ia[k]] * d[ib[k]] - (k < 3 ? r2 : 0.0)) * ir5;
        h[2 * k] = fma(st[0], g, h[2 * k]);
        h[2 * k + 1] = fma(st[1], g, h[2 * k + 1]);
      }
    }
    if (DIP) {
      // 3 [ (v.d)(5 d_a d_b / r^2 - delta_ab) - (v_a d_b + v_b d_a) ] / r^5
      const double dr = st[2] * dx + st[4] * dy + st[6] * dz, di = st[3] * dx + st[5] * dy + st[7] * dz;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double g = 5.0 * d[ia[k]] * d[ib[k]] * ir2 - (k < 3 ? 1.0 : 0.0);
        const double vr = st[2 + 2 * ia[k]] * d[ib[k]] + st[2 + 2 * ib[k]] * d[ia[k]];
        const double vi = st[3 + 2 * ia[k]] * d[ib[k]] + st[3 + 2 * ib[k]] * d[ia[k]];
        h[2 * k] = fma(3.0 * ir5, dr * g - vr, h[2 * k]);
        h[2 * k + 1] = fma(3.0 * ir5, di * g - vi, h[2 * k + 1]);
      }
    }
  };
  {
    const int s0 = P.isrc[b], s1 = s0 + P.nsrc[b];
    for (int s = s0; s < s1; s++) add(s, kind == 0 && s == me);
  }
  for (int jj = P.l1_off[b]; jj < P.l1_off[b + 1]; jj++) {
    const int j = P.l1_idx[jj];
    const int s0 = P.isrc[j], s1 = s0 + P.nsrc[j];
    for (int s = s0; s < s1; s++) add(s, false);
  }
  if (!active) return;
  double* o = hess + 12 * (size_t)me;
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] += h[k];
}

// The six second-derivative sums (complex) as the Cartesian Hessian, times s,
// added to o[12].
__device__ __forceinline__ void hess_cartesian(const C2<double> (&X)[6], double s, double* o) {
  // X: ZZ, ZU, ZD, UU, UD, DD
  const C2<double> zz = X[0], zu = X[1], zd = X[2], uu = X[3], ud = X[4], dd = X[5];
  const double xx_r = 0.25 * (uu.re + 2.0 * ud.re + dd.re), xx_i = 0.25 * (uu.im + 2.0 * ud.im + dd.im);
  const double yy_r = -0.25 * (uu.re - 2.0 * ud.re + dd.re), yy_i = -0.25 * (uu.im - 2.0 * ud.im + dd.im);
  // (a + i b) / (4i) = (b - i a) / 4 ; (a + i b) / (2i) = (b - i a) / 2
  const double xy_r = 0.25 * (uu.im - dd.im), xy_i = -0.25 * (uu.re - dd.re);
  const double xz_r = 0.5 * (zu.re + zd.re), xz_i = 0.5 * (zu.im + zd.im);
  const double yz_r = 0.5 * (zu.im - zd.im), yz_i = -0.5 * (zu.re - zd.re);
  o[0] += xx_r * s; o[1] += xx_i * s;
  o[2] += yy_r * s; o[3] += yy_i * s;
  o[4] += zz.re * s; o[5] += zz.im * s;
  o[6] += xy_r * s; o[7] += xy_i * s;
  o[8] += xz_r * s; o[9] += xz_i * s;
  o[10] += yz_r * s; o[11] += yz_i * s;
}

// coefficient f1(n1,m1) f2(n2,m2) E(n2,m2) of an expansion E of order pe,
// zero where an index is outside (|m| > n, n < 0 or n > pe)
__device__ __forceinline__ C2<double> hess_coef(const double* fac, int o1, int n1, int m1, int o2, int n2, int m2,
                                                const double* E, int pe) {
  if (n1 < 0 || m1 > n1 || m1 < -n1 || n2 < 0 || n2 > pe || m2 > n2 || m2 < -n2) return {0.0, 0.0};
  const double f = fac[o1 + eidx(n1, m1)] * fac[o2 + eidx(n2, m2)];
  return {f * E[2 * eidx(n2, m2)], f * E[2 * eidx(n2, m2) + 1]};
}

// out += c+ h + c- conj(h) for the (n, a) and (n, -a) coefficients
__device__ __forceinline__ void hess_fma(C2<double>& out, C2<double> cp, C2<double> cm, C2<double> h) {
  out.re += cp.re * h.re - cp.im * h.im + cm.re * h.re + cm.im * h.im;
  out.im += cp.re * h.im + cp.im * h.re - cm.re * h.im + cm.im * h.re;
}

// ---- far field, local expansions: one L2P work item (box, kind, first
// point) per wave; the six second-derivative expansions of degree <= p - 2
// are formed on the fly from the box's local (wave-uniform loads)
__global__ __launch_bounds__(256) void k_l2p_hess(DevPlan P, const double* __restrict__ fac2,
                                                  const double* __restrict__ lo, const double* __restrict__ spos,
                                                  const double* __restrict__ tpos, int want0, int want1,
                                                  double* __restrict__ hess0, double* __restrict__ hess1) {
  const int p = P.p;
  const FacLayout L(p + 1);
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= P.n_l2pi) return;
  const int ib = P.l2p_items[3 * w], kind = P.l2p_items[3 * w + 1], start = P.l2p_items[3 * w + 2];
  if (ib < 0) return;
  if (!(kind ? want1 : want0)) return;
  const int b = P.l2p_boxes[ib];
  const double* E = lo + 2 * (size_t)b * P.K;
  const double ih = P.inv_side[P.level[b]];
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + (int)(threadIdx.x & 63);
  const bool active = i < nn;
  double x = cx, y = cy, z = cz;
  if (active) {
    const double* q = (kind ? tpos : spos) + 4 * (size_t)(n0 + i);
    x = q[0]; y = q[1]; z = q[2];
  }
  const double ux = (x - cx) * ih, uy = (y - cy) * ih, uz = (z - cz) * ih;
  const double r2 = ux * ux + uy * uy + uz * uz;
  C2<double> X[6];
#pragma unroll
  for (int k = 0; k < 6; k++) X[k] = {0.0, 0.0};
  const int pe = p - 2;
  C2<double> hmm = {1.0, 0.0};
  for (int a = 0; a <= pe; a++) {
    if (a > 0) {
      const double f = fac2[L.MM + a];
      const C2<double> v = {-ux * f, -uy * f};
      hmm = cmul(v, hmm);
    }
    C2<double> h0 = hmm, h1 = hmm;
    for (int n = a; n <= pe; n++) {
      C2<double> h;
      if (n == a) {
        h = hmm;
      } else if (n == a + 1) {
        const double g = fac2[L.M1 + a] * uz;
        h = {g * hmm.re, g * hmm.im};
        h1 = h;
      } else {
        const double aa = (2 * n - 1) * uz, bq = fac2[L.RA + hidx(n, a)] * r2, cc = fac2[L.RB + hidx(n, a)];
        h = {(aa * h1.re - bq * h0.re) * cc, (aa * h1.im - bq * h0.im) * cc};
        h0 = h1;
        h1 = h;
      }
      // second-derivative coefficients at (n, a) and (n, -a)
      for (int sgn = 0; sgn < 2; sgn++) {
        if (sgn == 1 && a == 0) break;
        const int m = sgn ? -a : a;
        C2<double> c[6];
        c[0] = hess_coef(fac2, L.AR, n + 1, m, L.AR, n + 2, m, E, p);          // ZZ
        c[1] = hess_coef(fac2, L.BR, n + 1, m - 1, L.AR, n + 2, m - 1, E, p);  // ZU
        c[2] = hess_coef(fac2, L.CR, n + 1, m + 1, L.AR, n + 2, m + 1, E, p);  // ZD
        c[3] = hess_coef(fac2, L.BR, n + 1, m - 1, L.BR, n + 2, m - 2, E, p);  // UU
        c[4] = hess_coef(fac2, L.CR, n + 1, m + 1, L.BR, n + 2, m, E, p);      // UD
        c[5] = hess_coef(fac2, L.CR, n + 1, m + 1, L.CR, n + 2, m + 2, E, p);  // DD
#pragma unroll
        for (int k = 0; k < 6; k++) {
          if (sgn == 0) hess_fma(X[k], c[k], {0.0, 0.0}, h);
          else hess_fma(X[k], {0.0, 0.0}, c[k], h);
        }
      }
    }
  }
  if (!active) return;
  double* o = (kind ? hess1 : hess0) + 12 * (size_t)(n0 + i);
  hess_cartesian(X, ih * ih, o);
}

// ---- far field, list 3: per work item (box, kind, first point) one point
// per lane, the multipoles of the box's list 3 in order; second-derivative
// expansions of degree 2..p+2 formed on the fly (wave-uniform loads)
__global__ __launch_bounds__(256) void k_m2p_hess(DevPlan P, const double* __restrict__ fac2,
                                                  const double* __restrict__ mp, const int* __restrict__ items,
                                                  int nitems, const double* __restrict__ spos,
                                                  const double* __restrict__ tpos, double* __restrict__ hess0,
                                                  double* __restrict__ hess1) {
  const int p = P.p;
  const FacLayout L(p + 1);
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= nitems) return;
  const int b = items[3 * w], kind = items[3 * w + 1], start = items[3 * w + 2];
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + (int)(threadIdx.x & 63);
  const bool active = i < nn;
  double x = 0.0, y = 0.0, z = 0.0;
  if (active) {
    const double* q = (kind ? tpos : spos) + 4 * (size_t)(n0 + i);
    x = q[0]; y = q[1]; z = q[2];
  }
  const int pd = p + 2;
  double out[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int jj = P.l3_off[b]; jj < P.l3_off[b + 1]; jj++) {
    const int j = P.l3_idx[jj];
    if (P.nsrc[j] == 0) continue;
    const double* M = mp + 2 * (size_t)j * P.K;
    const double ih = P.inv_side[P.level[j]];
    // inactive lanes evaluate at a harmless point and are not written
    const double ux = active ? (x - P.center[3 * j]) * ih : 1.0;
    const double uy = active ? (y - P.center[3 * j + 1]) * ih : 1.0;
    const double uz = active ? (z - P.center[3 * j + 2]) * ih : 1.0;
    const double r2 = ux * ux + uy * uy + uz * uz;
    const double ir2 = 1.0 / r2, ir = sqrt(ir2);
    C2<double> X[6];
#pragma unroll
    for (int k = 0; k < 6; k++) X[k] = {0.0, 0.0};
    C2<double> hmm = {ir, 0.0};
    for (int a = 0; a <= pd; a++) {
      if (a > 0) {
        const double f = fac2[L.MM + a] * ir2;
        const C2<double> v = {-ux * f, -uy * f};
        hmm = cmul(v, hmm);
      }
      C2<double> h0 = hmm, h1 = hmm;
      for (int n = a; n <= pd; n++) {
        C2<double> h;
        if (n == a) {
          h = hmm;
        } else if (n == a + 1) {
          const double g = fac2[L.M1 + a] * uz * ir2;
          h = {g * hmm.re, g * hmm.im};
          h1 = h;
        } else {
          const double aa = (2 * n - 1) * uz, bq = fac2[L.RA + hidx(n, a)], cc = fac2[L.RB + hidx(n, a)] * ir2;
          h = {(aa * h1.re - bq * h0.re) * cc, (aa * h1.im - bq * h0.im) * cc};
          h0 = h1;
          h1 = h;
        }
        if (n < 2) continue;
        for (int sgn = 0; sgn < 2; sgn++) {
          if (sgn == 1 && a == 0) break;
          const int m = sgn ? -a : a;
          C2<double> c[6];
          c[0] = hess_coef(fac2, L.AS, n - 1, m, L.AS, n - 2, m, M, p);          // ZZ
          c[1] = hess_coef(fac2, L.BS, n - 1, m - 1, L.AS, n - 2, m - 1, M, p);  // ZU
          c[2] = hess_coef(fac2, L.CS, n - 1, m + 1, L.AS, n - 2, m + 1, M, p);  // ZD
          c[3] = hess_coef(fac2, L.BS, n - 1, m - 1, L.BS, n - 2, m - 2, M, p);  // UU
          c[4] = hess_coef(fac2, L.CS, n - 1, m + 1, L.BS, n - 2, m, M, p);      // UD
          c[5] = hess_coef(fac2, L.CS, n - 1, m + 1, L.CS, n - 2, m + 2, M, p);  // DD
#pragma unroll
          for (int k = 0; k < 6; k++) {
            if (sgn == 0) hess_fma(X[k], c[k], {0.0, 0.0}, h);
            else hess_fma(X[k], {0.0, 0.0}, c[k], h);
          }
        }
      }
    }
    hess_cartesian(X, ih * ih * ih, out);
  }
  if (!active) return;
  double* o = (kind ? hess1 : hess0) + 12 * (size_t)(n0 + i);
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] += out[k];
}

}  // namespace fmm3d
