// Laplace quadrupole sources (FP64): gather, P2M / P2L and the near field.
//
// A quadrupole at y with complex strength quadstr and real components
// quadvec = (Qxx, Qyy, Qzz, Qxy, Qxz, Qyz) has the potential
//   pot(x) = quadstr sum_k quadvec_k V_k(x - y),  V_ab = d_a d_b (1/r)
// (V_xx = -1/r^3 + 3 dx^2/r^5, V_xy = 3 dx dy/r^5, ...) and fld = -grad pot,
// as src/laprouts3d.f:8691 (lpotfld3dhess_qp) defines it.
//
// Its expansions (src/laprouts3d.f:7914 l3dformmp_qp_trunc, :8261
// l3dformta_qp_trunc) are those of a charge differentiated twice in the
// source position.  With d+ = d/dx + i d/dy, d- = d/dx - i d/dy,
//   D_Q = c_{+2} d+^2 + c_{-2} d-^2 + c_0 dz^2 + c_{+1} dz d+ + c_{-1} dz d-,
//   c_{+-2} = (Qxx - Qyy)/4 -+ i Qxy/4,  c_0 = Qzz - (Qxx + Qyy)/2,
//   c_{+-1} = Qxz/2 -+ i Qyz/2
// (d+ d- = -dz^2 on harmonic functions takes up the trace), and the ladder
// relations of the solid harmonics (R_n^m = r^n Yhat_n^m, S_n^m =
// Yhat_n^m / r^{n+1}, the (-1)^m phase of the expansions):
//   dz R_n^m = sqrt((n-m)(n+m)) R_{n-1}^m
//   d+ R_n^m = s(m)  sqrt((n-m)(n-m-1)) R_{n-1}^{m+1}
//   d- R_n^m = s'(m) sqrt((n+m)(n+m-1)) R_{n-1}^{m-1}
//   dz S_n^m = -sqrt((n+1-m)(n+1+m)) S_{n+1}^m
//   d+ S_n^m = s(m)  sqrt((n+m+1)(n+m+2)) S_{n+1}^{m+1}
//   d- S_n^m = s'(m) sqrt((n-m+1)(n-m+2)) S_{n+1}^{m-1}
// with s(m) = 1 for m >= 0 (else -1) and s'(m) = 1 for m <= 0 (else -1), give
// D_Q R_n^m = sum_j c_j L_j(n,m) R_{n-2}^{m+j} and D_Q S_n^m = sum_j c_j
// L'_j(n,m) S_{n+2}^{m+j}, j = -2..2.  For the channel charges q_j =
// quadstr conj(c_j) and u = (y - c)/h (expansions carry 1/side):
//   M_n^m += h^-2 sum_j L_j(n,m)  q_j conj(R_{n-2}^{m+j}(u)),
//   L_n^m += h^-3 sum_j L'_j(n,m) q_j conj(S_{n+2}^{m+j}(u)).
#pragma once
#include <hip/hip_runtime.h>

#include "kern_common.hpp"

namespace fmm3d {

// Sorted quadrupoles: 8 doubles each: quadstr (re, im), Qxx, Qyy, Qzz, and
// Qxy/2, Qxz/2, Qyz/2 (the off-diagonal entries of the symmetric matrix Q;
// halving is exact).
static __global__ void k_gather_qp(int ns, const int* __restrict__ perm, const double* __restrict__ qs,
                                   const double* __restrict__ qv, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int o = perm[i];
  double* d = out + 8 * (size_t)i;
  d[0] = qs[2 * (size_t)o];
  d[1] = qs[2 * (size_t)o + 1];
  for (int k = 0; k < 3; k++) d[2 + k] = qv[6 * (size_t)o + k];
  for (int k = 3; k < 6; k++) d[2 + k] = 0.5 * qv[6 * (size_t)o + k];
}

__device__ __forceinline__ double qp_sp(int m) { return m >= 0 ? 1.0 : -1.0; }
__device__ __forceinline__ double qp_sm(int m) { return m <= 0 ? 1.0 : -1.0; }

// L_j(n,m) (IRREG false) or L'_j(n,m) (true), index j + 2
template <bool IRREG>
__device__ __forceinline__ void qp_ladder(int n, int m, double (&Lj)[5]) {
  auto sq = [](double a) { return a > 0.0 ? sqrt(a) : 0.0; };
  if (IRREG) {
    const double az0 = -sqrt((double)(n + 1 - m) * (n + 1 + m));
    const double bp0 = qp_sp(m) * sqrt((double)(n + m + 1) * (n + m + 2));
    const double bm0 = qp_sm(m) * sqrt((double)(n - m + 1) * (n - m + 2));
    Lj[4] = bp0 * qp_sp(m + 1) * sqrt((double)(n + m + 3) * (n + m + 4));
    Lj[0] = bm0 * qp_sm(m - 1) * sqrt((double)(n - m + 3) * (n - m + 4));
    Lj[2] = az0 * -sqrt((double)(n + 2 - m) * (n + 2 + m));
    Lj[3] = bp0 * -sqrt((double)(n + 1 - m) * (n + 3 + m));
    Lj[1] = bm0 * -sqrt((double)(n + 3 - m) * (n + 1 + m));
  } else {
    // factors vanish where the ladder leaves the degree (|m'| > n')
    const double az0 = sq((double)(n - m) * (n + m));
    const double bp0 = (n - m >= 1) ? qp_sp(m) * sq((double)(n - m) * (n - m - 1)) : 0.0;
    const double bm0 = (n + m >= 1) ? qp_sm(m) * sq((double)(n + m) * (n + m - 1)) : 0.0;
    Lj[4] = (n - m - 2 >= 0) ? bp0 * qp_sp(m + 1) * sq((double)(n - m - 2) * (n - m - 3)) : 0.0;
    Lj[0] = (n + m - 2 >= 0) ? bm0 * qp_sm(m - 1) * sq((double)(n + m - 2) * (n + m - 3)) : 0.0;
    Lj[2] = az0 * sq((double)(n - 1 - m) * (n - 1 + m));
    Lj[3] = bp0 * sq((double)(n - 2 - m) * (n + m));
    Lj[1] = bm0 * sq((double)(n - m) * (n - 2 + m));
  }
}

// LDS doubles of k_qp_form for a tile of ts sources (table degree pt)
inline size_t qp_form_lds_doubles(int p, int pt, int ts) {
  const size_t H = (size_t)(pt + 1) * (pt + 2) / 2;
  return 2 * (size_t)(p + 1) * (p + 1) + (size_t)ts * (2 * H + 13);
}

// P2M (IRREG false: the leaf's own quadrupoles, regular harmonics to degree
// p - 2) or P2L (true: the quadrupoles of list 4, irregular harmonics to
// degree p + 2) of one box per workgroup, added into its expansion.  fac2:
// the factor table of order p + 1 (degrees up to p + 2).  Tiles of ts
// sources: phase A builds their harmonic half tables in LDS (one thread per
// (source, order) column), phase B contracts them, one thread per (n, m).
template <bool IRREG>
__global__ __launch_bounds__(256) void k_qp_form(DevPlan P, const double* __restrict__ fac2,
                                                 const double* __restrict__ spos, const double* __restrict__ qp,
                                                 int ts, double* __restrict__ expn) {
  extern __shared__ unsigned char smem_raw[];
  double* smem = reinterpret_cast<double*>(smem_raw);
  const int p = P.p, K = P.K;
  const int pt = IRREG ? p + 2 : p - 2;
  const int H = (pt + 1) * (pt + 2) / 2;
  const FacLayout L(p + 1);
  double* acc = smem;                     // 2K
  double* tab = acc + 2 * K;              // ts * H complex
  double* qc = tab + 2 * (size_t)ts * H;  // ts * 10: q_j, j = -2..2
  double* pos = qc + 10 * ts;             // ts * 3 scaled offsets
  const int b = IRREG ? P.p2l_boxes[blockIdx.x] : P.p2m_boxes[blockIdx.x];
  const double ih = P.inv_side[P.level[b]];
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) acc[k] = 0.0;
  const int jbeg = IRREG ? P.l4_off[b] : 0, jend = IRREG ? P.l4_off[b + 1] : 1;
  for (int jj = jbeg; jj < jend; jj++) {
    const int j = IRREG ? P.l4_idx[jj] : b;
    const int s0 = P.isrc[j], sn = P.nsrc[j];
    for (int t0 = 0; t0 < sn; t0 += ts) {
      const int tn = min(ts, sn - t0);
      __syncthreads();
      for (int s = threadIdx.x; s < tn; s += blockDim.x) {
        const double* x = spos + 4 * (size_t)(s0 + t0 + s);
        const double* q = qp + 8 * (size_t)(s0 + t0 + s);
        pos[3 * s] = (x[0] - cx) * ih;
        pos[3 * s + 1] = (x[1] - cy) * ih;
        pos[3 * s + 2] = (x[2] - cz) * ih;
        const double xx = q[2], yy = q[3], zz = q[4], xy = 2.0 * q[5], xz = 2.0 * q[6], yz = 2.0 * q[7];
        // conj(c_j), j = -2..2
        const double cr[5] = {0.25 * (xx - yy), 0.5 * xz, zz - 0.5 * (xx + yy), 0.5 * xz, 0.25 * (xx - yy)};
        const double ci[5] = {-0.25 * xy, -0.5 * yz, 0.0, 0.5 * yz, 0.25 * xy};
        for (int k = 0; k < 5; k++) {
          qc[10 * s + 2 * k] = q[0] * cr[k] - q[1] * ci[k];
          qc[10 * s + 2 * k + 1] = q[0] * ci[k] + q[1] * cr[k];
        }
      }
      __syncthreads();
      for (int task = threadIdx.x; task < tn * (pt + 1); task += blockDim.x) {
        const int s = task / (pt + 1), m = task - s * (pt + 1);
        const double ux = pos[3 * s], uy = pos[3 * s + 1], uz = pos[3 * s + 2];
        const double r2 = ux * ux + uy * uy + uz * uz;
        if (IRREG) {
          const double ir2 = 1.0 / r2, ir = sqrt(ir2);
          irregular_column<double>(fac2, L, pt, m, ux, uy, uz, ir2, ir, tab + 2 * (size_t)s * H, 1);
        } else {
          regular_column<double>(fac2, L, pt, m, ux, uy, uz, r2, tab + 2 * (size_t)s * H, 1);
        }
      }
      __syncthreads();
      for (int k = threadIdx.x; k < K; k += blockDim.x) {
        int n = (int)sqrtf((float)k);
        while (n * n > k) n--;
        while ((n + 1) * (n + 1) <= k) n++;
        const int m = k - n * n - n;
        const int kd = IRREG ? n + 2 : n - 2;
        if (kd < 0) continue;
        double Lj[5];
        qp_ladder<IRREG>(n, m, Lj);
        C2<double> a = {0.0, 0.0};
        for (int s = 0; s < tn; s++) {
          const double* tb = tab + 2 * (size_t)s * H;
#pragma unroll
          for (int jq = 0; jq < 5; jq++) {
            const int mm = m + jq - 2;
            if (mm < -kd || mm > kd || Lj[jq] == 0.0) continue;
            // conj(H_kd^mm) = H_kd^{-mm}
            const C2<double> h = half_get(tb, kd, -mm, 1);
            const C2<double> q = {Lj[jq] * qc[10 * s + 2 * jq], Lj[jq] * qc[10 * s + 2 * jq + 1]};
            cfma(a, q, h);
          }
        }
        acc[2 * k] += a.re;
        acc[2 * k + 1] += a.im;
      }
    }
  }
  __syncthreads();
  const double sc = IRREG ? ih * ih * ih : ih * ih;
  double* out = expn + 2 * (size_t)b * K;
  for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) out[k] += acc[k] * sc;
}

// Near field of quadrupoles for the P2P work items (the self box and list 1),
// one point per lane, sources in order, the self pair excluded by index;
// pot and fld (when fld is not null) added to the point's near-field sums.
__global__ __launch_bounds__(256) void k_p2p_qp(DevPlan P, const double* __restrict__ spos,
                                                const double* __restrict__ qp, const double* __restrict__ tpos,
                                                int kind, double* __restrict__ pot, double* __restrict__ fld) {
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
  double pr = 0, pi = 0, fr[3] = {0, 0, 0}, fi[3] = {0, 0, 0};
  auto add = [&](int s) {
    const double* sp = spos + 4 * (size_t)s;
    const double* q = qp + 8 * (size_t)s;
    const double dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir = rsqrt_t<double>(r2), ir2 = ir * ir;
    const double d3 = ir * ir2, d5 = d3 * ir2, d7 = d5 * ir2;
    const double xx = q[2], yy = q[3], zz = q[4], hxy = q[5], hxz = q[6], hyz = q[7];
    // Q d, d.Q.d and the trace of Q
    const double qdx = fma(xx, dx, fma(hxy, dy, hxz * dz));
    const double qdy = fma(yy, dy, fma(hxy, dx, hyz * dz));
    const double qdz = fma(zz, dz, fma(hxz, dx, hyz * dy));
    const double dqd = fma(dx, qdx, fma(dy, qdy, dz * qdz)), tr = xx + yy + zz;
    // pot = 3 d.Q.d / r^5 - tr / r^3; grad = 6 Q d / r^5 - 15 d.Q.d d / r^7 + 3 tr d / r^5
    const double a5 = 3.0 * d5;
    const double v = fma(dqd, a5, -tr * d3);
    const double ga = fma(5.0 * dqd, d7, -tr * d5) * 3.0;
    const double b5 = 2.0 * a5;
    const double gx = fma(b5, qdx, -ga * dx), gy = fma(b5, qdy, -ga * dy), gz = fma(b5, qdz, -ga * dz);
    pr = fma(q[0], v, pr);
    pi = fma(q[1], v, pi);
    fr[0] = fma(-q[0], gx, fr[0]); fi[0] = fma(-q[1], gx, fi[0]);
    fr[1] = fma(-q[0], gy, fr[1]); fi[1] = fma(-q[1], gy, fi[1]);
    fr[2] = fma(-q[0], gz, fr[2]); fi[2] = fma(-q[1], gz, fi[2]);
  };
  {
    const int s0 = P.isrc[b], s1 = s0 + P.nsrc[b];
    for (int s = s0; s < s1; s++)
      if (!(kind == 0 && s == me)) add(s);
  }
  for (int jj = P.l1_off[b]; jj < P.l1_off[b + 1]; jj++) {
    const int j = P.l1_idx[jj];
    const int s0 = P.isrc[j], s1 = s0 + P.nsrc[j];
#pragma unroll 2
    for (int s = s0; s < s1; s++) add(s);
  }
  if (!active) return;
  pot[2 * (size_t)me] += pr;
  pot[2 * (size_t)me + 1] += pi;
  if (fld) {
    double* o = fld + 6 * (size_t)me;
    for (int k = 0; k < 3; k++) {
      o[2 * k] += fr[k];
      o[2 * k + 1] += fi[k];
    }
  }
}

}  // namespace fmm3d
