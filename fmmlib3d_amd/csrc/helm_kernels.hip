// Helmholtz particle-FMM kernels for gfx950 (MI355X), FP64.
//
// Expansions use the reference's convention (src/helmrouts3d.f:3119 and
// :3438): M_n^m = ik sum q j_n(kr)/sc^n Y_n^{-m}(u), L_n^m = ik sum q h_n(kr)
// sc^n Y_n^{-m}(u), Y_n^m = sqrt(2n+1) Yhat_n^m, sc = the level's rscale.
// Each box stores (pmax+1)^2 coefficients and uses the first (p_l+1)^2,
// p_l = the order of its level.  Gradients use the exact identities
//   dz f_n Yhat_n^m = k/(2n+1) [AR(n,m) f_{n-1} Yhat_{n-1}^m + AS(n,m) f_{n+1} Yhat_{n+1}^m]
//   U  f_n Yhat_n^m = k/(2n+1) [BR(n,m) f_{n-1} Yhat_{n-1}^{m+1} + BS(n,m) f_{n+1} Yhat_{n+1}^{m+1}]
//   D  f_n Yhat_n^m = k/(2n+1) [CR(n,m) f_{n-1} Yhat_{n-1}^{m-1} + CS(n,m) f_{n+1} Yhat_{n+1}^{m-1}]
// (f = j or h, U = d/dx + i d/dy, D = d/dx - i d/dy, AR..CS the factor table
// of the Laplace kernels), so the field of an expansion is itself a short
// sum over the same basis.
#include <hip/hip_runtime.h>

#include "kern_common.hpp"
#include "m2l_grp.hpp"

namespace fmm3d {

using Cd = C2<double>;

// Helmholtz-specific device data (all pointers into the uploaded tables).
struct HelmDev {
  double kre, kim;
  int nd;                  // distance classes of the list-2 offsets
  const int* plev;         // order per level
  const double* sclev;     // rscale per level
  const int* ptrunc;       // (nlev+1) x 343 list-2 orders
  const long* coax;        // per table: base, pin, pout
  const double* z;         // coaxial blocks, complex
  // grouped M2L (k_hm2l_grp): rotation tiles and work items at pmax, and the
  // coaxial M2L tables in complex 4x4 tiles, (pmax+1) offsets per table
  const double* rot_half;
  const long* rot_half_off;
  const int *grp_ritems, *grp_rcount, *grp_zitems, *grp_zoff;
  const double* zt;
  const long* zt_off;
};

__device__ __forceinline__ Cd cadd2(Cd a, Cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cd csub2(Cd a, Cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cd cscale(Cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ double cnorm(Cd a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ Cd cdiv(Cd a, Cd b) {
  const double d = 1.0 / cnorm(b);
  return {(a.re * b.re + a.im * b.im) * d, (a.im * b.re - a.re * b.im) * d};
}

// h_i(z) sc^i, i = 0..n, upward recurrence (src/helmrouts3d.f:690 h3dall)
__device__ void dev_hankel(int n, Cd z, double sc, Cd* out, int st) {
  if (cnorm(z) < 1.0e-30) {
    for (int i = 0; i <= n; i++) out[(size_t)i * st] = {0.0, 0.0};
    return;
  }
  double s, c;
  sincos(z.re, &s, &c);
  const double e = exp(-z.im);
  const Cd h0 = cdiv(Cd{e * c, e * s}, Cd{-z.im, z.re});   // e^{iz}/(iz)
  out[0] = h0;
  if (n < 1) return;
  const Cd zi = cdiv(Cd{1.0, 0.0}, z);
  Cd h1 = cmul(h0, Cd{zi.re, zi.im - 1.0});
  h1 = cscale(h1, sc);
  out[st] = h1;
  const Cd zinv = cscale(zi, sc);
  const double sc2 = sc * sc;
  Cd hm = h0, hc = h1;
  for (int i = 1; i < n; i++) {
    const Cd zt = cscale(zinv, 2.0 * i + 1.0);
    const Cd hn = csub2(cmul(zt, hc), cscale(hm, sc2));
    out[(size_t)(i + 1) * st] = hn;
    hm = hc;
    hc = hn;
  }
}

// j_i(z)/sc^i, i = 0..n (1 <= n <= 127), Miller's backward recurrence as
// src/cdjseval3d.f:15 jfuns3d: start where the forward recurrence passes
// 1e20, rescale by 1e-40 whenever the backward sweep would overflow,
// normalise with j_0 or j_1.
__device__ void dev_bessel(int n, Cd z, double sc, Cd* out, int st) {
  if (fabs(z.re) < 1e-200 && fabs(z.im) < 1e-200) {
    out[0] = {1.0, 0.0};
    for (int i = 1; i <= n; i++) out[(size_t)i * st] = {0.0, 0.0};
    return;
  }
  const Cd zinv = cdiv(Cd{1.0, 0.0}, z);
  Cd fm = {0.0, 0.0}, fc = {1.0, 0.0};
  int ntop = 0;
  for (int i = n; i <= n + 1000; i++) {
    const Cd t = csub2(cmul(cscale(zinv, 2.0 * i + 1.0), fc), fm);
    fm = fc;
    fc = t;
    if (cnorm(t) > 1.0e40) { ntop = i + 1; break; }
  }
  if (ntop == 0) {  // |z| beyond ~1000: outside any FMM box of this code
    for (int i = 0; i <= n; i++) out[(size_t)i * st] = {0.0, 0.0};
    return;
  }
  unsigned long long ev0 = 0, ev1 = 0;   // rescale events at index e (e <= 127)
  Cd fa = {0.0, 0.0}, fb = {1.0, 0.0};    // f[i+1], f[i] with i = ntop-1
  if (ntop - 1 <= n) out[(size_t)(ntop - 1) * st] = fb;
  for (int i = ntop - 1; i >= 1; i--) {
    Cd t = csub2(cmul(cscale(zinv, 2.0 * i + 1.0), fb), fa);
    if (cnorm(t) > 1.0e40) {
      fb = cscale(fb, 1.0e-40);
      t = cscale(t, 1.0e-40);
      if (i <= n) out[(size_t)i * st] = fb;
      if (i < 64) ev0 |= 1ull << i;
      else if (i < 128) ev1 |= 1ull << (i - 64);
    }
    if (i - 1 <= n) out[(size_t)(i - 1) * st] = t;
    fa = fb;
    fb = t;
  }
  const double scinv = 1.0 / sc;
  double sctot = 1.0;
  for (int i = 1; i <= n; i++) {
    sctot *= scinv;
    const int e = i - 1;
    if ((e < 64 && ((ev0 >> e) & 1)) || (e >= 64 && ((ev1 >> (e - 64)) & 1))) sctot *= 1.0e-40;
    out[(size_t)i * st] = cscale(out[(size_t)i * st], sctot);
  }
  // j0 = sin z / z, j1 = j0 / z - cos z / z
  double sx, cx;
  sincos(z.re, &sx, &cx);
  const double ch = cosh(z.im), sh = sinh(z.im);
  const Cd sz = {sx * ch, cx * sh}, cz = {cx * ch, -sx * sh};
  const Cd j0 = cmul(sz, zinv);
  const Cd j1 = csub2(cmul(j0, zinv), cmul(cz, zinv));
  Cd zs;
  if (cnorm(j1) > cnorm(j0)) zs = cdiv(j1, cscale(out[st], sc));
  else zs = cdiv(j0, out[0]);
  for (int i = 0; i <= n; i++) out[(size_t)i * st] = cmul(out[(size_t)i * st], zs);
}

// Yhat_n^{mu} from a half table of Yhat_n^m (m >= 0): conj for mu < 0
__device__ __forceinline__ Cd yget(const double* Y, int n, int mu) {
  if (mu >= 0) return cld(Y, hidx(n, mu));
  return cconj(cld(Y, hidx(n, -mu)));
}
// f_n Yhat_n^mu, zero outside 0 <= |mu| <= n <= nmax
__device__ __forceinline__ Cd fyget(const double* Y, const double* f, int nmax, int n, int mu) {
  if (n < 0 || n > nmax || mu > n || mu < -n) return {0.0, 0.0};
  return cmul(cld(f, n), yget(Y, n, mu));
}

// ---------------------------------------------------------------------------
// P2M (src/helmrouts3d.f:2972, :3580) and P2L (:3299, :3978).  One workgroup
// per box; sources in tiles of `ts`.  Phase A: per source, the unit-vector
// harmonics Yhat (one thread per column m) and the radial series f_n(k r)
// (one thread per source); phase B: one thread per coefficient.
// ---------------------------------------------------------------------------
template <bool IRREG>
__global__ void k_hform(DevPlan P, HelmDev Hd, const double* __restrict__ fac,
                        const double* __restrict__ spos, const double* __restrict__ str, int ifcharge,
                        int ifdipole, int ts, double* __restrict__ expn) {
  extern __shared__ unsigned char smem_raw[];
  double* smem = reinterpret_cast<double*>(smem_raw);
  const int b = IRREG ? P.p2l_boxes[blockIdx.x] : P.p2m_boxes[blockIdx.x];
  const int lev = P.level[b];
  if (lev < 2) return;
  const int pmaxg = P.p, Kg = P.K;
  const int p = Hd.plev[lev], K = (p + 1) * (p + 1);
  const int pt = ifdipole ? p + 1 : p;            // degrees the tile needs
  const int Hh = (pmaxg + 2) * (pmaxg + 3) / 2;   // half-table stride (fixed by the launch)
  const int Fn = pmaxg + 2;
  FacLayout L(pmaxg);
  double* acc = smem;                       // 2*Kg: charge part, then dipole bracket
  double* accd = acc + 2 * (size_t)Kg;      // 2*Kg
  double* tab = accd + 2 * (size_t)Kg;      // ts * Hh complex
  double* fser = tab + 2 * (size_t)ts * Hh; // ts * Fn complex
  double* sst = fser + 2 * (size_t)ts * Fn; // ts * 8: q, wz, wU, wD
  double* geo = sst + 8 * (size_t)ts;       // ts * 4: unit vector, r
  const double sc = Hd.sclev[lev];
  const double lo = IRREG ? sc : 1.0 / sc, hi = IRREG ? 1.0 / sc : sc;
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  for (int k = threadIdx.x; k < 4 * Kg; k += blockDim.x) acc[k] = 0.0;

  int jbeg = 0, jend = 1;
  if (IRREG) { jbeg = P.l4_off[b]; jend = P.l4_off[b + 1]; }
  for (int jj = jbeg; jj < jend; jj++) {
    const int j = IRREG ? P.l4_idx[jj] : b;
    const int s0 = P.isrc[j], sn = P.nsrc[j];
    for (int t0 = 0; t0 < sn; t0 += ts) {
      const int tn = min(ts, sn - t0);
      __syncthreads();
      for (int s = threadIdx.x; s < tn; s += blockDim.x) {
        const double* x = spos + 4 * (size_t)(s0 + t0 + s);
        const double* r = str + 8 * (size_t)(s0 + t0 + s);
        const double ux = x[0] - cx, uy = x[1] - cy, uz = x[2] - cz;
        const double rr = sqrt(ux * ux + uy * uy + uz * uz);
        const double ir = rr > 0.0 ? 1.0 / rr : 0.0;
        geo[4 * s] = rr > 0.0 ? ux * ir : 0.0;
        geo[4 * s + 1] = rr > 0.0 ? uy * ir : 0.0;
        geo[4 * s + 2] = rr > 0.0 ? uz * ir : 1.0;
        geo[4 * s + 3] = rr;
        sst[8 * s] = r[0]; sst[8 * s + 1] = r[1];
        sst[8 * s + 2] = r[6]; sst[8 * s + 3] = r[7];
        sst[8 * s + 4] = 0.5 * (r[2] + r[5]); sst[8 * s + 5] = 0.5 * (r[3] - r[4]);
        sst[8 * s + 6] = 0.5 * (r[2] - r[5]); sst[8 * s + 7] = 0.5 * (r[3] + r[4]);
      }
      __syncthreads();
      // phase A
      for (int task = threadIdx.x; task < tn * (pt + 2); task += blockDim.x) {
        const int s = task / (pt + 2), m = task - s * (pt + 2);
        if (m <= pt) {
          regular_column<double>(fac, L, pt, m, geo[4 * s], geo[4 * s + 1], geo[4 * s + 2], 1.0,
                                 tab + 2 * (size_t)s * Hh, 1);
        } else {
          const double rr = geo[4 * s + 3];
          const Cd zz = {Hd.kre * rr, Hd.kim * rr};
          Cd* f = reinterpret_cast<Cd*>(fser + 2 * (size_t)s * Fn);
          if (IRREG) {
            if (cnorm(zz) < 1.0e-30) {      // h3dformta_trunc0 skips such a source
              for (int i = 0; i <= pt; i++) f[i] = {0.0, 0.0};
            } else {
              dev_hankel(pt, zz, sc, f, 1);
            }
          } else {
            dev_bessel(max(pt, 1), zz, sc, f, 1);
          }
        }
      }
      __syncthreads();
      // phase B
      for (int k = threadIdx.x; k < K; k += blockDim.x) {
        int n = (int)sqrtf((float)k);
        while (n * n > k) n--;
        while ((n + 1) * (n + 1) <= k) n++;
        const int m = k - n * n - n, mu = -m;
        Cd a = {0.0, 0.0}, g = {0.0, 0.0};
        const double far_ = (double)fac[L.AR + eidx(n, mu)], fas = (double)fac[L.AS + eidx(n, mu)];
        const double fbr = n >= 1 ? (double)fac[L.BR + eidx(n, mu)] : 0.0, fbs = (double)fac[L.BS + eidx(n, mu)];
        const double fcr = n >= 1 ? (double)fac[L.CR + eidx(n, mu)] : 0.0, fcs = (double)fac[L.CS + eidx(n, mu)];
        for (int s = 0; s < tn; s++) {
          const double* Y = tab + 2 * (size_t)s * Hh;
          const double* f = fser + 2 * (size_t)s * Fn;
          if (ifcharge) cfma(a, Cd{sst[8 * s], sst[8 * s + 1]}, fyget(Y, f, pt, n, mu));
          if (ifdipole) {
            Cd tz = cadd2(cscale(fyget(Y, f, pt, n - 1, mu), far_ * lo), cscale(fyget(Y, f, pt, n + 1, mu), fas * hi));
            Cd tu = cadd2(cscale(fyget(Y, f, pt, n - 1, mu + 1), fbr * lo), cscale(fyget(Y, f, pt, n + 1, mu + 1), fbs * hi));
            Cd td = cadd2(cscale(fyget(Y, f, pt, n - 1, mu - 1), fcr * lo), cscale(fyget(Y, f, pt, n + 1, mu - 1), fcs * hi));
            cfma(g, Cd{sst[8 * s + 2], sst[8 * s + 3]}, tz);
            cfma(g, Cd{sst[8 * s + 4], sst[8 * s + 5]}, tu);
            cfma(g, Cd{sst[8 * s + 6], sst[8 * s + 7]}, td);
          }
        }
        acc[2 * k] += a.re; acc[2 * k + 1] += a.im;
        accd[2 * k] += g.re; accd[2 * k + 1] += g.im;
      }
    }
  }
  __syncthreads();
  // out += ik sqrt(2n+1) (charge part + k/(2n+1) dipole bracket)
  double* out = expn + 2 * (size_t)b * Kg;
  const Cd ik = {-Hd.kim, Hd.kre}, kk = {Hd.kre, Hd.kim};
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    Cd v = {acc[2 * k], acc[2 * k + 1]};
    Cd d = cmul(kk, Cd{accd[2 * k], accd[2 * k + 1]});
    v = cadd2(v, cscale(d, 1.0 / (2.0 * n + 1.0)));
    v = cscale(cmul(ik, v), (double)fac[L.M1 + n]);
    out[2 * k] += v.re;
    out[2 * k + 1] += v.im;
  }
}

// ---------------------------------------------------------------------------
// P2M / P2L with per-thread sums (the scheme of the Laplace k_form2).  With
// B_n^mu = f_n(k r) Yhat_n^mu(u) (f = j_n/sc^n for P2M, h_n sc^n for P2L),
// the 256 threads form G source groups of 256/G threads; each thread owns up
// to kHFormItems half-table entries (n', a >= 0) and accumulates, over the
// sources of its group, S^w(n', +-a) = sum_s w_s B_{n'}^{+-a} (Yhat^{-a} =
// conj Yhat^a) for w = q, v_x, v_y, v_z, the weights read with wave-uniform
// loads.  G = 4, 2 or 1: the most groups for which the entries of the box's
// order fit.  Sources are taken in tiles of ts (as many as the LDS holds at
// the box's order) that run across the boxes of list 4; per tile the
// unit-vector harmonics and the radial series are built in LDS.  The
// coefficients are then
//   out(n,m) += ik sqrt(2n+1) [S^q(n,mu) + k/(2n+1) (AR lo S^z(n-1,mu)
//     + AS hi S^z(n+1,mu) + BR lo S^u(n-1,mu+1) + BS hi S^u(n+1,mu+1)
//     + CR lo S^d(n-1,mu-1) + CS hi S^d(n+1,mu-1))],   mu = -m,
// with S^u, S^d = (S^x -+ i S^y)/2 (the same sums as k_hform, by linearity),
// combined through LDS in two passes (q and z, then x and y), the groups
// added in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kHFormItems = 4;   // half-table entries per thread of a group
constexpr int kHFormOut = 8;     // output coefficients per thread in the combine
constexpr int kHFormTile = 64;   // most sources per tile

// LDS head of the kernel: the factor tables [0, AR) and cm[m] = R_m^m /
// (-(ux + i uy))^m = prod_{k<=m} sqrt((2k-1)/(2k)), m = 0..pmax+2
__host__ __device__ inline size_t hform2_fac_bytes(int pmax) {
  return ((size_t)(FacLayout(pmax).AR + pmax + 3) * sizeof(double) + 15) / 16 * 16;
}

// One complex table entry as a single 16-byte LDS store
__device__ __forceinline__ void cst_v2(double* p, int k, Cd v) {
  double2 t;
  t.x = v.re;
  t.y = v.im;
  *reinterpret_cast<double2*>(p + 2 * k) = t;
}

// Column m of Yhat_n^m(u), n = m..pt, of a unit vector into a half table:
// Yhat_m^m = cm (-(ux + i uy))^m with the power taken by binary powers, then
// the three-term recurrence in n of regular_column (with r^2 = 1).
__device__ __forceinline__ void hform_column(const double* facl, const FacLayout& L, int pt, int m, double cm,
                                             double ux, double uy, double uz, double* out) {
  Cd base = {-ux, -uy}, r = {cm, 0.0};
  for (int e = m; e; e >>= 1) {
    if (e & 1) r = cmul(r, base);
    base = cmul(base, base);
  }
  cst_v2(out, hidx(m, m), r);
  if (m + 1 > pt) return;
  const double f1 = facl[L.M1 + m] * uz;
  Cd r1 = {f1 * r.re, f1 * r.im};
  cst_v2(out, hidx(m + 1, m), r1);
  Cd r0 = r;
  for (int n = m + 2; n <= pt; n++) {
    const double a = (double)(2 * n - 1) * uz, b = facl[L.RA + hidx(n, m)], c = facl[L.RB + hidx(n, m)];
    const Cd rn = {(a * r1.re - b * r0.re) * c, (a * r1.im - b * r0.im) * c};
    cst_v2(out, hidx(n, m), rn);
    r0 = r1;
    r1 = rn;
  }
}

template <bool IRREG, bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_hform2(DevPlan P, HelmDev Hd, const double* __restrict__ fac,
                                                const double* __restrict__ spos, const double* __restrict__ str,
                                                int region, double* __restrict__ expn) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int gidx[kHFormTile];
  const int b = IRREG ? P.p2l_boxes[blockIdx.x] : P.p2m_boxes[blockIdx.x];
  const int lev = P.level[b];
  if (lev < 2) return;
  const int tid = threadIdx.x;
  const int pmaxg = P.p, Kg = P.K;
  const int p = Hd.plev[lev];
  const int pt = DIP ? p + 1 : p;
  const int H = (pt + 1) * (pt + 2) / 2;
  const int P2 = (pt + 1) * (pt + 1);
  const int G = H <= 64 * kHFormItems ? 4 : (H <= 128 * kHFormItems ? 2 : 1);
  const int gs = 256 / G, grp = tid / gs, lt = tid - grp * gs;
  // region = doubles of LDS after the factor head; every order fits 4 sources
  // tile row stride: H entries padded so that consecutive sources start one
  // entry (4 banks) apart modulo the 32 LDS banks
  int HS = H;
  while ((HS * 4) % 32 != 4) HS++;
  const int ts = min(kHFormTile, region / (2 * (HS + pt + 1) + 8));
  FacLayout L(pmaxg);
  double* facl = reinterpret_cast<double*>(smem_raw);
  double* cm = facl + L.AR;
  for (int k = tid; k < L.AR; k += 256) facl[k] = fac[k];
  if (tid == 0) {
    double c = 1.0;
    cm[0] = 1.0;
    for (int k = 1; k <= pt; k++) { c *= fac[L.MM + k]; cm[k] = c; }
  }
  double* tab = reinterpret_cast<double*>(smem_raw + hform2_fac_bytes(pmaxg));  // ts x H complex
  double* fser = tab + 2 * (size_t)ts * HS;                                     // ts x (pt+1) complex
  double* wsh = fser + 2 * (size_t)ts * (pt + 1);                               // ts x 8 source weights
  const double sc = Hd.sclev[lev];
  const double lo = IRREG ? sc : 1.0 / sc, hi = IRREG ? 1.0 / sc : sc;
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  double acc[kHFormItems][16];
  int hn[kHFormItems];   // degree n' of each owned entry h = n'(n'+1)/2 + a
#pragma unroll
  for (int q = 0; q < kHFormItems; q++) {
    const int h = lt + q * gs;
    int n = (int)((sqrtf(8.0f * h + 1.0f) - 1.0f) * 0.5f);
    while (n * (n + 1) / 2 > h) n--;
    while ((n + 1) * (n + 2) / 2 <= h) n++;
    hn[q] = n;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = 0.0;
  }
  // the source boxes: list 4 of b for P2L, b itself for P2M
  const int jbeg = IRREG ? P.l4_off[b] : 0, jend = IRREG ? P.l4_off[b + 1] : 1;
  auto jbox = [&](int jj) { return IRREG ? P.l4_idx[jj] : b; };
  int cj = jbeg, co = 0;   // cursor: box position in the list, offset in the box
  while (cj < jend && P.nsrc[jbox(cj)] == 0) cj++;
  while (cj < jend) {
    // this tile: up to ts sources from the cursor on
    int tn = 0;
    {
      int nj = cj, no = co, rem = ts;
      while (nj < jend) {
        const int left = P.nsrc[jbox(nj)] - no;
        if (rem < left) { no += rem; tn += rem; rem = 0; break; }
        rem -= left; tn += left; nj++; no = 0;
      }
      __syncthreads();
      if (tid < tn) {
        int wj = cj, wo = co + tid;
        while (wo >= P.nsrc[jbox(wj)]) { wo -= P.nsrc[jbox(wj)]; wj++; }
        gidx[tid] = P.isrc[jbox(wj)] + wo;
      }
      cj = nj; co = no;
      while (cj < jend && co >= P.nsrc[jbox(cj)]) { cj++; co = 0; }
    }
    __syncthreads();
    // phase A: the columns m of Yhat on the unit vector (task m * tn + s,
    // so a wave holds few orders and its loop lengths are close), then the
    // radial series (tasks after them, in waves of their own)
    // the tile's weights (8 per source, two values per thread), read while
    // the tables are built and staged in LDS for the sums
    double wr0 = 0.0, wr1 = 0.0;
    if (tid < 8 * tn) wr0 = str[8 * (size_t)gidx[tid >> 3] + (tid & 7)];
    if (tid + 256 < 8 * tn) wr1 = str[8 * (size_t)gidx[(tid + 256) >> 3] + (tid & 7)];
    const int ncolt = (pt + 1) * tn;
    for (int task = tid; task < ncolt + tn; task += 256) {
      const int m = task < ncolt ? task / tn : -1;
      const int s = task < ncolt ? task - m * tn : task - ncolt;
      const double* x = spos + 4 * (size_t)gidx[s];
      const double ux0 = x[0] - cx, uy0 = x[1] - cy, uz0 = x[2] - cz;
      const double rr = sqrt(ux0 * ux0 + uy0 * uy0 + uz0 * uz0);
      if (m >= 0) {
        const double ir = rr > 0.0 ? 1.0 / rr : 0.0;
        const double ux = rr > 0.0 ? ux0 * ir : 0.0, uy = rr > 0.0 ? uy0 * ir : 0.0, uz = rr > 0.0 ? uz0 * ir : 1.0;
        hform_column(facl, L, pt, m, cm[m], ux, uy, uz, tab + 2 * (size_t)s * HS);
      } else {
        const Cd zz = {Hd.kre * rr, Hd.kim * rr};
        Cd* fcol = reinterpret_cast<Cd*>(fser + 2 * (size_t)s * (pt + 1));
        if (IRREG) {
          if (cnorm(zz) < 1.0e-30) {      // h3dformta_trunc0 skips such a source
            for (int i = 0; i <= pt; i++) fcol[i] = {0.0, 0.0};
          } else {
            dev_hankel(pt, zz, sc, fcol, 1);
          }
        } else {
          // the series needs two entries; the tile row holds pt+1
          if (pt >= 1) {
            dev_bessel(pt, zz, sc, fcol, 1);
          } else {
            Cd tmp[2];
            dev_bessel(1, zz, sc, tmp, 1);
            fcol[0] = tmp[0];
          }
        }
      }
    }
    if (tid < 8 * tn) wsh[tid] = wr0;
    if (tid + 256 < 8 * tn) wsh[tid + 256] = wr1;
    __syncthreads();
    // phase B: per-thread sums over the group's sources of the tile
    for (int s = grp; s < tn; s += G) {
      const double* w = wsh + 8 * (size_t)s;
      const double* Y = tab + 2 * (size_t)s * HS;
      const double* fr = fser + 2 * (size_t)s * (pt + 1);
#pragma unroll
      for (int q = 0; q < kHFormItems; q++) {
        const int h = lt + q * gs;
        if (h < H) {
          const int n = hn[q];
          const double2 yv = *reinterpret_cast<const double2*>(Y + 2 * h);
          const double yr = yv.x, yi = yv.y;
          const double fre = fr[2 * n], fim = fr[2 * n + 1];
          const double bpr = fre * yr - fim * yi, bpi = fre * yi + fim * yr;   // f Y^a
          const double bmr = fre * yr + fim * yi, bmi = fim * yr - fre * yi;   // f conj(Y^a)
#pragma unroll
          for (int c = 0; c < 4; c++) {
            if ((c == 0 && !CHG) || (c > 0 && !DIP)) continue;
            const double wr = w[2 * c], wi = w[2 * c + 1];
            double* A = &acc[q][4 * c];
            A[0] = fma(wr, bpr, fma(-wi, bpi, A[0]));
            A[1] = fma(wr, bpi, fma(wi, bpr, A[1]));
            A[2] = fma(wr, bmr, fma(-wi, bmi, A[2]));
            A[3] = fma(wr, bmi, fma(wi, bmr, A[3]));
          }
        }
      }
    }
  }
  // combine through LDS, two weights at a time: c = 0 (q) with 3 (v_z), then 1, 2 (v_x, v_y)
  const int K = (p + 1) * (p + 1);
  double* SA = tab;
  double* SB = tab + 2 * (size_t)P2;
  Cd part[kHFormOut], dzp[kHFormOut];
#pragma unroll
  for (int r = 0; r < kHFormOut; r++) { part[r] = {0.0, 0.0}; dzp[r] = {0.0, 0.0}; }
  auto sget = [&](const double* S, int n1, int m1) -> Cd {
    if (n1 < 0 || n1 > pt || m1 > n1 || m1 < -n1) return Cd{0.0, 0.0};
    return Cd{S[2 * eidx(n1, m1)], S[2 * eidx(n1, m1) + 1]};
  };
  auto faf = [&](int off, int n1, int m1) -> double { return fac[off + eidx(n1, m1)]; };
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int ca = pass == 0 ? 0 : 1, cb = pass == 0 ? 3 : 2;
    if (pass == 1 && !DIP) break;
    for (int g = 0; g < G; g++) {
      __syncthreads();
      if (grp == g) {
#pragma unroll
        for (int q = 0; q < kHFormItems; q++) {
          const int h = lt + q * gs;
          if (h < H) {
            const int n = hn[q];
            const int a = h - n * (n + 1) / 2;
            // (a = 0: the two sums are the same entry, taken once)
            const int ip = 2 * eidx(n, a), im = 2 * eidx(n, -a);
            if (g == 0) {
              SA[ip] = acc[q][4 * ca]; SA[ip + 1] = acc[q][4 * ca + 1];
              SB[ip] = acc[q][4 * cb]; SB[ip + 1] = acc[q][4 * cb + 1];
              if (a > 0) {
                SA[im] = acc[q][4 * ca + 2]; SA[im + 1] = acc[q][4 * ca + 3];
                SB[im] = acc[q][4 * cb + 2]; SB[im + 1] = acc[q][4 * cb + 3];
              }
            } else {
              SA[ip] += acc[q][4 * ca]; SA[ip + 1] += acc[q][4 * ca + 1];
              SB[ip] += acc[q][4 * cb]; SB[ip + 1] += acc[q][4 * cb + 1];
              if (a > 0) {
                SA[im] += acc[q][4 * ca + 2]; SA[im + 1] += acc[q][4 * ca + 3];
                SB[im] += acc[q][4 * cb + 2]; SB[im + 1] += acc[q][4 * cb + 3];
              }
            }
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kHFormOut; r++) {
      const int k = tid + r * 256;
      if (k >= K) continue;
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      const int mu = -(k - n * n - n);
      if (pass == 0) {
        if (CHG) part[r] = sget(SA, n, mu);
        if (DIP) {
          const Cd z1 = sget(SB, n - 1, mu), z2 = sget(SB, n + 1, mu);
          dzp[r] = cadd2(cscale(z1, faf(L.AR, n, mu) * lo), cscale(z2, faf(L.AS, n, mu) * hi));
        }
      } else if (DIP) {
        // S^u = (S^x - i S^y)/2, S^d = (S^x + i S^y)/2
        auto su = [&](int n1, int m1) {
          const Cd x = sget(SA, n1, m1), y = sget(SB, n1, m1);
          return Cd{0.5 * (x.re + y.im), 0.5 * (x.im - y.re)};
        };
        auto sd = [&](int n1, int m1) {
          const Cd x = sget(SA, n1, m1), y = sget(SB, n1, m1);
          return Cd{0.5 * (x.re - y.im), 0.5 * (x.im + y.re)};
        };
        const double fbr = n >= 1 ? faf(L.BR, n, mu) : 0.0, fcr = n >= 1 ? faf(L.CR, n, mu) : 0.0;
        Cd g = dzp[r];
        g = cadd2(g, cadd2(cscale(su(n - 1, mu + 1), fbr * lo), cscale(su(n + 1, mu + 1), faf(L.BS, n, mu) * hi)));
        g = cadd2(g, cadd2(cscale(sd(n - 1, mu - 1), fcr * lo), cscale(sd(n + 1, mu - 1), faf(L.CS, n, mu) * hi)));
        dzp[r] = g;
      }
    }
  }
  // out += ik sqrt(2n+1) (charge part + k/(2n+1) dipole bracket)
  double* out = expn + 2 * (size_t)b * Kg;
  const Cd ik = {-Hd.kim, Hd.kre}, kk = {Hd.kre, Hd.kim};
#pragma unroll
  for (int r = 0; r < kHFormOut; r++) {
    const int k = tid + r * 256;
    if (k >= K) continue;
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    Cd v = part[r];
    if (DIP) v = cadd2(v, cscale(cmul(kk, dzp[r]), 1.0 / (2.0 * n + 1.0)));
    v = cscale(cmul(ik, v), fac[L.M1 + n]);
    out[2 * k] += v.re;
    out[2 * k + 1] += v.im;
  }
}

// ---------------------------------------------------------------------------
// M2P / L2P from packed expansions.  For every box whose expansion is
// evaluated, the expansion Ew = sqrt(2n+1) E and the three gradient
// expansions Gz, Gu, Gd (the z-derivative and U, D of the field written as
// expansions of degree pd + 1 through the identities at the top of this
// file) are stored per half-table entry (n, a >= 0)
// in the order of the evaluation walk (column a, then n = a..pd): 8 complex
//   Ew(n,a) Ew(n,-a) Gz(n,a) Gz(n,-a) Gu(n,a) Gu(n,-a) Gd(n,a) Gd(n,-a)
// (the -a slots zero for a = 0; Gz, Gu, Gd zero without the gradient).  The
// evaluation kernel reads them with wave-uniform loads, one thread per point,
// the point's radial series in its own LDS column.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int hwalk_index(int pd, int a, int n) { return a * (pd + 1) - a * (a - 1) / 2 + (n - a); }

template <bool IRREG>
__global__ void k_hpack(DevPlan P, HelmDev Hd, const double* __restrict__ fac, const double* __restrict__ expn,
                        const int* __restrict__ boxes, const int* __restrict__ offs, int grad,
                        double* __restrict__ pack) {
  const int j = boxes[blockIdx.x];
  const int lev = P.level[j];
  const int p = Hd.plev[lev];
  const int pd = grad ? p + 1 : p;
  const int H = (pd + 1) * (pd + 2) / 2;
  FacLayout L(P.p);
  const double sc = Hd.sclev[lev];
  const double lo = IRREG ? sc : 1.0 / sc, hi = IRREG ? 1.0 / sc : sc;
  const Cd kk = {Hd.kre, Hd.kim};
  const double* E = expn + 2 * (size_t)j * P.K;
  double* out = pack + 16 * (size_t)offs[blockIdx.x];
  auto ew = [&](int n1, int m1) -> Cd {
    if (n1 < 0 || n1 > p || m1 > n1 || m1 < -n1) return Cd{0.0, 0.0};
    return cscale(Cd{E[2 * eidx(n1, m1)], E[2 * eidx(n1, m1) + 1]}, fac[L.M1 + n1]);
  };
  auto ek = [&](int n1, int m1) -> Cd {   // k Ew(n1, m1) / (2 n1 + 1)
    if (n1 < 0 || n1 > p || m1 > n1 || m1 < -n1) return Cd{0.0, 0.0};
    return cmul(kk, cscale(ew(n1, m1), 1.0 / (2.0 * n1 + 1.0)));
  };
  auto fa = [&](int off, int n1, int m1) -> double {
    if (n1 < 0 || n1 > p || m1 > n1 || m1 < -n1) return 0.0;
    return fac[off + eidx(n1, m1)];
  };
  for (int w = threadIdx.x; w < H; w += blockDim.x) {
    int a = 0;
    while (a < pd && hwalk_index(pd, a + 1, a + 1) <= w) a++;
    const int n = a + (w - hwalk_index(pd, a, a));
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = 0.0;
    for (int sg = 0; sg < (a > 0 ? 2 : 1); sg++) {
      const int m = sg ? -a : a;
      const Cd e0 = ew(n, m);
      v[2 * sg] = e0.re; v[2 * sg + 1] = e0.im;
      if (grad) {
        const Cd gz = cadd2(cscale(ek(n + 1, m), fa(L.AR, n + 1, m) * lo), cscale(ek(n - 1, m), fa(L.AS, n - 1, m) * hi));
        const Cd gu = cadd2(cscale(ek(n + 1, m - 1), fa(L.BR, n + 1, m - 1) * lo),
                            cscale(ek(n - 1, m - 1), fa(L.BS, n - 1, m - 1) * hi));
        const Cd gd = cadd2(cscale(ek(n + 1, m + 1), fa(L.CR, n + 1, m + 1) * lo),
                            cscale(ek(n - 1, m + 1), fa(L.CS, n - 1, m + 1) * hi));
        v[4 + 2 * sg] = gz.re; v[5 + 2 * sg] = gz.im;
        v[8 + 2 * sg] = gu.re; v[9 + 2 * sg] = gu.im;
        v[12 + 2 * sg] = gd.re; v[13 + 2 * sg] = gd.im;
      }
    }
    double* o = out + 16 * (size_t)w;
#pragma unroll
    for (int e = 0; e < 16; e++) o[e] = v[e];
  }
}

// out += c+ B + c- Bm
__device__ __forceinline__ void hpack_fma(Cd& out, const double* c, const Cd& B, const Cd& Bm) {
  out.re = fma(c[0], B.re, fma(-c[1], B.im, fma(c[2], Bm.re, fma(-c[3], Bm.im, out.re))));
  out.im = fma(c[0], B.im, fma(c[1], B.re, fma(c[2], Bm.im, fma(c[3], Bm.re, out.im))));
}

// One 64-thread workgroup per work item.  L2P (IRREG = false): item (box,
// kind, first point), the box's own local expansion, added to the outputs.
// M2P (IRREG = true): item (box, kind, first point, first and last+1 index
// into the box's list 3), a slice of list 3 whose sums go to `partial`
// (8 values per point), added to the outputs in slice order by k_slice_reduce.
// All sums stay in registers until the single store at the end.
template <bool IRREG>
__global__ __launch_bounds__(64) void k_heval3(DevPlan P, HelmDev Hd, const double* __restrict__ fac,
                                               const double* __restrict__ pack, const int* __restrict__ slot,
                                               const int* __restrict__ items, int grad,
                                               const double* __restrict__ spos, const double* __restrict__ tpos,
                                               int wpot0, int wfld0, int wpot1, int wfld1,
                                               double* __restrict__ pot0, double* __restrict__ fld0,
                                               double* __restrict__ pot1, double* __restrict__ fld1,
                                               double* __restrict__ partial) {
  constexpr int NT = 64;
  constexpr int IS = IRREG ? 5 : 3;
  extern __shared__ unsigned char smem_raw[];
  double* fcol = reinterpret_cast<double*>(smem_raw) + 2 * threadIdx.x;  // stride NT complex
  FacLayout L(P.p);
  const int b = items[IS * blockIdx.x], kind = items[IS * blockIdx.x + 1], start = items[IS * blockIdx.x + 2];
  const int wp = kind ? wpot1 : wpot0, wf = kind ? wfld1 : wfld0;
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + threadIdx.x;
  const bool active = i < nn;
  double x = 0.0, y = 0.0, z = 0.0;
  if (active) {
    const double* q = (kind ? tpos : spos) + 4 * (size_t)(n0 + i);
    x = q[0]; y = q[1]; z = q[2];
  }
  Cd pot = {0.0, 0.0}, gz = {0.0, 0.0}, gu = {0.0, 0.0}, gd = {0.0, 0.0};
  int jbeg = 0, jend = 1;
  if (IRREG) { jbeg = items[IS * blockIdx.x + 3]; jend = items[IS * blockIdx.x + 4]; }
  for (int jj = jbeg; jj < jend; jj++) {
    const int j = IRREG ? P.l3_idx[jj] : b;
    const int lev = P.level[j];
    if (lev < 2) continue;
    if (IRREG && P.nsrc[j] == 0) continue;
    const int p = Hd.plev[lev];
    const int pd = grad ? p + 1 : p;
    const double sc = Hd.sclev[lev];
    const double* pk = pack + 16 * (size_t)slot[j];
    const double ux0 = x - P.center[3 * j], uy0 = y - P.center[3 * j + 1], uz0 = z - P.center[3 * j + 2];
    const double rr = sqrt(ux0 * ux0 + uy0 * uy0 + uz0 * uz0);
    const Cd zz = {Hd.kre * rr, Hd.kim * rr};
    // h3dmpeval_trunc0 returns early for a point at the centre
    const bool skip = !active || (IRREG && cnorm(zz) < 1.0e-30);
    const double ir = rr > 0.0 ? 1.0 / rr : 0.0;
    const double ux = rr > 0.0 ? ux0 * ir : 0.0, uy = rr > 0.0 ? uy0 * ir : 0.0, uz = rr > 0.0 ? uz0 * ir : 1.0;
    // radial factors: M2P runs the upward recurrence of h_n(kr) sc^n
    // (dev_hankel) along each column from the column's first two values, so
    // nothing is stored; L2P keeps j_n/sc^n (backward recurrence) in LDS
    Cd cs0 = {0.0, 0.0}, cs1 = {0.0, 0.0}, zinv = {0.0, 0.0};
    const double sc2 = sc * sc;
    if (IRREG) {
      if (!skip) {
        double sn, cn;
        sincos(zz.re, &sn, &cn);
        const double e = exp(-zz.im);
        cs0 = cdiv(Cd{e * cn, e * sn}, Cd{-zz.im, zz.re});   // e^{iz}/(iz)
        const Cd zi = cdiv(Cd{1.0, 0.0}, zz);
        cs1 = cscale(cmul(cs0, Cd{zi.re, zi.im - 1.0}), sc);
        zinv = cscale(zi, sc);
      }
    } else {
      dev_bessel(max(pd, 1), zz, sc, reinterpret_cast<Cd*>(fcol), NT);
    }
    Cd tp = {0.0, 0.0}, tz = {0.0, 0.0}, tu = {0.0, 0.0}, td = {0.0, 0.0};
    Cd ymm = {1.0, 0.0};
    const double* c = pk;
    for (int m = 0; m <= pd; m++) {
      if (m > 0) {
        const double fm = fac[L.MM + m];
        ymm = cmul(Cd{-ux * fm, -uy * fm}, ymm);
      }
      Cd y0 = ymm, y1 = ymm;
      Cd hc = cs0, hn1 = cs1;  // h_n, h_{n+1} along the column
      if (IRREG && m > 0) {
        // advance the column start: (h_m, h_{m+1}) from (h_{m-1}, h_m)
        const Cd hnew = csub2(cmul(cscale(zinv, 2.0 * m + 1.0), cs1), cscale(cs0, sc2));
        cs0 = cs1;
        cs1 = hnew;
        hc = cs0;
        hn1 = cs1;
      }
      for (int n = m; n <= pd; n++, c += 16) {
        Cd yv;
        if (n == m) {
          yv = ymm;
        } else if (n == m + 1) {
          yv = cscale(ymm, fac[L.M1 + m] * uz);
          y1 = yv;
        } else {
          const double aa = (2.0 * n - 1.0) * uz, bq = fac[L.RA + hidx(n, m)], cc = fac[L.RB + hidx(n, m)];
          yv = {(aa * y1.re - bq * y0.re) * cc, (aa * y1.im - bq * y0.im) * cc};
          y0 = y1;
          y1 = yv;
        }
        Cd fn;
        if (IRREG) {
          fn = hc;
          const Cd hnew = csub2(cmul(cscale(zinv, 2.0 * n + 3.0), hn1), cscale(hc, sc2));
          hc = hn1;
          hn1 = hnew;
        } else {
          fn = {fcol[2 * (size_t)NT * n], fcol[2 * (size_t)NT * n + 1]};
        }
        const Cd B = {fn.re * yv.re - fn.im * yv.im, fn.re * yv.im + fn.im * yv.re};    // f_n Y_n^m
        const Cd Bm = {fn.re * yv.re + fn.im * yv.im, fn.im * yv.re - fn.re * yv.im};   // f_n Y_n^{-m}
        hpack_fma(tp, c, B, Bm);
        if (grad) {
          hpack_fma(tz, c + 4, B, Bm);
          hpack_fma(tu, c + 8, B, Bm);
          hpack_fma(td, c + 12, B, Bm);
        }
      }
    }
    if (!skip) {
      pot = cadd2(pot, tp);
      gz = cadd2(gz, tz);
      gu = cadd2(gu, tu);
      gd = cadd2(gd, td);
    }
  }
  if (IRREG) {
    double* o = partial + 8 * ((size_t)blockIdx.x * NT + threadIdx.x);
    o[0] = pot.re; o[1] = pot.im; o[2] = gz.re; o[3] = gz.im;
    o[4] = gu.re; o[5] = gu.im; o[6] = gd.re; o[7] = gd.im;
    return;
  }
  if (!active) return;
  double* po = kind ? pot1 : pot0;
  double* fo = kind ? fld1 : fld0;
  const size_t id = (size_t)(n0 + i);
  if (wp) { po[2 * id] += pot.re; po[2 * id + 1] += pot.im; }
  if (wf) {
    const Cd dx = {0.5 * (gu.re + gd.re), 0.5 * (gu.im + gd.im)};
    const Cd dy = {0.5 * (gu.im - gd.im), -0.5 * (gu.re - gd.re)};
    fo[6 * id + 0] -= dx.re; fo[6 * id + 1] -= dx.im;
    fo[6 * id + 2] -= dy.re; fo[6 * id + 3] -= dy.im;
    fo[6 * id + 4] -= gz.re; fo[6 * id + 5] -= gz.im;
  }
}

// ---------------------------------------------------------------------------
// Translations: phase, rotate, coaxial shift (complex packed block), rotate
// back, phase back (src/h3dtrans.f:211, :1229, :1998 apply the same
// operators by projection).  `acc` accumulates in LDS or global memory.
// ---------------------------------------------------------------------------
__device__ void hzshift_block(const HelmDev& Hd, int t, int pin, int pout, const double* in, double* out) {
  // orders m and -m share the block of |m|: one thread produces both
  const long base = Hd.coax[3 * t];
  const int tin = (int)Hd.coax[3 * t + 1], tout = (int)Hd.coax[3 * t + 2];
  const double* Z = Hd.z + 2 * base;
  const int Kh = (pout + 1) * (pout + 2) / 2;
  for (int k = threadIdx.x; k < Kh; k += blockDim.x) {
    int l = (int)((sqrtf(8.0f * k + 1.0f) - 1.0f) * 0.5f);
    while (l * (l + 1) / 2 > k) l--;
    while ((l + 1) * (l + 2) / 2 <= k) l++;
    const int am = k - l * (l + 1) / 2;
    Cd a = {0.0, 0.0}, b2 = {0.0, 0.0};
    if (am <= pin) {
      const long A = tout + 1, Bc = tin + 1, M = am;
      const long mb = M * A * Bc - (A + Bc) * M * (M - 1) / 2 + (M - 1) * M * (2 * M - 1) / 6;
      const double* zr = Z + 2 * (mb + (long)(l - am) * (tin + 1 - am));
      for (int n = am; n <= pin; n++) {
        const Cd zz = {zr[2 * (n - am)], zr[2 * (n - am) + 1]};
        cfma(a, zz, cld(in, eidx(n, am)));
        cfma(b2, zz, cld(in, eidx(n, -am)));
      }
    }
    out[2 * eidx(l, am)] = a.re;
    out[2 * eidx(l, am) + 1] = a.im;
    if (am > 0) {
      out[2 * eidx(l, -am)] = b2.re;
      out[2 * eidx(l, -am) + 1] = b2.im;
    }
  }
}

__device__ void hpoint_and_shoot(const DevTables& Tb, const HelmDev& Hd, int rc, const double* phase, int t,
                                 int pin, int pout, const double* in, double* out, double* A, double* B) {
  const int pm = pin > pout ? pin : pout;
  const int Ki = (pin + 1) * (pin + 1);
  for (int k = threadIdx.x; k < (pm + 1) * (pm + 1); k += blockDim.x) {
    if (k < Ki) {
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      const int m = k - n * n - n, am = m < 0 ? -m : m;
      double er = phase[2 * am], ei = phase[2 * am + 1];
      if (m < 0) ei = -ei;
      const double xr = in[2 * k], xi = in[2 * k + 1];
      A[2 * k] = xr * er - xi * ei;
      A[2 * k + 1] = xr * ei + xi * er;
    } else {
      A[2 * k] = 0.0;
      A[2 * k + 1] = 0.0;
    }
  }
  __syncthreads();
  rotate_block<double>(Tb, rc, pin, A, B, false);
  __syncthreads();
  hzshift_block(Hd, t, pin, pout, B, A);
  __syncthreads();
  rotate_block<double>(Tb, rc, pout, A, B, true);
  __syncthreads();
  const int Ko = (pout + 1) * (pout + 1);
  for (int k = threadIdx.x; k < Ko; k += blockDim.x) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    const int m = k - n * n - n, am = m < 0 ? -m : m;
    double er = phase[2 * am], ei = -phase[2 * am + 1];
    if (m < 0) ei = -ei;
    const double xr = B[2 * k], xi = B[2 * k + 1];
    out[2 * k] += xr * er - xi * ei;
    out[2 * k + 1] += xr * ei + xi * er;
  }
  __syncthreads();
}

// M2M into the parents of one level (src/hfmm3drouts.f:193-266)
__global__ void k_hm2m(DevPlan P, DevTables Tb, HelmDev Hd, int first, int tab, double* __restrict__ mp) {
  extern __shared__ unsigned char smem_raw[];
  double* A = reinterpret_cast<double*>(smem_raw);
  double* B = A + 2 * P.K;
  const int b = P.m2m_boxes[first + blockIdx.x];
  const int lev = P.level[b];
  const int pp = Hd.plev[lev], pc = Hd.plev[lev + 1];
  for (int c = 0; c < P.nchild[b]; c++) {
    const int ch = P.child[8 * b + c];
    if (P.nsrc[ch] == 0) continue;
    const int o = (P.ijk[3 * ch] & 1 ? 0 : 1) | (P.ijk[3 * ch + 1] & 1 ? 0 : 2) | (P.ijk[3 * ch + 2] & 1 ? 0 : 4);
    const double* ph = Tb.oct_phase + (size_t)(o * 2 + 0) * 2 * (Tb.p + 1);
    hpoint_and_shoot(Tb, Hd, Tb.oct_rot[o * 2 + 0], ph, tab, pc, pp, mp + 2 * (size_t)ch * P.K,
                     mp + 2 * (size_t)b * P.K, A, B);
  }
}

// L2L from the parents of one level (src/hfmm3drouts.f:483-556)
__global__ void k_hl2l(DevPlan P, DevTables Tb, HelmDev Hd, int first, int tab, double* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  double* A = reinterpret_cast<double*>(smem_raw);
  double* B = A + 2 * P.K;
  const int b = P.l2l_boxes[first + blockIdx.x];
  const int lev = P.level[b];
  const int pp = Hd.plev[lev], pc = Hd.plev[lev + 1];
  for (int c = 0; c < P.nchild[b]; c++) {
    const int ch = P.child[8 * b + c];
    const int o = (P.ijk[3 * ch] & 1 ? 0 : 1) | (P.ijk[3 * ch + 1] & 1 ? 0 : 2) | (P.ijk[3 * ch + 2] & 1 ? 0 : 4);
    const double* ph = Tb.oct_phase + (size_t)(o * 2 + 1) * 2 * (Tb.p + 1);
    hpoint_and_shoot(Tb, Hd, Tb.oct_rot[o * 2 + 1], ph, tab, pp, pc, lo + 2 * (size_t)b * P.K,
                     lo + 2 * (size_t)ch * P.K, A, B);
  }
}

// M2L over the pruned list 2 of each target box (src/hfmm3drouts.f:330-477);
// the target's local accumulates in LDS and is added to memory once.
template <bool LDSACC>
__global__ void k_hm2l(DevPlan P, DevTables Tb, HelmDev Hd, const double* __restrict__ mp,
                       double* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  double* A = reinterpret_cast<double*>(smem_raw);
  double* B = A + 2 * P.K;
  const int b = P.m2l_boxes[blockIdx.x];
  const int lev = P.level[b];
  const int p = Hd.plev[lev], K = (p + 1) * (p + 1);
  double* acc = LDSACC ? B + 2 * P.K : lo + 2 * (size_t)b * P.K;
  if (LDSACC)
    for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) acc[k] = 0.0;
  __syncthreads();
  for (int jj = P.l2_off[b]; jj < P.l2_off[b + 1]; jj++) {
    const int j = P.l2_idx[jj];
    const int di = P.ijk[3 * j] - P.ijk[3 * b] + 3, dj = P.ijk[3 * j + 1] - P.ijk[3 * b + 1] + 3,
              dk = P.ijk[3 * j + 2] - P.ijk[3 * b + 2] + 3;
    const int code = (di * 7 + dj) * 7 + dk;
    const int pt = Hd.ptrunc[lev * 343 + code];
    const int tab = lev * Hd.nd + Tb.off_dist[code];
    hpoint_and_shoot(Tb, Hd, Tb.off_rot[code], Tb.off_phase + (size_t)code * 2 * (Tb.p + 1), tab, pt, pt,
                     mp + 2 * (size_t)j * P.K, acc, A, B);
  }
  if (LDSACC) {
    double* out = lo + 2 * (size_t)b * P.K;
    for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) out[k] += acc[k];
  }
}

// ---------------------------------------------------------------------------
// M2L grouped by translation class (the scheme of the Laplace k_m2l_grp):
// the list-2 sources of a target box with the same (dx^2+dy^2, dz) share the
// rotation class, the coaxial table (per level and distance class) and the
// truncation order; up to kM2LGroup of them are translated together, the
// lanes of a work item holding the members.  The coaxial shift is a complex
// table (grp_z_item_c); no final scale (the tables carry the level's sc).
// src/hfmm3drouts.f:330-477 applies the same operators pair by pair.
// ---------------------------------------------------------------------------
template <int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_hm2l_grp(DevPlan P, DevTables Tb, HelmDev Hd, int gmax,
                                                   const double* __restrict__ mp, double* __restrict__ lo) {
  using T = double;
  extern __shared__ unsigned char smem_raw[];
  int* s_key = reinterpret_cast<int*>(smem_raw);
  int* s_box = s_key + kM2LChunk;
  int* s_code = s_box + kM2LChunk;
  int* t_key = s_code + kM2LChunk;
  int* t_box = t_key + kM2LChunk;
  int* t_code = t_box + kM2LChunk;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int b = P.m2l_boxes[xcd_block()];
  const int p = P.p, P1 = p + 1, K2 = 2 * P.K;
  // degree of each packed index k < (p+4)^2, and the members' phases
  // e^{i am phi}, am = 0..p, of the group in hand
  int* s_kn = t_code + kM2LChunk;
  T* s_ph = reinterpret_cast<T*>(s_kn + grp_kn_ints(p));
  T* acc = s_ph + 2 * kM2LGroup * P1;
  // members: (p+4)^2 complex each (degrees p+1..p+3 stay zero or hold the
  // zeros written for a truncated member), after a zero guard of 8 values
  const int MS = grp_member_stride<T>(p);
  T* A = acc + K2 + 8;
  const T ih = T(1);   // the coaxial tables carry the level scales
  const int lev = P.level[b];
  for (int k = tid; k < K2 + 8 + gmax * MS; k += bs) acc[k] = T(0);
  for (int k = tid; k < (p + 4) * (p + 4); k += bs) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    s_kn[k] = n;
  }
  const T* rtab = Hd.rot_half;
  const int off0 = P.l2_off[b], n2 = P.l2_off[b + 1] - off0;
  const int bi = P.ijk[3 * b], bj = P.ijk[3 * b + 1], bk = P.ijk[3 * b + 2];
  for (int c0 = 0; c0 < n2; c0 += kM2LChunk) {
    const int cn = min(kM2LChunk, n2 - c0);
    __syncthreads();
    if (tid < cn) {
      const int j = P.l2_idx[off0 + c0 + tid];
      const int di = P.ijk[3 * j] - bi, dj = P.ijk[3 * j + 1] - bj, dk = P.ijk[3 * j + 2] - bk;
      s_box[tid] = j;
      s_code[tid] = ((di + 3) * 7 + (dj + 3)) * 7 + (dk + 3);
      s_key[tid] = (di * di + dj * dj) * 7 + (dk + 3);
    }
    __syncthreads();
    if (tid < cn) {  // stable rank by key
      const int key = s_key[tid];
      int rank = 0;
      for (int t = 0; t < cn; t++) {
        const int kt = s_key[t];
        rank += (kt < key) || (kt == key && t < tid);
      }
      t_key[rank] = key;
      t_box[rank] = s_box[tid];
      t_code[rank] = s_code[tid];
    }
    __syncthreads();
    for (int s0 = 0; s0 < cn;) {
      const int key0 = t_key[s0];
      int g = 1;
      while (g < gmax && s0 + g < cn && t_key[s0 + g] == key0) g++;
      const int gs = g > 2 ? 4 : g;  // lanes per item
      const int lg = gs == 4 ? 2 : gs - 1;
      const int code0 = t_code[s0];
      const int pt = Hd.ptrunc[lev * 343 + code0];
      const int rc = Tb.off_rot[code0];
      const int tz = lev * Hd.nd + Tb.off_dist[code0];
      const long* zto = Hd.zt_off + (size_t)tz * P1;
      const int pz = (int)Hd.coax[3 * tz + 2];   // order of the table's tiles
      const int Kp = (pt + 1) * (pt + 1), Kz = (pt + 4) * (pt + 4);
      for (int idx = tid; idx < g * (pt + 1); idx += bs) {
        const int jj = idx / (pt + 1), am = idx - jj * (pt + 1);
        const double* ph = Tb.off_phase + (size_t)t_code[s0 + jj] * 2 * P1;
        s_ph[2 * (jj * P1 + am)] = (T)ph[2 * am];
        s_ph[2 * (jj * P1 + am) + 1] = (T)ph[2 * am + 1];
      }
      __syncthreads();
      // (a) members times e^{i m phi}; degrees pt+1..pt+3 set to zero
      for (int jj = 0; jj < g; jj++) {
        const T* src = mp + (size_t)t_box[s0 + jj] * K2;
        const T* phj = s_ph + 2 * jj * P1;
        T* Aj = A + jj * MS;
        for (int k = tid; k < Kz; k += bs) {
          T vr = 0, vi = 0;
          if (k < Kp) {
            const int n = s_kn[k];
            const int m = k - n * n - n, am = m < 0 ? -m : m;
            const T er = phj[2 * am], ei = m < 0 ? -phj[2 * am + 1] : phj[2 * am + 1];
            const T xr = src[2 * k], xi = src[2 * k + 1];
            vr = xr * er - xi * ei;
            vi = xr * ei + xi * er;
          }
          Aj[2 * k] = vr;
          Aj[2 * k + 1] = vi;
        }
      }
      __syncthreads();
      // Work items of a stage (rotation items by degree, z-shift items by
      // order; the two counts are equal): member jj of item u0 in the first
      // round and of item NI-1-u0 in the second, which pairs the cheapest
      // items with the most expensive ones.
      const int NI = Hd.grp_rcount[pt], U0 = bs >> lg;
      const int u0 = tid >> lg, jj = tid & (gs - 1);
      int uq[ITEMS];
      uq[0] = u0 < NI ? u0 : -1;
      if (ITEMS > 1) uq[ITEMS - 1] = NI - 1 - u0 >= U0 ? NI - 1 - u0 : -1;
      const bool mem = jj < g;
      T r[ITEMS][16];
      // (b) rotate to the z axis
#pragma unroll
      for (int q = 0; q < ITEMS; q++) {
#pragma unroll
        for (int e = 0; e < 16; e++) r[q][e] = T(0);
        if (uq[q] >= 0 && mem) {
          const int it = Hd.grp_ritems[uq[q]], n = it >> 8, qc = it & 255;
          grp_rot_item<T, false>(rtab + Hd.rot_half_off[rc * P1 + n] + 16 * qc, n / 4 + 1, n, qc, A + jj * MS, r[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ITEMS; q++)
        if (uq[q] >= 0 && mem) {
          const int it = Hd.grp_ritems[uq[q]];
          grp_rot_store<T>(it >> 8, 4 * (it & 255), r[q], A + jj * MS);
        }
      __syncthreads();
      // (c) shift along z
      const int zo = Hd.grp_zoff[pt];
#pragma unroll
      for (int q = 0; q < ITEMS; q++) {
#pragma unroll
        for (int e = 0; e < 16; e++) r[q][e] = T(0);
        if (uq[q] >= 0 && mem) {
          const int it = Hd.grp_zitems[zo + uq[q]], am = it >> 8;
          grp_z_item_c<T>(Hd.zt + zto[am], (pz - am) / 4 + 1, am, it & 255, pt, A + jj * MS, r[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ITEMS; q++)
        if (uq[q] >= 0 && mem) {
          const int it = Hd.grp_zitems[zo + uq[q]], am = it >> 8;
          grp_z_store<T>(am, am + 4 * (it & 255), pt, r[q], A + jj * MS);
        }
      __syncthreads();
      // (d) rotate back, apply e^{-i m phi} and the scale, sum the members
#pragma unroll
      for (int q = 0; q < ITEMS; q++) {
#pragma unroll
        for (int e = 0; e < 16; e++) r[q][e] = T(0);
        int n = 0, mp0 = 0;
        if (uq[q] >= 0) {
          const int it = Hd.grp_ritems[uq[q]];
          n = it >> 8;
          mp0 = 4 * (it & 255);
          if (mem) {
            grp_rot_item<T, true>(rtab + Hd.rot_half_off[rc * P1 + n] + 4 * mp0, n / 4 + 1, n, mp0 / 4,
                                  A + jj * MS, r[q]);
            const T* ph = s_ph + 2 * jj * P1;
#pragma unroll
            for (int c = 0; c < 4; c++) {
              const int mq = mp0 + c <= n ? mp0 + c : n;
              const T er = ph[2 * mq] * ih, ei = ph[2 * mq + 1] * ih;
              const T pr = r[q][4 * c], pi = r[q][4 * c + 1], qr = r[q][4 * c + 2], qi = r[q][4 * c + 3];
              r[q][4 * c] = pr * er + pi * ei;       // (pr + i pi) e^{-i m' phi}
              r[q][4 * c + 1] = pi * er - pr * ei;
              r[q][4 * c + 2] = qr * er - qi * ei;   // (qr + i qi) e^{+i m' phi}
              r[q][4 * c + 3] = qi * er + qr * ei;
            }
          }
        }
#pragma unroll
        for (int e = 0; e < 16; e++) {
          T v = r[q][e];
          if (gs > 1) v += __shfl_xor(v, 1);
          if (gs > 2) v += __shfl_xor(v, 2);
          r[q][e] = v;
        }
        if (uq[q] >= 0 && jj == 0) {
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const int mq = mp0 + c;
            if (mq > n) break;
            acc[2 * eidx(n, mq)] += r[q][4 * c];
            acc[2 * eidx(n, mq) + 1] += r[q][4 * c + 1];
            if (mq > 0) {
              acc[2 * eidx(n, -mq)] += r[q][4 * c + 2];
              acc[2 * eidx(n, -mq) + 1] += r[q][4 * c + 3];
            }
          }
        }
      }
      __syncthreads();
      s0 += g;
    }
  }
  __syncthreads();
  T* out = lo + (size_t)b * K2;
  for (int k = tid; k < K2; k += bs) out[k] += acc[k];
}

// ---------------------------------------------------------------------------
// P2P near field (src/hfmm3dpart.f:1092-1262 hfmm3dpart_direct_self/_targ with
// hpotfld3d / hpotfld3d_dp, src/helmrouts3d.f:1335, :1844):
//   charge  pot = q e^{ikr}/r,  fld = q (1-ikr) e^{ikr} x / r^3
//   dipole  pot = (v.x) (1-ikr) e^{ikr}/r^3,
//           fld = -(v (1-ikr) e^{ikr}/r^3 + (v.x) x e^{ikr} (k^2 r^2 + 3ikr - 3)/r^5)
// with v = dipstr dipvec.  Same work decomposition as the Laplace kernel.
// ---------------------------------------------------------------------------
template <typename TN>
__device__ __forceinline__ void sincos_t(TN x, TN* s, TN* c);
template <>
__device__ __forceinline__ void sincos_t<double>(double x, double* s, double* c) { sincos(x, s, c); }
template <>
__device__ __forceinline__ void sincos_t<float>(float x, float* s, float* c) { sincosf(x, s, c); }
template <typename TN>
__device__ __forceinline__ TN exp_t(TN x);
template <>
__device__ __forceinline__ double exp_t<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float exp_t<float>(float x) { return expf(x); }

// Offsets and r^2 in FP64 always; with TN = float the rest of the pair is
// FP32 (sums folded into FP64 by the caller every kFold sources), except for
// the closest pairs (r^2 < rc2, in practice r below an eighth of the leaf
// side): their terms dominate the sums at the points they are near, and the
// few FP32 roundings of such a pair (of v and d too, which a cancelling v.d
// multiplies) would be the largest error of the result.  add() leaves them
// out and reports them; the caller adds the close pairs of a flagged run in
// FP64 from the FP64 strengths (close_hpairs).
template <typename TN, bool CHG, bool DIP, bool FLD>
struct HPairAcc {
  TN a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // returns whether the pair is a close one (r^2 < rc2), which is then left out
  template <typename TS>
  __device__ __forceinline__ bool add(double kre_d, double kim_d, double xt, double yt, double zt,
                                      const double* __restrict__ sp, const TS* __restrict__ sts,
                                      bool mask, double rc2 = 0.0) {
    const double dxd = xt - sp[0], dyd = yt - sp[1], dzd = zt - sp[2];
    const double r2d = fma(dxd, dxd, fma(dyd, dyd, dzd * dzd));
    const bool close = sizeof(TN) == sizeof(float) && r2d < rc2;
    TN st[8];
    for (int k = 0; k < 8; k++) st[k] = (TN)sts[k];
    const TN dx = (TN)dxd, dy = (TN)dyd, dz = (TN)dzd, r2 = (TN)r2d;
    const TN kre = (TN)kre_d, kim = (TN)kim_d;
    TN ir = rsqrt_t<TN>(r2);
    // FP32: one Newton step on the hardware estimate (FP64 rsqrt_t is refined already)
    if (sizeof(TN) == sizeof(float)) ir = ir * (TN(1.5) - TN(0.5) * r2 * ir * ir);
    if (mask || close) ir = TN(0);                   // a source never acts on itself
    TN r, s, c;
    if (sizeof(TN) == sizeof(float)) {
      // The phase k r needs an absolute accuracy that FP32 cannot give once
      // k r is large (its rounding is k r 2^-24), so r comes from one FP64
      // Newton step on r^2 and the phase is reduced to [-pi, pi] in FP64,
      // then split into an FP32 head and tail: sin/cos of the head, first
      // order in the tail.  (Masked pair: ir = 0 gives r = 0.)
      const double ird = (double)ir;
      const double r0 = r2d * ird;
      const double rd = fma(0.5 * ird, fma(-r0, r0, r2d), r0);
      const double th = kre_d * rd;
      const double nrev = rint(th * 0.15915494309189535);     // th / (2 pi)
      const double thr = fma(-nrev, 2.4492935982947064e-16, fma(-nrev, 6.283185307179586, th));
      const TN hi = (TN)thr, lo = (TN)(thr - (double)hi);
      TN s0, c0;
      sincos_t<TN>(hi, &s0, &c0);
      s = fma(lo, c0, s0);
      c = fma(-lo, s0, c0);
      r = (TN)rd;
    } else {
      r = r2 * ir;
      sincos_t<TN>(kre * r, &s, &c);
    }
    const TN damp = kim == TN(0) ? TN(1) : exp_t<TN>(-kim * r);
    const TN gr = c * damp * ir, gi = s * damp * ir;   // G = e^{ikr}/r
    const TN ir2 = ir * ir;
    const TN omr = TN(1) + kim * r, omi = -kre * r;     // 1 - ikr
    const TN g_r = (omr * gr - omi * gi) * ir2, g_i = (omr * gi + omi * gr) * ir2;
    if (CHG) {
      a[0] += st[0] * gr - st[1] * gi;
      a[1] += st[0] * gi + st[1] * gr;
      if (FLD) {
        const TN cr = st[0] * g_r - st[1] * g_i, ci = st[0] * g_i + st[1] * g_r;
        a[2] += cr * dx; a[3] += ci * dx;
        a[4] += cr * dy; a[5] += ci * dy;
        a[6] += cr * dz; a[7] += ci * dz;
      }
    }
    if (DIP) {
      const TN dotr = st[2] * dx + st[4] * dy + st[6] * dz;
      const TN doti = st[3] * dx + st[5] * dy + st[7] * dz;
      a[0] += dotr * g_r - doti * g_i;
      a[1] += dotr * g_i + doti * g_r;
      if (FLD) {
        // h = G (k^2 r^2 + 3 i k r - 3) / r^4
        const TN k2r = (kre * kre - kim * kim) * r2, k2i = TN(2) * kre * kim * r2;
        const TN wr = k2r - TN(3) * kim * r - TN(3), wi = k2i + TN(3) * kre * r;
        const TN ir4 = ir2 * ir2;
        const TN hr = (gr * wr - gi * wi) * ir4, hi = (gr * wi + gi * wr) * ir4;
        const TN ur = dotr * hr - doti * hi, ui = dotr * hi + doti * hr;
        a[2] -= (st[2] * g_r - st[3] * g_i) + ur * dx;
        a[3] -= (st[2] * g_i + st[3] * g_r) + ui * dx;
        a[4] -= (st[4] * g_r - st[5] * g_i) + ur * dy;
        a[5] -= (st[4] * g_i + st[5] * g_r) + ui * dy;
        a[6] -= (st[6] * g_r - st[7] * g_i) + ur * dz;
        a[7] -= (st[6] * g_i + st[7] * g_r) + ui * dz;
      }
    }
    return close && !mask;
  }
  __device__ __forceinline__ void fold_into(double* t) {
    for (int k = 0; k < 8; k++) {
      t[k] += (double)a[k];
      a[k] = TN(0);
    }
  }
};

// The close pairs (r^2 < rc2) of sources s0..s1-1 at (xt, yt, zt), except
// source `self`, in FP64 from the FP64 strengths, added to tot.
template <bool CHG, bool DIP, bool FLD>
__device__ __forceinline__ void close_hpairs(double kre, double kim, double xt, double yt, double zt,
                                             const double* __restrict__ spos, const double* __restrict__ str64,
                                             int s0, int s1, int self, double rc2, double* tot) {
  HPairAcc<double, CHG, DIP, FLD> c64;
  for (int s = s0; s < s1; s++) {
    const double* sp = spos + 4 * (size_t)s;
    const double dx = xt - sp[0], dy = yt - sp[1], dz = zt - sp[2];
    if (s != self && fma(dx, dx, fma(dy, dy, dz * dz)) < rc2)
      c64.add(kre, kim, xt, yt, zt, sp, str64 + 8 * (size_t)s, false);
  }
  c64.fold_into(tot);
}

template <typename TN, bool CHG, bool DIP, bool FLD>
__global__ __launch_bounds__(256) void k_hp2p(DevPlan P, double kre, double kim, const double* __restrict__ spos,
                                              const TN* __restrict__ str, const double* __restrict__ str64,
                                              const double* __restrict__ tpos, int kind, double* __restrict__ pot,
                                              double* __restrict__ fld) {
  // one work item (64 points) per wave
  const int w = __builtin_amdgcn_readfirstlane((int)(xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= P.n_p2p) return;
  const int b = P.p2p_items[3 * w], start = P.p2p_items[3 * w + 2];
  if (b < 0) return;   // padding item
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + (int)(threadIdx.x & 63);
  const bool active = i < nn;
  const int me = n0 + (active ? i : 0);
  const double* mp = (kind ? tpos : spos) + 4 * (size_t)me;
  const double x = mp[0], y = mp[1], z = mp[2];
  HPairAcc<TN, CHG, DIP, FLD> acc;
  double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  constexpr bool fold = sizeof(TN) < sizeof(double);
  // FP32 pairs: pairs closer than an eighth of this leaf's side run in FP64
  const double hside = 1.0 / P.inv_side[P.level[b]];
  const double rc2 = fold ? hside * hside * (1.0 / 64.0) : 0.0;
  {
    const int s0 = P.isrc[b], s1 = s0 + P.nsrc[b];
    const int self = kind ? -1 : me;
    for (int c = s0; c < s1; c += (fold ? kFold : (1 << 30))) {
      const int ce = fold ? min(c + kFold, s1) : s1;
      bool close = false;
      for (int s = c; s < ce; s++)
        close |= acc.add(kre, kim, x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, s == self, rc2);
      if (fold) {
        acc.fold_into(tot);
        if (close) close_hpairs<CHG, DIP, FLD>(kre, kim, x, y, z, spos, str64, c, ce, self, rc2, tot);
      }
    }
  }
  for (int jj = P.l1_off[b]; jj < P.l1_off[b + 1]; jj++) {
    const int j = P.l1_idx[jj];
    const int s0 = P.isrc[j], s1 = s0 + P.nsrc[j];
    for (int c = s0; c < s1; c += (fold ? kFold : (1 << 30))) {
      const int ce = fold ? min(c + kFold, s1) : s1;
      bool close = false;
      for (int s = c; s < ce; s++)
        close |= acc.add(kre, kim, x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, false, rc2);
      if (fold) {
        acc.fold_into(tot);
        if (close) close_hpairs<CHG, DIP, FLD>(kre, kim, x, y, z, spos, str64, c, ce, -1, rc2, tot);
      }
    }
  }
  if (!fold) acc.fold_into(tot);
  if (!active) return;
  pot[2 * (size_t)me] += tot[0];
  pot[2 * (size_t)me + 1] += tot[1];
  if (FLD)
    for (int k = 0; k < 6; k++) fld[6 * (size_t)me + k] += tot[2 + k];
}

// h3dpartdirect (src/hfmm3dpart.f:1268): every source on every point
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_hdirect(double kre, double kim, int ns, const double* __restrict__ spos,
                                                 const double* __restrict__ str, int npts,
                                                 const double* __restrict__ pts, int pts_are_sources,
                                                 double* __restrict__ pot, double* __restrict__ fld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < npts;
  const int me = active ? i : 0;
  const double* mp = (pts_are_sources ? spos : pts) + 4 * (size_t)me;
  const double x = mp[0], y = mp[1], z = mp[2];
  HPairAcc<double, CHG, DIP, true> acc;
  const int self = pts_are_sources ? me : -1;
  for (int s = 0; s < ns; s++) acc.add(kre, kim, x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, s == self);
  if (!active) return;
  pot[2 * (size_t)i] = acc.a[0];
  pot[2 * (size_t)i + 1] = acc.a[1];
  for (int k = 0; k < 6; k++) fld[6 * (size_t)i + k] = acc.a[2 + k];
}

}  // namespace fmm3d
