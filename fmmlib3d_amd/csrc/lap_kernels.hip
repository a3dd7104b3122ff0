// Laplace particle-FMM kernels for gfx950 (MI355X).
//
// Every kernel is templated on the arithmetic type T (double for the FP64
// path, float for the mixed path).  Expansions are packed complex arrays,
// coefficient (n,m) at eidx(n,m), and carry rscale = 1/side of their box, so
// translation tables are level independent.  Conventions match the
// reference's "_trunc" routines (src/laprouts3d.f:12-66 with the (-1)^m
// phase of ylgndrufw): R_n^m = r^n Yhat_n^m, S_n^m = Yhat_n^m / r^{n+1},
// multipole M_n^m = sum q conj(R_n^m(u)), local L_n^m = sum q conj(S_n^m(u))/h.
#include <hip/hip_runtime.h>

#include "kern_common.hpp"
#include "m2l_grp.hpp"

namespace fmm3d {

// ---------------------------------------------------------------------------
// P2M (src/laprouts3d.f:3210, :3835) and P2L (src/laprouts3d.f:3550, :4207)
// One workgroup per box.  Sources are processed in tiles of `ts`: phase A
// fills an LDS table of harmonics (one thread per (source, m) column),
// phase B contracts it against the strengths (one thread per coefficient).
// IRREG selects P2L (irregular harmonics at the target box centre, sources
// taken from the boxes of list 4) instead of P2M.
// ---------------------------------------------------------------------------
template <typename T, bool IRREG>
__global__ void k_form(DevPlan P, const double* __restrict__ fac, const double* __restrict__ spos,
                       const double* __restrict__ str, int ifcharge, int ifdipole, int ts,
                       T* __restrict__ expn) {
  extern __shared__ unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  const int p = P.p, K = P.K;
  const int pt = IRREG ? p + 1 : p;  // table degree (dipoles need one more for S)
  const int H = (pt + 1) * (pt + 2) / 2;
  FacLayout L(p);
  T* acc = smem;                    // 2*K
  T* tab = acc + 2 * K;             // ts * H complex
  T* sst = tab + 2 * (size_t)ts * H; // ts * 8 : q, dz, du, dd (complex)
  T* pos = sst + 8 * ts;            // ts * 3 scaled offsets

  const int b = IRREG ? P.p2l_boxes[blockIdx.x] : P.p2m_boxes[blockIdx.x];
  const double ihd = P.inv_side[P.level[b]];
  const T ih = (T)ihd;
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) acc[k] = T(0);

  int jbeg, jend;
  if (IRREG) { jbeg = P.l4_off[b]; jend = P.l4_off[b + 1]; }
  else { jbeg = 0; jend = 1; }
  for (int jj = jbeg; jj < jend; jj++) {
    const int j = IRREG ? P.l4_idx[jj] : b;
    const int s0 = P.isrc[j], sn = P.nsrc[j];
    for (int t0 = 0; t0 < sn; t0 += ts) {
      const int tn = min(ts, sn - t0);
      __syncthreads();
      for (int s = threadIdx.x; s < tn; s += blockDim.x) {
        const double* x = spos + 4 * (size_t)(s0 + t0 + s);
        const double* r = str + 8 * (size_t)(s0 + t0 + s);
        pos[3 * s] = (T)((x[0] - cx) * ihd);
        pos[3 * s + 1] = (T)((x[1] - cy) * ihd);
        pos[3 * s + 2] = (T)((x[2] - cz) * ihd);
        // charge, and dipole components dz = v_z, du = (v_x - i v_y)/2, dd = (v_x + i v_y)/2
        sst[8 * s] = (T)r[0]; sst[8 * s + 1] = (T)r[1];
        sst[8 * s + 2] = (T)r[6]; sst[8 * s + 3] = (T)r[7];
        sst[8 * s + 4] = (T)(0.5 * (r[2] + r[5])); sst[8 * s + 5] = (T)(0.5 * (r[3] - r[4]));
        sst[8 * s + 6] = (T)(0.5 * (r[2] - r[5])); sst[8 * s + 7] = (T)(0.5 * (r[3] + r[4]));
      }
      __syncthreads();
      // phase A: harmonic columns
      for (int task = threadIdx.x; task < tn * (pt + 1); task += blockDim.x) {
        int s = task / (pt + 1), m = task - s * (pt + 1);
        T ux = pos[3 * s], uy = pos[3 * s + 1], uz = pos[3 * s + 2];
        T r2 = ux * ux + uy * uy + uz * uz;
        if (IRREG) {
          T ir2 = T(1) / r2, ir = sqrt(ir2);
          irregular_column<T>(fac, L, pt, m, ux, uy, uz, ir2, ir, tab + 2 * (size_t)s * H, 1);
        } else {
          regular_column<T>(fac, L, pt, m, ux, uy, uz, r2, tab + 2 * (size_t)s * H, 1);
        }
      }
      __syncthreads();
      // phase B: contraction, one thread per coefficient (n,m)
      for (int k = threadIdx.x; k < K; k += blockDim.x) {
        int n = (int)sqrtf((float)k);
        while (n * n > k) n--;
        while ((n + 1) * (n + 1) <= k) n++;
        int m = k - n * n - n;
        C2<T> a = {T(0), T(0)};
        // derivative factors of the harmonic with order -m (grad of conj(H_n^m) = grad H_n^{-m})
        int km = -m;
        T fa = 0, fb = 0, fcf = 0;
        if (IRREG) {
          fa = (T)fac[L.AS + eidx(n, km)]; fb = (T)fac[L.BS + eidx(n, km)]; fcf = (T)fac[L.CS + eidx(n, km)];
        } else if (n >= 1) {
          fa = (T)fac[L.AR + eidx(n, km)]; fb = (T)fac[L.BR + eidx(n, km)]; fcf = (T)fac[L.CR + eidx(n, km)];
        }
        for (int s = 0; s < tn; s++) {
          const T* tb = tab + 2 * (size_t)s * H;
          if (ifcharge) {
            C2<T> hv = half_get(tb, n, km, 1);  // conj(H_n^m) = H_n^{-m}
            cfma(a, C2<T>{sst[8 * s], sst[8 * s + 1]}, hv);
          }
          if (ifdipole) {
            C2<T> g = {T(0), T(0)};
            int nn = IRREG ? n + 1 : n - 1;
            if (IRREG || n >= 1) {
              if (abs(km) <= nn) { C2<T> h = half_get(tb, nn, km, 1); cfma(g, C2<T>{fa * sst[8 * s + 2], fa * sst[8 * s + 3]}, h); }
              if (abs(km + 1) <= nn) { C2<T> h = half_get(tb, nn, km + 1, 1); cfma(g, C2<T>{fb * sst[8 * s + 4], fb * sst[8 * s + 5]}, h); }
              if (abs(km - 1) <= nn) { C2<T> h = half_get(tb, nn, km - 1, 1); cfma(g, C2<T>{fcf * sst[8 * s + 6], fcf * sst[8 * s + 7]}, h); }
            }
            // d/dy of H(u) with u = (y-c)/h gives a factor 1/h
            a.re += g.re * ih;
            a.im += g.im * ih;
          }
        }
        acc[2 * k] += a.re;
        acc[2 * k + 1] += a.im;
      }
    }
  }
  __syncthreads();
  // P2L: L = sum q conj(S(u)) / h ; P2M: M = sum q conj(R(u))
  const T sc = IRREG ? ih : T(1);
  T* out = expn + 2 * (size_t)b * K;
  for (int k = threadIdx.x; k < 2 * K; k += blockDim.x) out[k] += acc[k] * sc;
}

// ---------------------------------------------------------------------------
// P2M / P2L with per-thread sums.  Each thread owns up to two half-table
// entries (n', a >= 0) and accumulates, in registers over every source of the
// box, S^w(n', +-a) = sum_s w_s H_{n'}^{+-a}(u_s) for the weights
// w = q, v_x, v_y, v_z (H^{-a} = conj(H^a)); the weights are wave-uniform
// loads.  The harmonic table of a tile of `ts` sources is built in LDS, one
// thread per (column m, source), m-major so that a wave holds few orders and
// its loop lengths are close; each source's scalars (form_scalars) are
// computed once, during the previous tile's sums, from positions read while
// the previous tile's table was built.  At the end the sums go to LDS
// and one thread per coefficient combines them: charge S^q(n, -m), dipole
// through the derivative factors and degree n-1 (P2M) or n+1 (P2L), with
// (v_x -+ i v_y)/2 formed from S^{v_x}, S^{v_y} by linearity.
// ---------------------------------------------------------------------------
// Threads of k_form2.  P2L: 512, since the tile's LDS admits two workgroups per
// CU and twice the waves of 256-thread groups hide the column recurrence.
// P2M: 256, since it runs beside the near field, where a 512-thread workgroup
// waits for two near-field workgroups to leave a CU.  Each thread owns
// 512 / threads half-table entries, twice that for larger orders.
template <bool IRREG>
__host__ __device__ constexpr int form2_threads() {
  return IRREG ? 512 : 256;
}
// LDS bytes k_form2 keeps ahead of the harmonic table, 16-aligned: the
// recurrence factors (FacLayout MM..RB) and cm[m] = prod_{k<=m} MM[k],
// m = 0..p+2
// Tile row stride (entries) of k_form2: the half table of one source, padded
// so that consecutive sources start one entry apart modulo the 32 LDS banks
// (the harmonic columns of neighbouring sources are written by neighbouring
// lanes)
template <typename T>
__host__ __device__ inline int form2_row_stride(int H) {
  const int words = 2 * (int)sizeof(T) / 4;  // per complex entry
  int hs = H;
  while ((hs * words) % 32 != words) hs++;
  return hs;
}

__host__ __device__ inline size_t form2_fac_bytes(int p) {
  return ((size_t)(FacLayout(p).AR + p + 3) * sizeof(double) + 15) / 16 * 16;
}

// LDS stride (T) of the per-source scalars: 8 used, padded so that the rows
// of neighbouring sources start in different banks
constexpr int kFormScStride = 10;

// Per-source scalars of the column tasks, 8 T per source: the power base
// w = -(ux + i uy) (times 1/|u|^2 for S), uz, the scale of the recurrence's
// second term (|u|^2 for R, 1 for S), the scale of each new entry (1 for R,
// 1/|u|^2 for S) and of the diagonal (1 for R, 1/|u| for S).
template <typename T, bool IRREG>
__device__ __forceinline__ void form_scalars(double x, double y, double z, double cx, double cy, double cz,
                                             double ihd, T* sc) {
  const T ux = (T)((x - cx) * ihd), uy = (T)((y - cy) * ihd), uz = (T)((z - cz) * ihd);
  const T r2 = ux * ux + uy * uy + uz * uz;
  T ir2 = T(1), ir = T(1);
  if (IRREG) {
    ir2 = T(1) / r2;
    ir = sqrt(ir2);
  }
  sc[0] = IRREG ? -ux * ir2 : -ux;
  sc[1] = IRREG ? -uy * ir2 : -uy;
  sc[2] = uz;
  sc[3] = IRREG ? T(1) : r2;
  sc[4] = ir2;
  sc[5] = ir;
  sc[6] = T(0);
  sc[7] = T(0);
}

// One complex table entry as a single LDS store (16 bytes for double, 8 for
// float; the k_form2 tile rows keep entries at that alignment)
template <typename T>
__device__ __forceinline__ void cst_vec(T* p, int k, C2<T> v) {
  using V2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  V2 t;
  t.x = v.re;
  t.y = v.im;
  *reinterpret_cast<V2*>(p + 2 * k) = t;
}

// Column m of the regular (R) or irregular (S) harmonics of one source into
// a half table, n = m..pt, from that source's scalars `sc` (form_scalars).
// The diagonal entry R_m^m = cm w^m, S_m^m = ir cm (w ir2)^m is taken by
// binary powers; then the three-term recurrences of regular_column and
// irregular_column.  The entry index advances by n + 1 per step, and the
// next step's two recurrence factors are read from LDS before this step's
// arithmetic, so the LDS latency overlaps the dependent chain.
template <typename T, bool IRREG>
__device__ __forceinline__ void form_column(const double* facl, const FacLayout& L, int pt, int m, T cm,
                                            const T* sc, T* out) {
  C2<T> base = {sc[0], sc[1]};
  const T uz = sc[2], bsc = sc[3], csc = sc[4];
  C2<T> r = {IRREG ? cm * sc[5] : cm, T(0)};
  for (int e = m; e; e >>= 1) {
    if (e & 1) r = cmul(r, base);
    base = cmul(base, base);
  }
  int k = hidx(m, m);
  cst_vec(out, k, r);
  if (m + 1 > pt) return;
  const T g = (T)facl[L.M1 + m] * uz * csc;
  C2<T> r1 = {g * r.re, g * r.im};
  k += m + 1;  // hidx(m + 1, m)
  cst_vec(out, k, r1);
  if (m + 2 > pt) return;
  C2<T> r0 = r;
  k += m + 2;  // hidx(m + 2, m)
  double fa = facl[L.RA + k], fb = facl[L.RB + k];
  T cn = (T)(2 * m + 3);  // 2n - 1, exact
#pragma unroll 2
  for (int n = m + 2; n <= pt; n++) {
    const T b = IRREG ? (T)fa : (T)fa * bsc;
    const T c = IRREG ? (T)fb * csc : (T)fb;
    const int kn = k + n + 1;  // hidx(n + 1, m)
    // at n = pt this reads past the column (still inside the factor and cm
    // tables in LDS); the values go unused
    fa = facl[L.RA + kn];
    fb = facl[L.RB + kn];
    const T a = cn * uz;
    cn += T(2);
    const C2<T> rn = {(a * r1.re - b * r0.re) * c, (a * r1.im - b * r0.im) * c};
    cst_vec(out, k, rn);
    r0 = r1;
    r1 = rn;
    k = kn;
  }
}

template <typename T, bool IRREG, bool CHG, bool DIP, int ITEMS>
__global__ __launch_bounds__(form2_threads<IRREG>()) void k_form2(DevPlan P, const double* __restrict__ fac,
                                               const double* __restrict__ spos, const double* __restrict__ str,
                                               int ts, T* __restrict__ expn) {
  extern __shared__ unsigned char smem_raw[];
  const int p = P.p, K = P.K;
  const int pt = IRREG ? p + 1 : p;
  const int H = (pt + 1) * (pt + 2) / 2;
  const int HS = form2_row_stride<T>(H);
  const int P2 = (pt + 1) * (pt + 1);
  FacLayout L(p);
  const int tid = threadIdx.x;
  // the recurrence factors (MM, M1, RA, RB: the head of the factor table) in
  // LDS, ahead of the harmonic table: each column task reads a different
  // entry; then the diagonal constants cm
  double* facl = reinterpret_cast<double*>(smem_raw);
  double* cmt = facl + L.AR;
  for (int k = tid; k < L.AR; k += blockDim.x) facl[k] = fac[k];
  if (tid == 0) {
    double c = 1.0;
    cmt[0] = 1.0;
    for (int k = 1; k <= pt; k++) { c *= fac[L.MM + k]; cmt[k] = c; }
  }
  T* tab = reinterpret_cast<T*>(smem_raw + form2_fac_bytes(p));
  const int b = IRREG ? P.p2l_boxes[blockIdx.x] : P.p2m_boxes[blockIdx.x];
  const double ihd = P.inv_side[P.level[b]];
  const T ih = (T)ihd;
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  T acc[ITEMS][16];
#pragma unroll
  for (int q = 0; q < ITEMS; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = T(0);
  // per-source scalars of the current tile, after the tile region
  T* scs = tab + 2 * (size_t)max(ts * HS, 4 * P2);
  T* wsh = scs + kFormScStride * ts;  // the tile's source weights
  int rot = blockIdx.x;
  using V2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  // tile cursor over the sources of the list-4 boxes (P2L) or of box b (P2M):
  // (jj, t0) moves to the first tile holding sources, s0/sn of its box
  const int jend = IRREG ? P.l4_off[b + 1] : 1;
  auto settle = [&](int& jj, int& t0, int& s0, int& sn) {
    for (; jj < jend; jj++, t0 = 0) {
      const int j = IRREG ? P.l4_idx[jj] : b;
      s0 = P.isrc[j];
      sn = P.nsrc[j];
      if (t0 < sn) return;
    }
  };
  int jj = IRREG ? P.l4_off[b] : 0, t0 = 0, s0 = 0, sn = 0;
  settle(jj, t0, s0, sn);
  if (jj < jend && tid < min(ts, sn - t0)) {
    const double* x = spos + 4 * (size_t)(s0 + t0 + tid);
    form_scalars<T, IRREG>(x[0], x[1], x[2], cx, cy, cz, ihd, scs + kFormScStride * tid);
  }
  while (jj < jend) {
    const int tn = min(ts, sn - t0), cs0 = s0 + t0;
    int njj = jj, nt0 = t0 + tn, ns0 = s0, nsn = sn;
    settle(njj, nt0, ns0, nsn);
    const int ntn = njj < jend ? min(ts, nsn - nt0) : 0;
    __syncthreads();  // this tile's scalars written; the last tile's sums done with the table
    // the next tile's positions, read while this tile's table is built
    double nx = 0, ny = 0, nz = 0;
    if (tid < ntn) {
      const double* x = spos + 4 * (size_t)(ns0 + nt0 + tid);
      nx = x[0];
      ny = x[1];
      nz = x[2];
    }
    // this tile's weights, one value per thread (8 ts <= 256 <= blockDim.x),
    // staged in LDS
    // for the sums
    T wreg = T(0);
    if (tid < 8 * tn) wreg = (T)str[8 * (size_t)cs0 + tid];
    // column tasks, m-major, in blocks of 64; block k of round r goes to wave
    // (k - r - rot) mod nw: the wave with the shortest columns of one round
    // takes the longest of the next, and the long low-m columns visit every
    // wave (and SIMD) in turn from tile to tile
    {
      const int ntask = tn * (pt + 1), nw = (int)blockDim.x >> 6;
      int vw = ((tid >> 6) + rot) % nw;
      for (int r0 = 0; r0 < ntask; r0 += blockDim.x) {
        const int task = r0 + 64 * vw + (tid & 63);
        if (task < ntask) {
          const int m = task / tn, s = task - m * tn;
          form_column<T, IRREG>(facl, L, pt, m, (T)cmt[m], scs + kFormScStride * s, tab + 2 * (size_t)s * HS);
        }
        vw = vw + 1 == nw ? 0 : vw + 1;
      }
      rot++;
    }
    if (tid < 8 * tn) wsh[tid] = wreg;
    __syncthreads();
    if (tid < ntn) form_scalars<T, IRREG>(nx, ny, nz, cx, cy, cz, ihd, scs + kFormScStride * tid);
#pragma unroll 2
    for (int s = 0; s < tn; s++) {
      const T* wv = wsh + 8 * s;
#pragma unroll
      for (int q = 0; q < ITEMS; q++) {
        const int h = tid + q * (int)blockDim.x;
        if (h < H) {
          const V2 hv = *reinterpret_cast<const V2*>(tab + 2 * ((size_t)s * HS + h));
          const T hr = hv.x, hi = hv.y;
#pragma unroll
          for (int c = 0; c < 4; c++) {
            if ((c == 0 && !CHG) || (c > 0 && !DIP)) continue;
            const T wr = wv[2 * c], wi = wv[2 * c + 1];
            T* A = &acc[q][4 * c];
            A[0] = fma(wr, hr, fma(-wi, hi, A[0]));  // w H^{+a}
            A[1] = fma(wr, hi, fma(wi, hr, A[1]));
            A[2] = fma(wr, hr, fma(wi, hi, A[2]));   // w H^{-a} = w conj(H^{a})
            A[3] = fma(wi, hr, fma(-wr, hi, A[3]));
          }
        }
      }
    }
    jj = njj;
    t0 = nt0;
    s0 = ns0;
    sn = nsn;
  }
  __syncthreads();
  // sums to LDS: S[c][eidx(n', m')] complex, c = q, v_x, v_y, v_z
#pragma unroll
  for (int q = 0; q < ITEMS; q++) {
    const int h = tid + q * (int)blockDim.x;
    if (h < H) {
      int n = (int)((sqrtf(8.0f * h + 1.0f) - 1.0f) * 0.5f);
      while (n * (n + 1) / 2 > h) n--;
      while ((n + 1) * (n + 2) / 2 <= h) n++;
      const int a = h - n * (n + 1) / 2;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        T* S = tab + 2 * (size_t)c * P2;
        S[2 * eidx(n, a)] = acc[q][4 * c];
        S[2 * eidx(n, a) + 1] = acc[q][4 * c + 1];
        S[2 * eidx(n, -a)] = acc[q][4 * c + 2];
        S[2 * eidx(n, -a) + 1] = acc[q][4 * c + 3];
      }
    }
  }
  __syncthreads();
  const T sc = IRREG ? ih : T(1);
  T* out = expn + 2 * (size_t)b * K;
  for (int k = tid; k < K; k += blockDim.x) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    const int m = k - n * n - n, km = -m;
    T ar = 0, ai = 0;
    if (CHG) {
      ar = tab[2 * eidx(n, km)];
      ai = tab[2 * eidx(n, km) + 1];
    }
    if (DIP && (IRREG || n >= 1)) {
      const int nn = IRREG ? n + 1 : n - 1;
      const T fa = (T)fac[(IRREG ? L.AS : L.AR) + eidx(n, km)];
      const T fb = (T)fac[(IRREG ? L.BS : L.BR) + eidx(n, km)] * T(0.5);
      const T fc = (T)fac[(IRREG ? L.CS : L.CR) + eidx(n, km)] * T(0.5);
      const T* SX = tab + 2 * (size_t)P2;
      const T* SY = SX + 2 * (size_t)P2;
      const T* SZ = SY + 2 * (size_t)P2;
      T gr = 0, gi = 0;
      if (km <= nn && km >= -nn) {
        gr += fa * SZ[2 * eidx(nn, km)];
        gi += fa * SZ[2 * eidx(nn, km) + 1];
      }
      if (km + 1 <= nn && km + 1 >= -nn) {  // (v_x - i v_y)/2
        const int e = eidx(nn, km + 1);
        gr += fb * (SX[2 * e] + SY[2 * e + 1]);
        gi += fb * (SX[2 * e + 1] - SY[2 * e]);
      }
      if (km - 1 <= nn && km - 1 >= -nn) {  // (v_x + i v_y)/2
        const int e = eidx(nn, km - 1);
        gr += fc * (SX[2 * e] - SY[2 * e + 1]);
        gi += fc * (SX[2 * e + 1] + SY[2 * e]);
      }
      ar += gr * ih;
      ai += gi * ih;
    }
    out[2 * k] += ar * sc;
    out[2 * k + 1] += ai * sc;
  }
}

// ---------------------------------------------------------------------------
// L2P (src/laprouts3d.f:2772 l3dtaevalall_trunc) from a packed expansion.
// For every box of the L2P list the local
// expansion L and the three derived expansions of the gradient,
//   Gz(n,m) = AR(n+1,m)   L(n+1,m),
//   Gu(n,m) = BR(n+1,m-1) L(n+1,m-1),
//   Gd(n,m) = CR(n+1,m+1) L(n+1,m+1)       (n <= p-1; zero at n = p),
// are stored per half-table entry (n, a >= 0) in the order the evaluation
// walks them (column a, then n = a..p): 8 complex values
//   L(n,a) L(n,-a) Gz(n,a) Gz(n,-a) Gu(n,a) Gu(n,-a) Gd(n,a) Gd(n,-a)
// (the -a slots are zero for a = 0).  The evaluation then reads them with
// wave-uniform loads and pays four complex products per harmonic:
//   pot = sum L R, d/dz = sum Gz R, (d/dx + i d/dy) = sum Gu R, (d/dx - i d/dy) = sum Gd R.
// ---------------------------------------------------------------------------
__host__ __device__ inline int walk_index(int p, int a, int n) { return a * (p + 1) - a * (a - 1) / 2 + (n - a); }

template <typename T>
__global__ void k_l2p_pack(DevPlan P, const double* __restrict__ fac, const T* __restrict__ lo,
                           T* __restrict__ pack) {
  const int p = P.p, K = P.K;
  const int H = (p + 1) * (p + 2) / 2;
  FacLayout L(p);
  const size_t total = (size_t)P.n_l2p * H;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / H), w = (int)(t - (size_t)i * H);
    int a = 0;
    while (a < p && walk_index(p, a + 1, a + 1) <= w) a++;
    const int n = a + (w - walk_index(p, a, a));
    const T* E = lo + 2 * (size_t)P.l2p_boxes[i] * K;
    T v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = T(0);
    v[0] = E[2 * eidx(n, a)];
    v[1] = E[2 * eidx(n, a) + 1];
    if (a > 0) { v[2] = E[2 * eidx(n, -a)]; v[3] = E[2 * eidx(n, -a) + 1]; }
    if (n + 1 <= p) {
      const int n1 = n + 1;
      for (int sgn = 0; sgn < (a > 0 ? 2 : 1); sgn++) {
        const int m = sgn ? -a : a;
        const T fz = (T)fac[L.AR + eidx(n1, m)], fu = (T)fac[L.BR + eidx(n1, m - 1)], fd = (T)fac[L.CR + eidx(n1, m + 1)];
        v[4 + 2 * sgn] = fz * E[2 * eidx(n1, m)];
        v[5 + 2 * sgn] = fz * E[2 * eidx(n1, m) + 1];
        v[8 + 2 * sgn] = fu * E[2 * eidx(n1, m - 1)];
        v[9 + 2 * sgn] = fu * E[2 * eidx(n1, m - 1) + 1];
        v[12 + 2 * sgn] = fd * E[2 * eidx(n1, m + 1)];
        v[13 + 2 * sgn] = fd * E[2 * eidx(n1, m + 1) + 1];
      }
    }
    T* o = pack + 16 * t;
#pragma unroll
    for (int e = 0; e < 16; e++) o[e] = v[e];
  }
}

// out += c+ h + c- conj(h)
template <typename T>
__device__ __forceinline__ void pack_fma(C2<T>& out, const T* c, const C2<T>& h) {
  out.re = fma(c[0], h.re, fma(-c[1], h.im, fma(c[2], h.re, fma(c[3], h.im, out.re))));
  out.im = fma(c[0], h.im, fma(c[1], h.re, fma(-c[2], h.im, fma(c[3], h.re, out.im))));
}

// One workgroup per L2P work item (box, kind, first point); the loops over
// the packed entries are wave uniform, so the entries are scalar loads.
template <typename T>
__global__ __launch_bounds__(256) void k_l2p2(DevPlan P, const double* __restrict__ fac, const T* __restrict__ pack,
                                               const double* __restrict__ spos, const double* __restrict__ tpos,
                                               int wpot0, int wfld0, int wpot1, int wfld1,
                                               double* __restrict__ pot0, double* __restrict__ fld0,
                                               double* __restrict__ pot1, double* __restrict__ fld1) {
  const int p = P.p;
  const int H = (p + 1) * (p + 2) / 2;
  FacLayout L(p);
  // one work item (64 points) per wave
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= P.n_l2pi) return;
  const int ib = P.l2p_items[3 * w], kind = P.l2p_items[3 * w + 1], start = P.l2p_items[3 * w + 2];
  if (ib < 0) return;   // padding item
  const int b = P.l2p_boxes[ib];
  const T* pk = pack + (size_t)ib * H * 16;
  const T ih = (T)P.inv_side[P.level[b]];
  const double cx = P.center[3 * b], cy = P.center[3 * b + 1], cz = P.center[3 * b + 2];
  const int wp = kind ? wpot1 : wpot0, wf = kind ? wfld1 : wfld0;
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + (int)(threadIdx.x & 63);
  const bool active = i < nn;
  double x = cx, y = cy, z = cz;
  if (active) {
    const double* q = (kind ? tpos : spos) + 4 * (size_t)(n0 + i);
    x = q[0]; y = q[1]; z = q[2];
  }
  const T ux = (T)((x - cx) * (double)ih), uy = (T)((y - cy) * (double)ih), uz = (T)((z - cz) * (double)ih);
  const T r2 = ux * ux + uy * uy + uz * uz;
  C2<T> pv = {0, 0}, gz = {0, 0}, gu = {0, 0}, gd = {0, 0};
  C2<T> hmm = {T(1), T(0)};
  const T* c = pk;
  for (int a = 0; a <= p; a++) {
    if (a > 0) {
      const T f = (T)fac[L.MM + a];
      const C2<T> v = {-ux * f, -uy * f};
      hmm = cmul(v, hmm);
    }
    C2<T> h0 = hmm, h1 = hmm;
    for (int n = a; n <= p; n++, c += 16) {
      C2<T> h;
      if (n == a) {
        h = hmm;
      } else if (n == a + 1) {
        const T g = (T)fac[L.M1 + a] * uz;
        h = {g * hmm.re, g * hmm.im};
        h1 = h;
      } else {
        const T aa = (T)(2 * n - 1) * uz, bq = (T)fac[L.RA + hidx(n, a)] * r2, cc = (T)fac[L.RB + hidx(n, a)];
        h = {(aa * h1.re - bq * h0.re) * cc, (aa * h1.im - bq * h0.im) * cc};
        h0 = h1;
        h1 = h;
      }
      pack_fma(pv, c, h);
      if (wf) {
        pack_fma(gz, c + 4, h);
        pack_fma(gu, c + 8, h);
        pack_fma(gd, c + 12, h);
      }
    }
  }
  if (!active) return;
  const size_t id = (size_t)(n0 + i);
  double* po = kind ? pot1 : pot0;
  double* fo = kind ? fld1 : fld0;
  if (wp) { po[2 * id] += pv.re; po[2 * id + 1] += pv.im; }
  if (wf) {
    // fld = -grad: d/dx = (U+D)/2, d/dy = (U-D)/(2i), d/dz; the gradient in x carries 1/h
    const C2<T> dx = {T(0.5) * (gu.re + gd.re), T(0.5) * (gu.im + gd.im)};
    const C2<T> dy = {T(0.5) * (gu.im - gd.im), -T(0.5) * (gu.re - gd.re)};
    fo[6 * id + 0] -= dx.re * ih; fo[6 * id + 1] -= dx.im * ih;
    fo[6 * id + 2] -= dy.re * ih; fo[6 * id + 3] -= dy.im * ih;
    fo[6 * id + 4] -= gz.re * ih; fo[6 * id + 5] -= gz.im * ih;
  }
}

// ---------------------------------------------------------------------------
// M2P (list 3) from packed multipoles.  For every box whose multipole is
// evaluated, the multipole M and the three derived expansions of the
// gradient over the irregular harmonics of degree n <= p+1,
//   Gz(n,m) = AS(n-1,m)   M(n-1,m),
//   Gu(n,m) = BS(n-1,m-1) M(n-1,m-1),
//   Gd(n,m) = CS(n-1,m+1) M(n-1,m+1)      (zero where (n-1, m') is outside),
// are stored per half-table entry (n, a >= 0) in the order of the walk
// (column a, then n = a..pd, pd = p+1 with the gradient, else p), 8 complex
// values as for L2P.  One wave evaluates 64 points of a box against a slice
// of its list 3 (src/laprouts3d.f:2432 l3dmpevalall_trunc): per source box
// the irregular harmonics S_n^a(u) by the column recurrence of
// irregular_column, four complex products per harmonic, scaled
// by 1/h (potential) and 1/h^2 (gradient) of the source box; the slice's sums
// go to `partial` and k_slice_reduce adds the slices in order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_m2p_pack(DevPlan P, const double* __restrict__ fac, const T* __restrict__ mp,
                           const int* __restrict__ boxes, const int* __restrict__ offs, int nbox, int grad,
                           T* __restrict__ pack) {
  const int p = P.p, K = P.K;
  const int pd = grad ? p + 1 : p;
  const int Hd = (pd + 1) * (pd + 2) / 2;
  FacLayout L(p);
  const size_t total = (size_t)nbox * Hd;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / Hd), w = (int)(t - (size_t)i * Hd);
    int a = 0;
    while (a < pd && walk_index(pd, a + 1, a + 1) <= w) a++;
    const int n = a + (w - walk_index(pd, a, a));
    const T* E = mp + 2 * (size_t)boxes[i] * K;
    auto ef = [&](int off, int n1, int m1, int e) -> T {  // factor * E(n1, m1), zero outside
      if (n1 < 0 || m1 > n1 || m1 < -n1) return T(0);
      return (T)fac[off + eidx(n1, m1)] * E[2 * eidx(n1, m1) + e];
    };
    T v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = T(0);
    if (n <= p) {
      v[0] = E[2 * eidx(n, a)];
      v[1] = E[2 * eidx(n, a) + 1];
      if (a > 0) { v[2] = E[2 * eidx(n, -a)]; v[3] = E[2 * eidx(n, -a) + 1]; }
    }
    if (grad && n >= 1) {
      const int n1 = n - 1;
      for (int sgn = 0; sgn < (a > 0 ? 2 : 1); sgn++) {
        const int m = sgn ? -a : a;
        for (int e = 0; e < 2; e++) {
          v[4 + 2 * sgn + e] = ef(L.AS, n1, m, e);
          v[8 + 2 * sgn + e] = ef(L.BS, n1, m - 1, e);
          v[12 + 2 * sgn + e] = ef(L.CS, n1, m + 1, e);
        }
      }
    }
    T* o = pack + 16 * ((size_t)offs[i] + w);
#pragma unroll
    for (int e = 0; e < 16; e++) o[e] = v[e];
  }
}

// items: 5 ints per wave (box, kind, first point, first and end position in
// list 3); slot[j]: the pack offset (in entries) of source box j
template <typename T>
__global__ __launch_bounds__(256) void k_m2p_eval(DevPlan P, const double* __restrict__ fac, const T* __restrict__ pack,
                                                  const int* __restrict__ slot, const int* __restrict__ items, int nitems,
                                                  int grad, const double* __restrict__ spos,
                                                  const double* __restrict__ tpos, double* __restrict__ partial) {
  const int p = P.p;
  const int pd = grad ? p + 1 : p;
  FacLayout L(p);
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= nitems) return;
  const int lane = (int)(threadIdx.x & 63);
  const int* it = items + 5 * w;
  const int b = it[0], kind = it[1], start = it[2], jbeg = it[3], jend = it[4];
  const int n0 = kind ? P.itrg[b] : P.isrc[b];
  const int nn = kind ? P.ntrg[b] : P.nsrc[b];
  const int i = start + lane;
  const bool active = i < nn;
  double x = 0.0, y = 0.0, z = 0.0;
  if (active) {
    const double* q = (kind ? tpos : spos) + 4 * (size_t)(n0 + i);
    x = q[0]; y = q[1]; z = q[2];
  }
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int jj = jbeg; jj < jend; jj++) {
    const int j = P.l3_idx[jj];
    if (P.nsrc[j] == 0) continue;
    const double ihd = P.inv_side[P.level[j]];
    const T ih = (T)ihd;
    // inactive lanes evaluate at a harmless point and are not written
    const T ux = active ? (T)((x - P.center[3 * j]) * ihd) : T(1);
    const T uy = active ? (T)((y - P.center[3 * j + 1]) * ihd) : T(1);
    const T uz = active ? (T)((z - P.center[3 * j + 2]) * ihd) : T(1);
    const T r2 = ux * ux + uy * uy + uz * uz;
    const T ir2 = T(1) / r2, ir = sqrt(ir2);
    C2<T> pv = {0, 0}, gz = {0, 0}, gu = {0, 0}, gd = {0, 0};
    C2<T> hmm = {ir, T(0)};
    const T* c = pack + 16 * (size_t)slot[j];
    for (int a = 0; a <= pd; a++) {
      if (a > 0) {
        const T f = (T)fac[L.MM + a] * ir2;
        const C2<T> v = {-ux * f, -uy * f};
        hmm = cmul(v, hmm);
      }
      C2<T> h0 = hmm, h1 = hmm;
      for (int n = a; n <= pd; n++, c += 16) {
        C2<T> h;
        if (n == a) {
          h = hmm;
        } else if (n == a + 1) {
          const T g = (T)fac[L.M1 + a] * uz * ir2;
          h = {g * hmm.re, g * hmm.im};
          h1 = h;
        } else {
          const T aa = (T)(2 * n - 1) * uz, bq = (T)fac[L.RA + hidx(n, a)], cc = (T)fac[L.RB + hidx(n, a)] * ir2;
          h = {(aa * h1.re - bq * h0.re) * cc, (aa * h1.im - bq * h0.im) * cc};
          h0 = h1;
          h1 = h;
        }
        pack_fma(pv, c, h);
        if (grad) {
          pack_fma(gz, c + 4, h);
          pack_fma(gu, c + 8, h);
          pack_fma(gd, c + 12, h);
        }
      }
    }
    // pot = sum M S(u)/h ; the gradient w.r.t. x adds another 1/h
    const T sg = ih * ih;
    acc[0] += (double)(pv.re * ih); acc[1] += (double)(pv.im * ih);
    acc[2] += (double)(gz.re * sg); acc[3] += (double)(gz.im * sg);
    acc[4] += (double)(gu.re * sg); acc[5] += (double)(gu.im * sg);
    acc[6] += (double)(gd.re * sg); acc[7] += (double)(gd.im * sg);
  }
  double* o = partial + 8 * ((size_t)w * 64 + lane);
#pragma unroll
  for (int e = 0; e < 8; e++) o[e] = acc[e];
}

// ---------------------------------------------------------------------------
// Point-and-shoot translation of one expansion inside a workgroup:
//   A = in * e^{i m phi}           (n <= pin)
//   B = rotate(A)                   (rotation class `rc`)
//   A = zshift(B)  with table Z    (n <= pin -> l <= pout)
//   B = rotate back(A)
//   out += B * e^{-i m phi} * scale
// src/l3dtrans.f:223 (M2M), :868 (M2L), :1676 (L2L) do the same sequence.
// ZMODE 0: Z[m][out][in] ; the same dense (p+1)^3 layout for every shift.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void zshift_block(const double* Z, int p, int pin, int pout, const T* in, T* out) {
  // orders m and -m share the row Z[|m|][l][.]: one thread produces both
  const int Kh = (pout + 1) * (pout + 2) / 2;
  const int P1 = p + 1;
  for (int k = threadIdx.x; k < Kh; k += blockDim.x) {
    int l = (int)((sqrtf(8.0f * k + 1.0f) - 1.0f) * 0.5f);
    while (l * (l + 1) / 2 > k) l--;
    while ((l + 1) * (l + 2) / 2 <= k) l++;
    const int am = k - l * (l + 1) / 2;
    const double* z = Z + ((size_t)am * P1 + l) * P1;
    T pr = 0, pi = 0, qr = 0, qi = 0;
    for (int n = am; n <= pin; n++) {
      const T c = (T)z[n];
      pr += c * in[2 * eidx(n, am)];
      pi += c * in[2 * eidx(n, am) + 1];
      qr += c * in[2 * eidx(n, -am)];
      qi += c * in[2 * eidx(n, -am) + 1];
    }
    out[2 * eidx(l, am)] = pr;
    out[2 * eidx(l, am) + 1] = pi;
    if (am > 0) {
      out[2 * eidx(l, -am)] = qr;
      out[2 * eidx(l, -am) + 1] = qi;
    }
  }
}

template <typename T>
__device__ void point_and_shoot(const DevTables& Tb, int rc, const double* phase, const double* Z,
                                int pin, int pout, const T* in, T* out, T scale, T* A, T* B) {
  const int pm = pin > pout ? pin : pout;
  const int Ki = (pin + 1) * (pin + 1);
  for (int k = threadIdx.x; k < (pm + 1) * (pm + 1); k += blockDim.x) {
    if (k < Ki) {
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      int m = k - n * n - n, am = m < 0 ? -m : m;
      T er = (T)phase[2 * am], ei = (T)phase[2 * am + 1];
      if (m < 0) ei = -ei;
      T xr = in[2 * k], xi = in[2 * k + 1];
      A[2 * k] = xr * er - xi * ei;
      A[2 * k + 1] = xr * ei + xi * er;
    } else {
      A[2 * k] = 0;
      A[2 * k + 1] = 0;
    }
  }
  __syncthreads();
  rotate_block<T>(Tb, rc, pin, A, B, false);
  __syncthreads();
  zshift_block<T>(Z, Tb.p, pin, pout, B, A);
  __syncthreads();
  rotate_block<T>(Tb, rc, pout, A, B, true);
  __syncthreads();
  const int Ko = (pout + 1) * (pout + 1);
  for (int k = threadIdx.x; k < Ko; k += blockDim.x) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    int m = k - n * n - n, am = m < 0 ? -m : m;
    T er = (T)phase[2 * am], ei = -(T)phase[2 * am + 1];  // e^{-i m phi}
    if (m < 0) ei = -ei;
    T xr = B[2 * k], xi = B[2 * k + 1];
    out[2 * k] += (xr * er - xi * ei) * scale;
    out[2 * k + 1] += (xr * ei + xi * er) * scale;
  }
  __syncthreads();
}

// M2M for the parents of one level (src/lfmm3drouts.f:195-258)
template <typename T>
__global__ void k_m2m(DevPlan P, DevTables Tb, int first, T* __restrict__ mp) {
  extern __shared__ unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + 2 * P.K;
  const int b = P.m2m_boxes[first + blockIdx.x];
  const int p = P.p;
  for (int c = 0; c < P.nchild[b]; c++) {
    const int ch = P.child[8 * b + c];
    if (P.nsrc[ch] == 0) continue;
    // octant of the child relative to the parent
    int o = (P.ijk[3 * ch] & 1 ? 0 : 1) | (P.ijk[3 * ch + 1] & 1 ? 0 : 2) | (P.ijk[3 * ch + 2] & 1 ? 0 : 4);
    const double* ph = Tb.oct_phase + (size_t)(o * 2 + 0) * 2 * (p + 1);
    point_and_shoot<T>(Tb, Tb.oct_rot[o * 2 + 0], ph, Tb.z_m2m, p, p, mp + 2 * (size_t)ch * P.K,
                       mp + 2 * (size_t)b * P.K, T(1), A, B);
  }
}

// L2L for the parents of one level (src/lfmm3drouts.f:498-558)
template <typename T>
__global__ void k_l2l(DevPlan P, DevTables Tb, int first, T* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + 2 * P.K;
  const int b = P.l2l_boxes[first + blockIdx.x];
  const int p = P.p;
  for (int c = 0; c < P.nchild[b]; c++) {
    const int ch = P.child[8 * b + c];
    int o = (P.ijk[3 * ch] & 1 ? 0 : 1) | (P.ijk[3 * ch + 1] & 1 ? 0 : 2) | (P.ijk[3 * ch + 2] & 1 ? 0 : 4);
    const double* ph = Tb.oct_phase + (size_t)(o * 2 + 1) * 2 * (p + 1);
    point_and_shoot<T>(Tb, Tb.oct_rot[o * 2 + 1], ph, Tb.z_l2l, p, p, lo + 2 * (size_t)b * P.K,
                       lo + 2 * (size_t)ch * P.K, T(1), A, B);
  }
}

// M2L over the (pruned) list 2 of each target box (src/lfmm3drouts.f:347-483).
// The target's local accumulates in LDS (the third 2K block) and is added to
// memory once, so HBM sees one read per source multipole and one
// read-modify-write per target.  Without LDS room (LDSACC = false, high
// orders) the local is accumulated in memory directly.
template <typename T, bool LDSACC>
__global__ void k_m2l(DevPlan P, DevTables Tb, const T* __restrict__ mp, T* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + 2 * P.K;
  const int b = P.m2l_boxes[blockIdx.x];
  T* acc = LDSACC ? B + 2 * P.K : lo + 2 * (size_t)b * P.K;
  const int p = P.p;
  const T ih = (T)P.inv_side[P.level[b]];
  const size_t zsz = (size_t)(p + 1) * (p + 1) * (p + 1);
  if (LDSACC)
    for (int k = threadIdx.x; k < 2 * P.K; k += blockDim.x) acc[k] = T(0);
  __syncthreads();
  for (int jj = P.l2_off[b]; jj < P.l2_off[b + 1]; jj++) {
    const int j = P.l2_idx[jj];
    int di = P.ijk[3 * j] - P.ijk[3 * b] + 3, dj = P.ijk[3 * j + 1] - P.ijk[3 * b + 1] + 3,
        dk = P.ijk[3 * j + 2] - P.ijk[3 * b + 2] + 3;
    int code = (di * 7 + dj) * 7 + dk;
    int pt = Tb.off_ptrunc[code];
    point_and_shoot<T>(Tb, Tb.off_rot[code], Tb.off_phase + (size_t)code * 2 * (p + 1),
                       Tb.z_m2l + (size_t)Tb.off_dist[code] * zsz, pt, pt,
                       mp + 2 * (size_t)j * P.K, acc, ih, A, B);
  }
  if (LDSACC) {
    T* out = lo + 2 * (size_t)b * P.K;
    for (int k = threadIdx.x; k < 2 * P.K; k += blockDim.x) out[k] += acc[k];
  }
}

// ---------------------------------------------------------------------------
// M2L grouped by translation class.  List-2 sources of one target whose
// offsets share (dx^2 + dy^2, dz) share the rotation block, the z-shift table
// and the truncation order; only the azimuthal phase differs.  Up to
// kM2LGroup of them are translated together: the gs lanes of one work item
// hold the group's members and read the same table entry, so every table
// entry is fetched once per group, and each lane produces four consecutive
// orders m' (rotation) or degrees l (z-shift) of one member, so one LDS read
// of an input coefficient feeds four FMAs per output part.  Stages run in
// place (results held in registers across a barrier); the back rotation
// applies each member's phase, sums the members across lanes in a fixed order
// and adds the sum to the target's local in LDS.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ const T* grp_rot_tab(const DevTables& Tb);
template <>
__device__ __forceinline__ const double* grp_rot_tab<double>(const DevTables& Tb) { return Tb.rot_half; }
template <>
__device__ __forceinline__ const float* grp_rot_tab<float>(const DevTables& Tb) { return Tb.rot_half_f; }
template <typename T>
__device__ __forceinline__ const T* grp_z_tab(const DevTables& Tb);
template <>
__device__ __forceinline__ const double* grp_z_tab<double>(const DevTables& Tb) { return Tb.z_m2l_t; }
template <>
__device__ __forceinline__ const float* grp_z_tab<float>(const DevTables& Tb) { return Tb.z_m2l_t_f; }

template <typename T, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_m2l_grp(DevPlan P, DevTables Tb, int gmax, const T* __restrict__ mp,
                                                  T* __restrict__ lo) {
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
  const T ih = (T)P.inv_side[P.level[b]];
  for (int k = tid; k < K2 + 8 + gmax * MS; k += bs) acc[k] = T(0);
  for (int k = tid; k < (p + 4) * (p + 4); k += bs) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    s_kn[k] = n;
  }
  const T* rtab = grp_rot_tab<T>(Tb);
  const T* ztab = grp_z_tab<T>(Tb);
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
      const int pt = Tb.off_ptrunc[code0];
      const int rc = Tb.off_rot[code0];
      const T* Z = ztab + (size_t)Tb.off_dist[code0] * Tb.zt_size;
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
      const int NI = Tb.grp_rcount[pt], U0 = bs >> lg;
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
          const int it = Tb.grp_ritems[uq[q]], n = it >> 8, qc = it & 255;
          grp_rot_item<T, false>(rtab + Tb.rot_half_off[rc * P1 + n] + 16 * qc, n / 4 + 1, n, qc, A + jj * MS, r[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ITEMS; q++)
        if (uq[q] >= 0 && mem) {
          const int it = Tb.grp_ritems[uq[q]];
          grp_rot_store<T>(it >> 8, 4 * (it & 255), r[q], A + jj * MS);
        }
      __syncthreads();
      // (c) shift along z
      const int zo = Tb.grp_zoff[pt];
#pragma unroll
      for (int q = 0; q < ITEMS; q++) {
#pragma unroll
        for (int e = 0; e < 16; e++) r[q][e] = T(0);
        if (uq[q] >= 0 && mem) {
          const int it = Tb.grp_zitems[zo + uq[q]], am = it >> 8;
          grp_z_item<T>(Z + Tb.zt_off[am], (p - am) / 4 + 1, am, it & 255, pt, A + jj * MS, r[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ITEMS; q++)
        if (uq[q] >= 0 && mem) {
          const int it = Tb.grp_zitems[zo + uq[q]], am = it >> 8;
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
          const int it = Tb.grp_ritems[uq[q]];
          n = it >> 8;
          mp0 = 4 * (it & 255);
          if (mem) {
            grp_rot_item<T, true>(rtab + Tb.rot_half_off[rc * P1 + n] + 4 * mp0, n / 4 + 1, n, mp0 / 4,
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
// M2L with factored rotations.  A rotation by theta about y is
//   d^n(theta) = P Delta^n E(theta) Delta^nT P*,  Delta^n = d^n(pi/2),
//   P = diag(i^m), E = diag(e^{i m theta}),
// so every pair uses the same Delta blocks (they stay in cache) and only a
// diagonal phase depends on the pair; each rotation costs two products with
// Delta instead of one with a per-class matrix streamed from memory.
// Both products read Delta column-wise through d_ab = (-1)^{a-b} d_ba.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mod4(int m) { return ((m % 4) + 4) % 4; }

// multiply (re, im) by i^r, r = 0..3
template <typename T>
__device__ __forceinline__ void mul_ipow(T& re, T& im, int r) {
  T a = re, b = im;
  if (r == 1) { re = -b; im = a; }
  else if (r == 2) { re = -a; im = -b; }
  else if (r == 3) { re = b; im = -a; }
}

// out_k = sum_m Delta^n_{mk} in_m for every (n, k) with n <= pmax
// (transpose product), then out_k *= e^{i s k theta} (s = +1 or -1).
template <typename T>
__device__ void delta_t_times(const DevTables& Tb, int pmax, const double* e, int s, const T* in, T* out) {
  const int Kt = (pmax + 1) * (pmax + 1);
  for (int idx = threadIdx.x; idx < Kt; idx += blockDim.x) {
    int n = (int)sqrtf((float)idx);
    while (n * n > idx) n--;
    while ((n + 1) * (n + 1) <= idx) n++;
    const int k = idx - n * n - n, dim = 2 * n + 1;
    const double* D = Tb.delta + Tb.delta_off[n];
    T ar = 0, ai = 0;
    for (int m = -n; m <= n; m++) {
      const T d = (T)D[(m + n) * dim + (k + n)];
      ar += d * in[2 * eidx(n, m)];
      ai += d * in[2 * eidx(n, m) + 1];
    }
    const int ak = k < 0 ? -k : k;
    const T er = (T)e[2 * ak];
    T ei = (T)e[2 * ak + 1];
    if ((k < 0) != (s < 0)) ei = -ei;
    out[2 * idx] = ar * er - ai * ei;
    out[2 * idx + 1] = ar * ei + ai * er;
  }
}

// out_a = sum_b Delta^n_{ab} in_b = sum_b (-1)^{a-b} Delta^n_{ba} in_b
template <typename T>
__device__ void delta_times(const DevTables& Tb, int pmax, const T* in, T* out) {
  const int Kt = (pmax + 1) * (pmax + 1);
  for (int idx = threadIdx.x; idx < Kt; idx += blockDim.x) {
    int n = (int)sqrtf((float)idx);
    while (n * n > idx) n--;
    while ((n + 1) * (n + 1) <= idx) n++;
    const int a = idx - n * n - n, dim = 2 * n + 1;
    const double* D = Tb.delta + Tb.delta_off[n];
    T ar = 0, ai = 0;
    for (int b = -n; b <= n; b++) {
      T d = (T)D[(b + n) * dim + (a + n)];
      if ((a - b) & 1) d = -d;
      ar += d * in[2 * eidx(n, b)];
      ai += d * in[2 * eidx(n, b) + 1];
    }
    out[2 * idx] = ar;
    out[2 * idx + 1] = ai;
  }
}

// x_m *= (-i)^m e_m over degrees <= pmax (input factor of a factored rotation)
template <typename T>
__device__ void pre_rot(int pmax, T* x) {
  const int Kt = (pmax + 1) * (pmax + 1);
  for (int idx = threadIdx.x; idx < Kt; idx += blockDim.x) {
    int n = (int)sqrtf((float)idx);
    while (n * n > idx) n--;
    while ((n + 1) * (n + 1) <= idx) n++;
    const int m = idx - n * n - n;
    T re = x[2 * idx], im = x[2 * idx + 1];
    mul_ipow(re, im, mod4(-m));
    if (m > 0 && (m & 1)) { re = -re; im = -im; }
    x[2 * idx] = re;
    x[2 * idx + 1] = im;
  }
}

template <typename T, bool LDSACC>
__global__ void k_m2l_fact(DevPlan P, DevTables Tb, const T* __restrict__ mp, T* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + 2 * P.K;
  const int b = P.m2l_boxes[blockIdx.x];
  T* acc = LDSACC ? B + 2 * P.K : lo + 2 * (size_t)b * P.K;
  const int p = P.p;
  const T ih = (T)P.inv_side[P.level[b]];
  const size_t zsz = (size_t)(p + 1) * (p + 1) * (p + 1);
  if (LDSACC)
    for (int k = threadIdx.x; k < 2 * P.K; k += blockDim.x) acc[k] = T(0);
  __syncthreads();
  for (int jj = P.l2_off[b]; jj < P.l2_off[b + 1]; jj++) {
    const int j = P.l2_idx[jj];
    const int di = P.ijk[3 * j] - P.ijk[3 * b] + 3, dj = P.ijk[3 * j + 1] - P.ijk[3 * b + 1] + 3,
              dk = P.ijk[3 * j + 2] - P.ijk[3 * b + 2] + 3;
    const int code = (di * 7 + dj) * 7 + dk;
    const int pt = Tb.off_ptrunc[code];
    const int Ki = (pt + 1) * (pt + 1);
    const double* phase = Tb.off_phase + (size_t)code * 2 * (p + 1);
    const double* e = Tb.rot_e + (size_t)Tb.off_rot[code] * 2 * (p + 1);
    const double* Z = Tb.z_m2l + (size_t)Tb.off_dist[code] * zsz;
    const T* in = mp + 2 * (size_t)j * P.K;
    // A = in e^{i m phi} (-i)^m e_m
    for (int k = threadIdx.x; k < Ki; k += blockDim.x) {
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      const int m = k - n * n - n, am = m < 0 ? -m : m;
      const T er = (T)phase[2 * am];
      const T ei = m < 0 ? -(T)phase[2 * am + 1] : (T)phase[2 * am + 1];
      const T xr = in[2 * k], xi = in[2 * k + 1];
      T re = xr * er - xi * ei, im = xr * ei + xi * er;
      mul_ipow(re, im, mod4(-m));
      if (m > 0 && (m & 1)) { re = -re; im = -im; }
      A[2 * k] = re;
      A[2 * k + 1] = im;
    }
    __syncthreads();
    delta_t_times<T>(Tb, pt, e, +1, A, B);   // B_k = e^{ik theta} sum_m Delta_mk A_m
    __syncthreads();
    delta_times<T>(Tb, pt, B, A);            // A_a = sum_k Delta_ak B_k
    __syncthreads();
    // the rotated vector is e_a i^a A_a; fold those factors into the shift
    for (int k = threadIdx.x; k < Ki; k += blockDim.x) {
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      const int m = k - n * n - n;
      T re = A[2 * k], im = A[2 * k + 1];
      mul_ipow(re, im, mod4(m));
      if (m > 0 && (m & 1)) { re = -re; im = -im; }
      A[2 * k] = re;
      A[2 * k + 1] = im;
    }
    __syncthreads();
    zshift_block<T>(Z, p, pt, pt, A, B);
    __syncthreads();
    pre_rot<T>(pt, B);
    __syncthreads();
    delta_t_times<T>(Tb, pt, e, -1, B, A);   // back rotation: conj(E)
    __syncthreads();
    delta_times<T>(Tb, pt, A, B);
    __syncthreads();
    // acc += e_a i^a B_a e^{-i a phi} / h
    for (int k = threadIdx.x; k < Ki; k += blockDim.x) {
      int n = (int)sqrtf((float)k);
      while (n * n > k) n--;
      while ((n + 1) * (n + 1) <= k) n++;
      const int m = k - n * n - n, am = m < 0 ? -m : m;
      T re = B[2 * k], im = B[2 * k + 1];
      mul_ipow(re, im, mod4(m));
      if (m > 0 && (m & 1)) { re = -re; im = -im; }
      const T er = (T)phase[2 * am];
      const T ei = m < 0 ? (T)phase[2 * am + 1] : -(T)phase[2 * am + 1];
      acc[2 * k] += (re * er - im * ei) * ih;
      acc[2 * k + 1] += (re * ei + im * er) * ih;
    }
    __syncthreads();
  }
  if (LDSACC) {
    T* out = lo + 2 * (size_t)b * P.K;
    for (int k = threadIdx.x; k < 2 * P.K; k += blockDim.x) out[k] += acc[k];
  }
}

// ---------------------------------------------------------------------------
// P2P near field (src/lfmm3dpart.f:1076-1141, lfmm3dpart_direct_self/_targ).
// One workgroup per (leaf, chunk of evaluation points); one point per thread.
// The source loop is wave-uniform, so positions and strengths stream through
// the scalar cache and every lane reuses them from SGPRs.  Offsets and r^2
// are formed in FP64 for every precision; with TN = float the rest of the
// pair is FP32, summed in FP32 over runs of at most kFold sources and folded
// into FP64 accumulators, except for the closest pairs (r^2 < rc2, r below
// 1/32 of the leaf side): such a pair dominates the sums at its point, and a
// dipole's v.d can cancel, which multiplies the FP32 rounding of v and d by
// |v||d|/|v.d|.  The FP32 loop only keeps the smallest r^2 of the run (one
// min per pair, no branch, so the unrolled source loads stay batched); a
// point whose run held a close pair drops that run's FP32 sums and runs it
// again with the close pairs in FP64 from the FP64 strengths (redo_run),
// which happens for a few runs only.
// ---------------------------------------------------------------------------

template <typename TN, bool CHG, bool DIP, bool FLD>
struct PairAcc {
  TN pr = 0, pi = 0, fxr = 0, fxi = 0, fyr = 0, fyi = 0, fzr = 0, fzi = 0;
  // pot += q/r + (v.d)/r^3 = (q + (v.d)/r^2)/r;
  // fld += (q + 3 (v.d)/r^2) d/r^3 - v/r^3 = (c d - v)/r^3, d = x_t - x_s
  // returns r^2 in TN
  __device__ __forceinline__ TN add(double xt, double yt, double zt, const double* __restrict__ sp,
                                    const TN* __restrict__ st, bool mask) {
    const double dxd = xt - sp[0], dyd = yt - sp[1], dzd = zt - sp[2];
    const double r2d = fma(dxd, dxd, fma(dyd, dyd, dzd * dzd));
    const TN dx = (TN)dxd, dy = (TN)dyd, dz = (TN)dzd, r2 = (TN)r2d;
    TN ri = rsqrt_t<TN>(r2);
    if (mask) ri = TN(0);
    const TN ri2 = ri * ri;
    if (!DIP) {
      pr = fma(st[0], ri, pr);
      pi = fma(st[1], ri, pi);
      if (FLD) {
        const TN ri3 = ri * ri2;
        const TN cr = st[0] * ri3, ci = st[1] * ri3;
        fxr = fma(cr, dx, fxr); fxi = fma(ci, dx, fxi);
        fyr = fma(cr, dy, fyr); fyi = fma(ci, dy, fyi);
        fzr = fma(cr, dz, fzr); fzi = fma(ci, dz, fzi);
      }
      return r2;
    }
    const TN dotr = fma(st[2], dx, fma(st[4], dy, st[6] * dz));
    const TN doti = fma(st[3], dx, fma(st[5], dy, st[7] * dz));
    const TN ar = CHG ? fma(dotr, ri2, st[0]) : dotr * ri2;
    const TN ai = CHG ? fma(doti, ri2, st[1]) : doti * ri2;
    pr = fma(ar, ri, pr);
    pi = fma(ai, ri, pi);
    if (FLD) {
      const TN ri3 = ri * ri2, t3 = TN(3) * ri2;
      const TN cr = CHG ? fma(t3, dotr, st[0]) : t3 * dotr;
      const TN ci = CHG ? fma(t3, doti, st[1]) : t3 * doti;
      fxr = fma(ri3, fma(cr, dx, -st[2]), fxr); fxi = fma(ri3, fma(ci, dx, -st[3]), fxi);
      fyr = fma(ri3, fma(cr, dy, -st[4]), fyr); fyi = fma(ri3, fma(ci, dy, -st[5]), fyi);
      fzr = fma(ri3, fma(cr, dz, -st[6]), fzr); fzi = fma(ri3, fma(ci, dz, -st[7]), fzi);
    }
    return r2;
  }
  __device__ __forceinline__ void fold_into(double* a) {
    a[0] += (double)pr; a[1] += (double)pi;
    if (FLD) {
      a[2] += (double)fxr; a[3] += (double)fxi; a[4] += (double)fyr;
      a[5] += (double)fyi; a[6] += (double)fzr; a[7] += (double)fzi;
    }
    pr = pi = fxr = fxi = fyr = fyi = fzr = fzi = TN(0);
  }
};

// A run of sources s0..s1-1 again at (xt, yt, zt) after its FP32 sums were
// dropped: the close pairs (r^2 < rc2, not source `self`) in FP64 from the
// FP64 strengths into tot, the rest in FP32 into acc as before.
template <typename TN, bool CHG, bool DIP, bool FLD>
__device__ __forceinline__ void redo_run(PairAcc<TN, CHG, DIP, FLD>& acc, double xt, double yt, double zt,
                                         const double* __restrict__ spos, const TN* __restrict__ str,
                                         const double* __restrict__ str64, int s0, int s1, int self, double rc2,
                                         double* tot) {
  acc = PairAcc<TN, CHG, DIP, FLD>();
  PairAcc<double, CHG, DIP, FLD> c64;
  for (int s = s0; s < s1; s++) {
    const double* sp = spos + 4 * (size_t)s;
    const double dx = xt - sp[0], dy = yt - sp[1], dz = zt - sp[2];
    if (s != self && fma(dx, dx, fma(dy, dy, dz * dz)) < rc2)
      c64.add(xt, yt, zt, sp, str64 + 8 * (size_t)s, false);
    else
      acc.add(xt, yt, zt, sp, str + 8 * (size_t)s, s == self);
  }
  c64.fold_into(tot);
}

template <typename TN, bool CHG, bool DIP, bool FLD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_p2p(DevPlan P, const double* __restrict__ spos,
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
  double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  PairAcc<TN, CHG, DIP, FLD> acc;
  constexpr bool fold = sizeof(TN) < sizeof(double);
  // FP32 pairs: pairs closer than 1/32 of this leaf's side run in FP64 (on a
  // uniform 1M cube 0.05 such pairs per point; the near-field error of the
  // FP32 pairs falls from 1.5e-7 to 8e-8 of the potential, and on a sphere
  // with a target next to a dipole from 7e-7 to 3e-9)
  const double hside = 1.0 / P.inv_side[P.level[b]];
  const double rc2 = fold ? hside * hside * (1.0 / 1024.0) : 0.0;
  const TN rc2n = (TN)rc2, kFar = TN(1e30);
  // own box first: a source never acts on itself (index test, lfmm3dpart.f:1193)
  {
    const int s0 = P.isrc[b], s1 = s0 + P.nsrc[b];
    const int self = kind ? -1 : me;
    for (int c = s0; c < s1; c += (fold ? kFold : (1 << 30))) {
      const int ce = fold ? min(c + kFold, s1) : s1;
      TN mn = kFar;
      for (int s = c; s < ce; s++) {
        const TN r2 = acc.add(x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, s == self);
        if (fold) mn = fmin(mn, s == self ? kFar : r2);
      }
      if (fold) {
        if (mn < rc2n) redo_run(acc, x, y, z, spos, str, str64, c, ce, self, rc2, tot);
        acc.fold_into(tot);
      }
    }
  }
  for (int jj = P.l1_off[b]; jj < P.l1_off[b + 1]; jj++) {
    const int j = P.l1_idx[jj];
    const int s0 = P.isrc[j], s1 = s0 + P.nsrc[j];
    for (int c = s0; c < s1; c += (fold ? kFold : (1 << 30))) {
      const int ce = fold ? min(c + kFold, s1) : s1;
      TN mn = kFar;
#pragma unroll 4
      for (int s = c; s < ce; s++) {
        const TN r2 = acc.add(x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, false);
        if (fold) mn = fmin(mn, r2);
      }
      if (fold) {
        if (mn < rc2n) redo_run(acc, x, y, z, spos, str, str64, c, ce, -1, rc2, tot);
        acc.fold_into(tot);
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

// Plain direct sum over all sources (l3dpartdirect, src/lfmm3dpart.f:1391):
// the same pair arithmetic in FP64 with every source acting on every point.
template <bool CHG, bool DIP>
__global__ __launch_bounds__(256) void k_direct(int ns, const double* __restrict__ spos,
                                                const double* __restrict__ str, int npts,
                                                const double* __restrict__ pts, int pts_are_sources,
                                                double* __restrict__ pot, double* __restrict__ fld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < npts;
  const int me = active ? i : 0;
  const double* mp = (pts_are_sources ? spos : pts) + 4 * (size_t)me;
  const double x = mp[0], y = mp[1], z = mp[2];
  PairAcc<double, CHG, DIP, true> acc;
  const int self = pts_are_sources ? me : -1;
  for (int s = 0; s < ns; s++) acc.add(x, y, z, spos + 4 * (size_t)s, str + 8 * (size_t)s, s == self);
  if (!active) return;
  double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  acc.fold_into(t);
  pot[2 * (size_t)i] = t[0]; pot[2 * (size_t)i + 1] = t[1];
  for (int k = 0; k < 6; k++) fld[6 * (size_t)i + k] = t[2 + k];
}

}  // namespace fmm3d
