// Device helpers shared by the Laplace and Helmholtz kernels: complex
// arithmetic, the factor table, solid-harmonic recurrences, rotations of
// expansions, and the gather / scatter / fill kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "lap_kernels.hpp"

namespace fmm3d {

// ---------------------------------------------------------------------------
// small complex helpers
// ---------------------------------------------------------------------------
template <typename T>
struct C2 {
  T re, im;
};
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ void cfma(C2<T>& acc, C2<T> a, C2<T> b) {
  acc.re += a.re * b.re - a.im * b.im;
  acc.im += a.re * b.im + a.im * b.re;
}
template <typename T>
__device__ __forceinline__ C2<T> cld(const T* p, int k) { return {p[2 * k], p[2 * k + 1]}; }
template <typename T>
__device__ __forceinline__ void cst(T* p, int k, C2<T> v) { p[2 * k] = v.re; p[2 * k + 1] = v.im; }
template <typename T>
__device__ __forceinline__ void cadd(T* p, int k, C2<T> v) { p[2 * k] += v.re; p[2 * k + 1] += v.im; }

// Factor table layout (built on the host for an order p, doubles):
//   [F_MM + m]      sqrt((2m-1)/(2m))          m=1..p+1
//   [F_M1 + m]      sqrt(2m+1)                 m=0..p+1
//   [F_RA + h(n,m)] sqrt((n+m-1)(n-m-1))       n>=m+2 (recurrence)
//   [F_RB + h(n,m)] 1/sqrt((n-m)(n+m))
//   [F_AR + e(n,m)], [F_BR..], [F_CR..]  derivative factors of R (n<=p+1)
//   [F_AS + e(n,m)], [F_BS..], [F_CS..]  derivative factors of S (n<=p+1)
struct FacLayout {
  int MM, M1, RA, RB, AR, BR, CR, AS, BS, CS, total;
  __host__ __device__ explicit FacLayout(int p) {
    int P = p + 2;                       // tables cover degrees up to p+1
    int H = P * (P + 1) / 2, E = P * P;
    MM = 0; M1 = MM + P + 1; RA = M1 + P + 1; RB = RA + H; AR = RB + H;
    BR = AR + E; CR = BR + E; AS = CR + E; BS = AS + E; CS = BS + E; total = CS + E;
  }
};

// Regular solid harmonics R_n^m(u), 0<=m<=n<=pmax, for one point into a half
// table (stride `st` between entries).  Column recurrence in n per m.
template <typename T>
__device__ __forceinline__ void regular_column(const double* fac, const FacLayout& L, int pmax,
                                               int m, T ux, T uy, T uz, T r2, T* out, int st) {
  // R_m^m = prod_k (-(ux+i uy)) sqrt((2k-1)/(2k))
  C2<T> rmm = {T(1), T(0)};
  for (int k = 1; k <= m; k++) {
    T f = (T)fac[L.MM + k];
    C2<T> w = {-ux * f, -uy * f};
    rmm = cmul(w, rmm);
  }
  cst(out, hidx(m, m) * st / 1, rmm);
  if (m + 1 > pmax) return;
  C2<T> r1 = {(T)fac[L.M1 + m] * uz * rmm.re, (T)fac[L.M1 + m] * uz * rmm.im};
  cst(out, hidx(m + 1, m) * st, r1);
  C2<T> r0 = rmm;
  for (int n = m + 2; n <= pmax; n++) {
    T a = (T)(2 * n - 1) * uz, b = (T)fac[L.RA + hidx(n, m)] * r2, c = (T)fac[L.RB + hidx(n, m)];
    C2<T> rn = {(a * r1.re - b * r0.re) * c, (a * r1.im - b * r0.im) * c};
    cst(out, hidx(n, m) * st, rn);
    r0 = r1;
    r1 = rn;
  }
}

// Irregular solid harmonics S_n^m(u) for 0<=m<=n<=pmax.
template <typename T>
__device__ __forceinline__ void irregular_column(const double* fac, const FacLayout& L, int pmax,
                                                 int m, T ux, T uy, T uz, T ir2, T ir, T* out,
                                                 int st) {
  C2<T> smm = {ir, T(0)};
  for (int k = 1; k <= m; k++) {
    T f = (T)fac[L.MM + k] * ir2;
    C2<T> w = {-ux * f, -uy * f};
    smm = cmul(w, smm);
  }
  cst(out, hidx(m, m) * st, smm);
  if (m + 1 > pmax) return;
  T g = (T)fac[L.M1 + m] * uz * ir2;
  C2<T> s1 = {g * smm.re, g * smm.im};
  cst(out, hidx(m + 1, m) * st, s1);
  C2<T> s0 = smm;
  for (int n = m + 2; n <= pmax; n++) {
    T a = (T)(2 * n - 1) * uz, b = (T)fac[L.RA + hidx(n, m)], c = (T)fac[L.RB + hidx(n, m)] * ir2;
    C2<T> sn = {(a * s1.re - b * s0.re) * c, (a * s1.im - b * s0.im) * c};
    cst(out, hidx(n, m) * st, sn);
    s0 = s1;
    s1 = sn;
  }
}

// value of a harmonic with signed order from a half table
template <typename T>
__device__ __forceinline__ C2<T> half_get(const T* tab, int n, int m, int st) {
  if (m >= 0) return cld(tab, hidx(n, m) * st);
  return cconj(cld(tab, hidx(n, -m) * st));
}

// ---------------------------------------------------------------------------
// gather / scatter / fill
// ---------------------------------------------------------------------------

// Sorted sources: positions in FP64 (x, y, z, 0) for every precision, and
// strengths as 8 values (q.re q.im vx.re vx.im vy.re vy.im vz.re vz.im)
// where v = dipstr * dipvec is the complex dipole moment vector; an FP32 copy
// of the strengths is written when the near field runs in FP32.
static __global__ void k_gather_sources(int ns, const int* __restrict__ perm, const double* __restrict__ src,
                                 const double* __restrict__ q, const double* __restrict__ ds,
                                 const double* __restrict__ dv, int ifcharge, int ifdipole,
                                 double* __restrict__ spos, double* __restrict__ str64,
                                 float* __restrict__ str32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  int o = perm[i];
  double* x = spos + 4 * (size_t)i;
  x[0] = src[3 * o]; x[1] = src[3 * o + 1]; x[2] = src[3 * o + 2]; x[3] = 0.0;
  double v[8];
  v[0] = ifcharge ? q[2 * o] : 0.0;
  v[1] = ifcharge ? q[2 * o + 1] : 0.0;
  for (int k = 0; k < 3; k++) {
    double d = ifdipole ? dv[3 * o + k] : 0.0;
    v[2 + 2 * k] = ifdipole ? ds[2 * o] * d : 0.0;
    v[3 + 2 * k] = ifdipole ? ds[2 * o + 1] * d : 0.0;
  }
  for (int k = 0; k < 8; k++) str64[8 * (size_t)i + k] = v[k];
  if (str32)
    for (int k = 0; k < 8; k++) str32[8 * (size_t)i + k] = (float)v[k];
}

static __global__ void k_gather_targets(int nt, const int* __restrict__ perm, const double* __restrict__ trg,
                                 double* __restrict__ tpos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  int o = perm[i];
  tpos[4 * (size_t)i] = trg[3 * o];
  tpos[4 * (size_t)i + 1] = trg[3 * o + 1];
  tpos[4 * (size_t)i + 2] = trg[3 * o + 2];
  tpos[4 * (size_t)i + 3] = 0.0;
}

// out[perm[i]] = sorted[i]; `w` complex values per point (1 = pot, 3 = fld)
static __global__ void k_scatter(int n, int w, const int* __restrict__ perm, const double* __restrict__ sorted,
                          double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = perm[i];
  for (int k = 0; k < 2 * w; k++) out[(size_t)2 * w * o + k] = sorted[(size_t)2 * w * i + k];
}

// out[perm[i]] = far[i] + near[i]: the far-field and near-field sums of a
// point, added in that order
static __global__ void k_scatter_sum(int n, int w, const int* __restrict__ perm, const double* __restrict__ far,
                                     const double* __restrict__ near, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = perm[i];
  for (int k = 0; k < 2 * w; k++)
    out[(size_t)2 * w * o + k] = far[(size_t)2 * w * i + k] + near[(size_t)2 * w * i + k];
}

// Sums of the list-3 slices of one evaluation item (box, kind, first point,
// first slice, number of slices), 8 values per point (pot, d/dz, U, D as
// complex), added in slice order to the outputs: pot += pot, fld -= grad
// with d/dx = (U+D)/2, d/dy = (U-D)/(2i).  One 64-thread workgroup per item.
static __global__ void k_slice_reduce(const int* __restrict__ red, const double* __restrict__ partial,
                                      const int* __restrict__ isrc, const int* __restrict__ nsrc,
                                      const int* __restrict__ itrg, const int* __restrict__ ntrg, int wpot0,
                                      int wfld0, int wpot1, int wfld1, double* __restrict__ pot0,
                                      double* __restrict__ fld0, double* __restrict__ pot1,
                                      double* __restrict__ fld1) {
  const int* r = red + 5 * blockIdx.x;
  const int b = r[0], kind = r[1], start = r[2], s0 = r[3], ns = r[4];
  const int wp = kind ? wpot1 : wpot0, wf = kind ? wfld1 : wfld0;
  const int n0 = kind ? itrg[b] : isrc[b];
  const int nn = kind ? ntrg[b] : nsrc[b];
  const int i = start + threadIdx.x;
  if (i >= nn) return;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < ns; c++) {
    const double* o = partial + 8 * ((size_t)(s0 + c) * 64 + threadIdx.x);
    for (int e = 0; e < 8; e++) v[e] += o[e];
  }
  double* po = kind ? pot1 : pot0;
  double* fo = kind ? fld1 : fld0;
  const size_t id = (size_t)(n0 + i);
  if (wp) { po[2 * id] += v[0]; po[2 * id + 1] += v[1]; }
  if (wf) {
    // v[2..7] = d/dz, U, D (complex)
    fo[6 * id + 0] -= 0.5 * (v[4] + v[6]); fo[6 * id + 1] -= 0.5 * (v[5] + v[7]);
    fo[6 * id + 2] -= 0.5 * (v[5] - v[7]); fo[6 * id + 3] -= -0.5 * (v[4] - v[6]);
    fo[6 * id + 4] -= v[2]; fo[6 * id + 5] -= v[3];
  }
}

template <typename T>
__global__ void k_fill_zero(size_t n, T* __restrict__ a) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = T(0);
}

// Rotation of the coefficients of one expansion (degrees <= pmax) by the
// rotation class `rc`: out_n^{m'} = sum_m D_{m m'} in_n^m, or the inverse
// rotation when `back` (same blocks, (-1)^{m+m'} signs).
template <typename T>
__device__ void rotate_block(const DevTables& Tb, int rc, int pmax, const T* in, T* out, bool back) {
  // The blocks satisfy D[-m][-m'] = D[m][m'], so one thread produces the
  // outputs m' and -m' of a degree from a single pass over column m' of D:
  //   out_{m'} = sum_m D[m][m'] in_m,   out_{-m'} = sum_m D[m][m'] in_{-m}.
  const int p = Tb.p;
  const int Kh = (pmax + 1) * (pmax + 2) / 2;
  for (int k = threadIdx.x; k < Kh; k += blockDim.x) {
    int n = (int)((sqrtf(8.0f * k + 1.0f) - 1.0f) * 0.5f);
    while (n * (n + 1) / 2 > k) n--;
    while ((n + 1) * (n + 2) / 2 <= k) n++;
    const int mp = k - n * (n + 1) / 2;   // 0..n
    const double* D = Tb.rot + Tb.rot_off[rc * (p + 1) + n];
    const int dim = 2 * n + 1;
    T pr = 0, pi = 0, qr = 0, qi = 0;
    for (int m = -n; m <= n; m++) {
      const T d = (T)D[(m + n) * dim + (mp + n)];
      T xr = in[2 * eidx(n, m)], xi = in[2 * eidx(n, m) + 1];
      T yr = in[2 * eidx(n, -m)], yi = in[2 * eidx(n, -m) + 1];
      if (back && (m & 1)) { xr = -xr; xi = -xi; yr = -yr; yi = -yi; }
      pr += d * xr;
      pi += d * xi;
      qr += d * yr;
      qi += d * yi;
    }
    if (back && (mp & 1)) { pr = -pr; pi = -pi; qr = -qr; qi = -qi; }
    out[2 * eidx(n, mp)] = pr;
    out[2 * eidx(n, mp) + 1] = pi;
    if (mp > 0) {
      out[2 * eidx(n, -mp)] = qr;
      out[2 * eidx(n, -mp) + 1] = qi;
    }
  }
}

// The block index remapped so that each XCD runs contiguous chunks of the
// grid: workgroups are handed to the 8 XCDs of an MI355X round-robin in
// block order, each XCD with its own L2, so block b runs on XCD b % 8.  In
// each group of 8 * kXcdChunk blocks, XCD x gets the kXcdChunk consecutive
// blocks x * kXcdChunk ..; a last, partial group keeps its order.  Work
// items in box (Morton) order then keep runs of neighbouring boxes -- which
// read the same sources or multipoles -- on one L2, while chunks of every
// part of the domain still go to every XCD (work per box varies in space).
constexpr int kXcdChunk = 64;
__device__ __forceinline__ int xcd_block() {
  const int b = (int)blockIdx.x, gsz = 8 * kXcdChunk;
  const int g = b / gsz;
  if ((g + 1) * gsz > (int)gridDim.x) return b;
  const int l = b - g * gsz;
  return g * gsz + (l & 7) * kXcdChunk + (l >> 3);
}

// Near-field precision helpers: the pair arithmetic after r^2 runs in TN
// (double or float); FP32 partial sums are folded into FP64 every kFold sources.
constexpr int kFold = 32;

template <typename TN>
__device__ __forceinline__ TN rsqrt_t(TN x);
// x^{-1/2} for normal positive x (callers mask x = 0 themselves).  FP64: the
// hardware estimate y refined by one second-order step of the Householder
// iteration for x^{-1/2}, e = 1 - x y^2, y <- y (1 + e/2 + 3e^2/8) (the
// Taylor series of (1 - e)^{-1/2}), which cubes the estimate's relative error.
// FP32: the hardware instruction alone (about 1 ulp), without the rescaling
// of subnormal arguments a library rsqrtf adds.
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) { return __builtin_amdgcn_rsqf(x); }

}  // namespace fmm3d
