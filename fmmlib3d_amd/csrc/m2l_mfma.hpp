// M2L on the matrix cores (Laplace).  Same work decomposition as k_m2l_grp:
// one workgroup per target box, its list-2 sources sorted by translation
// class (dx^2 + dy^2, dz), up to four same-class members translated together
// (src/lfmm3drouts.f:347-483 -> l3dmplocquad_trunc0, src/l3dtrans.f:868-986:
// phase, rotate, shift along z, rotate back, phase back).  Each of the three
// linear stages is a sequence of 16x16x4 MFMA products:
//
//   rows    = output orders m' = 0..n (rotation) or degrees l (z-shift),
//             16 per row tile;
//   columns = 16 = member (4) x part (4): real and imaginary parts of the
//             input at +m ("x") and at -m ("y"); the y columns produce the
//             negative output orders through F[-m'][-m] = F[m'][m];
//   k       = input orders (rotation) or input degrees (z-shift), 4 per
//             MFMA.
//
// The A operands come from tables stored in operand order (64 values per
// block, one per lane, build_mfma_tables); the B operands are read from the
// members in LDS.  A wave owns whole degrees (rotations) or whole orders
// (z-shift), so it writes its results back in place without waiting for the
// other waves; the back rotation applies e^{-i m' phi} and the scale, sums
// the members across lanes in a fixed order and adds into the target's local
// in LDS, which is added to HBM once.  The results do not depend on
// scheduling.
//
// Selected with FMM3D_M2L_ROTATION=mfma.  Measured on MI355X (1M + 1M
// uniform, iprec 2, M2L alone): 8.3 ms in FP32 and 13.6 ms in FP64, against
// 4.9 and 8.3 ms for the VALU form k_m2l_grp, which stays the default: a
// group holds at most four members (16 columns), so every A block feeds one
// product, and each wave walks its degrees one after another with a table
// round trip per chunk; the products are about a tenth of the run time.
#pragma once
#include <hip/hip_runtime.h>

#include "m2l_grp.hpp"

namespace fmm3d {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ const double* rot(const DevTables& Tb) { return Tb.mf_rot; }
  static __device__ __forceinline__ const double* z(const DevTables& Tb) { return Tb.mf_z; }
  // row of C value r held by a lane of row group kr (lane / 16): the FP64
  // product interleaves the four row groups
  static __device__ __forceinline__ int row(int kr, int r) { return kr + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ const float* rot(const DevTables& Tb) { return Tb.mf_rot_f; }
  static __device__ __forceinline__ const float* z(const DevTables& Tb) { return Tb.mf_z_f; }
  static __device__ __forceinline__ int row(int kr, int r) { return 4 * kr + r; }
};

// column blocks whose operands are loaded together
constexpr int kMfChunk = 8;

// wave that owns the t-th item of a stage (items in decreasing cost order,
// dealt out back and forth over the waves)
__device__ __forceinline__ int mf_owner(int t, int nw) {
  const int r = t % (2 * nw);
  return r < nw ? r : 2 * nw - 1 - r;
}

// Lane roles in the 16x16x4 operands: B row kr = lane / 16 (k index within
// the block), column = lane % 16 = member jj and part; C holds four rows of
// that column (Mfma<T>::row).
struct MfLane {
  int lane, kr, col, jj, part, sg, ri;
  __device__ __forceinline__ MfLane() {
    lane = (int)(threadIdx.x & 63);
    kr = lane >> 4;
    col = lane & 15;
    jj = col >> 2;
    part = col & 3;
    sg = part < 2 ? 1 : -1;
    ri = part & 1;
  }
};

// Degree n of a rotation for the members of the group (g live members at
// X + jj * MS).  BACK = false: the result is written back in place.  BACK =
// true: the result times e^{-i m' phi_jj} ih is summed over the members and
// added to acc.
template <typename T, int MAXRT, bool BACK>
__device__ __forceinline__ void mf_rot_degree(const T* __restrict__ tab, int n, T* X, int MS, int g,
                                              const T* s_ph, int P1, T ih, T* acc) {
  using V = typename Mfma<T>::v4;
  const MfLane L;
  const bool live = L.jj < g;
  const int nrt = n / 16 + 1, nkb = (2 * n) / 4 + 1;
  T* xb = X + L.jj * MS + 2 * (n * n + n) + L.ri;  // order m of this column's input: xb[2 * sg * m]
  V c[MAXRT];
#pragma unroll
  for (int rt = 0; rt < MAXRT; rt++) c[rt] = V{T(0), T(0), T(0), T(0)};
  // operands in chunks of kMfChunk column blocks: every load of a chunk is
  // issued before its products, so a chunk waits for memory once
  for (int kb0 = 0; kb0 < nkb; kb0 += kMfChunk) {
    T a[kMfChunk][MAXRT], b[kMfChunk];
#pragma unroll
    for (int u = 0; u < kMfChunk; u++) {
      const int kb = kb0 + u, kap = 4 * kb + L.kr;  // input order kap - n
      b[u] = (kb < nkb && live && kap <= 2 * n) ? xb[2 * L.sg * (kap - n)] : T(0);
#pragma unroll
      for (int rt = 0; rt < MAXRT; rt++)
        a[u][rt] = (kb < nkb && rt < nrt) ? tab[((size_t)rt * nkb + kb) * 64 + L.lane] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kMfChunk; u++)
      if (kb0 + u < nkb)
#pragma unroll
        for (int rt = 0; rt < MAXRT; rt++)
          if (rt < nrt) c[rt] = Mfma<T>::mma(a[u][rt], b[u], c[rt]);
  }
#pragma unroll
  for (int rt = 0; rt < MAXRT; rt++) {
    if (rt >= nrt) break;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int mq = 16 * rt + Mfma<T>::row(L.kr, r);    // output order m' >= 0 of this value
      const bool keep = mq <= n && !(L.part >= 2 && mq == 0);
      if (!BACK) {
        if (live && keep) xb[2 * L.sg * mq] = c[rt][r];
      } else {
        const T v = c[rt][r];
        const T w = __shfl_xor(v, 1);                    // the other part of the same complex value
        const int mc = mq <= n ? mq : n;
        const T er = s_ph[2 * (L.jj * P1 + mc)] * ih, ei = s_ph[2 * (L.jj * P1 + mc) + 1] * ih;
        // x: (pr + i pi) e^{-i m' phi}, y: (qr + i qi) e^{+i m' phi}
        T o = (L.part == 0 || L.part == 3) ? fma(w, ei, v * er) : fma(-w, ei, v * er);
        if (!live) o = T(0);
        o += __shfl_xor(o, 4);
        o += __shfl_xor(o, 8);
        if (L.col < 4 && keep) acc[2 * (n * n + n + L.sg * mq) + L.ri] += o;
      }
    }
  }
}

// Order am of the z-shift, degrees am..pt, in place (table blocks of the full
// order p; inputs of degrees pt+1..pt+3 are zero in LDS).
template <typename T, int MAXRT>
__device__ __forceinline__ void mf_z_order(const T* __restrict__ tab, int p, int pt, int am, T* X, int MS, int g) {
  using V = typename Mfma<T>::v4;
  const MfLane L;
  const bool live = L.jj < g;
  const int nlt = (pt - am) / 16 + 1, nkb = (pt - am) / 4 + 1, nkbf = (p - am) / 4 + 1;
  T* xb = X + L.jj * MS + 2 * L.sg * am + L.ri;      // degree n of this column's input: xb[2 * (n * n + n)]
  V c[MAXRT];
#pragma unroll
  for (int rt = 0; rt < MAXRT; rt++) c[rt] = V{T(0), T(0), T(0), T(0)};
  for (int kb0 = 0; kb0 < nkb; kb0 += kMfChunk) {
    T a[kMfChunk][MAXRT], b[kMfChunk];
#pragma unroll
    for (int u = 0; u < kMfChunk; u++) {
      const int kb = kb0 + u, n = am + 4 * kb + L.kr;  // input degree
      b[u] = (kb < nkb && live) ? xb[2 * (n * n + n)] : T(0);
#pragma unroll
      for (int rt = 0; rt < MAXRT; rt++)
        a[u][rt] = (kb < nkb && rt < nlt) ? tab[((size_t)rt * nkbf + kb) * 64 + L.lane] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kMfChunk; u++)
      if (kb0 + u < nkb)
#pragma unroll
        for (int rt = 0; rt < MAXRT; rt++)
          if (rt < nlt) c[rt] = Mfma<T>::mma(a[u][rt], b[u], c[rt]);
  }
#pragma unroll
  for (int rt = 0; rt < MAXRT; rt++) {
    if (rt >= nlt) break;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int l = am + 16 * rt + Mfma<T>::row(L.kr, r);
      if (live && l <= pt && !(L.part >= 2 && am == 0)) xb[2 * (l * l + l)] = c[rt][r];
    }
  }
}

// LDS bytes of k_m2l_mfma (the layout of k_m2l_grp with four members)
template <typename T>
inline size_t m2l_mfma_lds(int p) {
  const size_t K = (size_t)(p + 1) * (p + 1);
  return (6 * kM2LChunk + (size_t)grp_kn_ints(p)) * sizeof(int) +
         sizeof(T) * (2 * (size_t)kM2LGroup * (p + 1) + 2 * K + 8 + (size_t)kM2LGroup * grp_member_stride<T>(p));
}

template <typename T, int NT, int MAXRT>
__global__ __launch_bounds__(NT) void k_m2l_mfma(DevPlan P, DevTables Tb, const T* __restrict__ mp,
                                                 T* __restrict__ lo) {
  extern __shared__ unsigned char smem_raw[];
  int* s_key = reinterpret_cast<int*>(smem_raw);
  int* s_box = s_key + kM2LChunk;
  int* s_code = s_box + kM2LChunk;
  int* t_key = s_code + kM2LChunk;
  int* t_box = t_key + kM2LChunk;
  int* t_code = t_box + kM2LChunk;
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, bs = blockDim.x, wave = tid >> 6;
  const int b = P.m2l_boxes[blockIdx.x];
  const int p = P.p, P1 = p + 1, K2 = 2 * P.K;
  int* s_kn = t_code + kM2LChunk;
  T* s_ph = reinterpret_cast<T*>(s_kn + grp_kn_ints(p));
  T* acc = s_ph + 2 * kM2LGroup * P1;
  const int MS = grp_member_stride<T>(p);
  T* X = acc + K2 + 8;
  const T ih = (T)P.inv_side[P.level[b]];
  for (int k = tid; k < K2 + 8 + kM2LGroup * MS; k += bs) acc[k] = T(0);
  for (int k = tid; k < (p + 4) * (p + 4); k += bs) {
    int n = (int)sqrtf((float)k);
    while (n * n > k) n--;
    while ((n + 1) * (n + 1) <= k) n++;
    s_kn[k] = n;
  }
  const T* rtab = Mfma<T>::rot(Tb);
  const T* ztab = Mfma<T>::z(Tb);
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
      while (g < kM2LGroup && s0 + g < cn && t_key[s0 + g] == key0) g++;
      const int code0 = t_code[s0];
      const int pt = Tb.off_ptrunc[code0];
      const int rc = Tb.off_rot[code0];
      const int dc = Tb.off_dist[code0];
      const int Kp = (pt + 1) * (pt + 1), Kz = (pt + 4) * (pt + 4);
      for (int idx = tid; idx < g * (pt + 1); idx += bs) {
        const int jj = idx / (pt + 1), am = idx - jj * (pt + 1);
        const double* ph = Tb.off_phase + (size_t)t_code[s0 + jj] * 2 * P1;
        s_ph[2 * (jj * P1 + am)] = (T)ph[2 * am];
        s_ph[2 * (jj * P1 + am) + 1] = (T)ph[2 * am + 1];
      }
      // (a) members times e^{i m phi}; degrees pt+1..pt+3 set to zero
      for (int jj = 0; jj < g; jj++) {
        const T* src = mp + (size_t)t_box[s0 + jj] * K2;
        const double* ph = Tb.off_phase + (size_t)t_code[s0 + jj] * 2 * P1;
        T* Xj = X + jj * MS;
        for (int k = tid; k < Kz; k += bs) {
          T vr = 0, vi = 0;
          if (k < Kp) {
            const int n = s_kn[k];
            const int m = k - n * n - n, am = m < 0 ? -m : m;
            const T er = (T)ph[2 * am], ei = m < 0 ? -(T)ph[2 * am + 1] : (T)ph[2 * am + 1];
            const T xr = src[2 * k], xi = src[2 * k + 1];
            vr = xr * er - xi * ei;
            vi = xr * ei + xi * er;
          }
          Xj[2 * k] = vr;
          Xj[2 * k + 1] = vi;
        }
      }
      __syncthreads();
      // (b) rotate to the z axis: degree pt - t on wave mf_owner(t)
      const long* ro = Tb.mf_rot_off + (size_t)rc * P1;
      for (int t = 0; t <= pt; t++)
        if (mf_owner(t, NW) == wave)
          mf_rot_degree<T, MAXRT, false>(rtab + ro[pt - t], pt - t, X, MS, g, s_ph, P1, ih, acc);
      __syncthreads();
      // (c) shift along z: order t on wave mf_owner(t)
      const T* zt = ztab + (size_t)dc * Tb.mf_z_size;
      for (int t = 0; t <= pt; t++)
        if (mf_owner(t, NW) == wave) mf_z_order<T, MAXRT>(zt + Tb.mf_z_off[t], p, pt, t, X, MS, g);
      __syncthreads();
      // (d) rotate back, phase back and scale, sum the members into acc
      const T* rb = rtab + Tb.mf_rot_back;
      for (int t = 0; t <= pt; t++)
        if (mf_owner(t, NW) == wave)
          mf_rot_degree<T, MAXRT, true>(rb + ro[pt - t], pt - t, X, MS, g, s_ph, P1, ih, acc);
      __syncthreads();
      s0 += g;
    }
  }
  __syncthreads();
  T* out = lo + (size_t)b * K2;
  for (int k = tid; k < K2; k += bs) out[k] += acc[k];
}

}  // namespace fmm3d
