// Shared pieces of the grouped M2L kernels (Laplace k_m2l_grp, Helmholtz
// k_hm2l_grp): a group is up to kM2LGroup list-2 sources of one target box
// with the same translation class; the lanes of a work item hold the members
// and read the same 4x4 table tile, each lane producing four outputs.
#pragma once
#include <hip/hip_runtime.h>

#include "lap_kernels.hpp"

namespace fmm3d {

constexpr int kM2LGroup = 4;   // members per group
constexpr int kM2LChunk = 192; // list-2 entries ordered at a time (list 2 has <= 189)

// Values per member in LDS: (p+4)^2 complex, padded so that the stride in
// 4-byte words is 8 mod 32 and the four members of a work item start in
// different quarters of the 32 LDS banks.
// ints of the per-index degree table of k_m2l_grp, kept a multiple of 4 so
// that what follows stays 16-byte aligned
__host__ __device__ inline int grp_kn_ints(int p) { return ((p + 4) * (p + 4) + 3) / 4 * 4; }

template <typename T>
__host__ __device__ inline int grp_member_stride(int p) {
  const int base = 2 * (p + 4) * (p + 4);
  const int wpe = (int)sizeof(T) / 4;  // words per value
  int ms = base;
  while ((ms * wpe) % 32 != 8) ms++;
  return ms;
}

// Table tiles: 4 rows x 4 columns, read as four 4-vectors (one per row).
template <typename T>
struct __attribute__((aligned(4 * sizeof(T)))) V4 {
  T v[4];
};
template <typename T>
__device__ __forceinline__ void grp_tile(const T* __restrict__ t, V4<T> (&d)[4]) {
#pragma unroll
  for (int u = 0; u < 4; u++) d[u] = reinterpret_cast<const V4<T>*>(t)[u];
}
// r[4c..4c+3] += d_c * (xr, xi, yr, yi), with the sign (-1)^u of row u when
// NEG (the transpose's (-1)^m, up to a factor applied afterwards)
template <bool NEG, typename T>
__device__ __forceinline__ void grp_fma_row(T (&r)[16], const V4<T>& d, T xr, T xi, T yr, T yi) {
  if (NEG) { xr = -xr; xi = -xi; yr = -yr; yi = -yi; }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    r[4 * c] = fma(d.v[c], xr, r[4 * c]);
    r[4 * c + 1] = fma(d.v[c], xi, r[4 * c + 1]);
    r[4 * c + 2] = fma(d.v[c], yr, r[4 * c + 2]);
    r[4 * c + 3] = fma(d.v[c], yi, r[4 * c + 3]);
  }
}

// Rotation rows 4rb..4rb+3 (input order m = row - n); xin: in_n^{-n}
// ascending, yin: in_n^{+n} descending.  Rows past 2n read neighbouring
// finite LDS values against zero table rows.
// The LDS inputs of four table rows, read one batch ahead like the tiles:
// per row u (xr, xi, yr, yi).
template <typename T>
struct Rows {
  T v[16];
};

// rotation rows 4rb..4rb+3 (input order m = row - n); xin: in_n^{-n}
// ascending, yin: in_n^{+n} descending.  Rows past 2n read neighbouring
// finite LDS values against zero table rows.
template <typename T>
__device__ __forceinline__ void grp_rot_rows(Rows<T>& x, int rb, const T* xin, const T* yin) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int row = 4 * rb + u;
    x.v[4 * u] = xin[2 * row];
    x.v[4 * u + 1] = xin[2 * row + 1];
    x.v[4 * u + 2] = yin[-2 * row];
    x.v[4 * u + 3] = yin[-2 * row + 1];
  }
}

template <typename T, bool BACK>
__device__ __forceinline__ void grp_rows_fma(T (&r)[16], const V4<T> (&d)[4], const Rows<T>& x) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    if (BACK && (u & 1))
      grp_fma_row<true>(r, d[u], x.v[4 * u], x.v[4 * u + 1], x.v[4 * u + 2], x.v[4 * u + 3]);
    else
      grp_fma_row<false>(r, d[u], x.v[4 * u], x.v[4 * u + 1], x.v[4 * u + 2], x.v[4 * u + 3]);
  }
}

// out_{m'} = sum_m D[m][m'] in_m, out_{-m'} = sum_m D[m][m'] in_{-m} for the
// four columns m' = 4q..4q+3 of degree n (tile column q at Db); BACK applies
// the transpose, D^T[m][m'] = (-1)^{m+m'} D[m'][m].  Two register buffers: the
// next tile and its input rows are in flight while the current ones are used.
template <typename T, bool BACK>
__device__ __forceinline__ void grp_rot_item(const T* __restrict__ Db, int nq, int n, int q, const T* in,
                                             T (&r)[16]) {
  const T* xin = in + 2 * (n * n);
  const T* yin = in + 2 * (n * n + 2 * n);
  const int nrb = (2 * n) / 4 + 1;
  const size_t stride = (size_t)nq * 16;
  V4<T> d0[4], d1[4];
  Rows<T> x0, x1;
  grp_tile(Db, d0);
  grp_rot_rows(x0, 0, xin, yin);
  for (int rb = 0;;) {
    const int nx = min(rb + 1, nrb - 1);
    grp_tile(Db + stride * (size_t)nx, d1);
    grp_rot_rows(x1, nx, xin, yin);
    __builtin_amdgcn_sched_barrier(0);
    grp_rows_fma<T, BACK>(r, d0, x0);
    if (++rb >= nrb) break;
    const int ny = min(rb + 1, nrb - 1);
    grp_tile(Db + stride * (size_t)ny, d0);
    grp_rot_rows(x0, ny, xin, yin);
    __builtin_amdgcn_sched_barrier(0);
    grp_rows_fma<T, BACK>(r, d1, x1);
    if (++rb >= nrb) break;
  }
  if (BACK) {
    // the row sign's remaining factor (-1)^n and the column sign (-1)^{m'}
#pragma unroll
    for (int c = 0; c < 4; c++)
      if ((n + 4 * q + c) & 1) { r[4 * c] = -r[4 * c]; r[4 * c + 1] = -r[4 * c + 1]; r[4 * c + 2] = -r[4 * c + 2]; r[4 * c + 3] = -r[4 * c + 3]; }
  }
}

template <typename T>
__device__ __forceinline__ void grp_rot_store(int n, int mp0, const T (&r)[16], T* out) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int mp = mp0 + c;
    if (mp > n) break;
    out[2 * eidx(n, mp)] = r[4 * c];
    out[2 * eidx(n, mp) + 1] = r[4 * c + 1];
    if (mp > 0) {
      out[2 * eidx(n, -mp)] = r[4 * c + 2];
      out[2 * eidx(n, -mp) + 1] = r[4 * c + 3];
    }
  }
}

// z-shift rows n = n0..n0+3 of order am: (in_n^{am}, in_n^{-am})
template <typename T>
__device__ __forceinline__ void grp_z_rows(Rows<T>& x, int n0, int am, const T* in) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int n = n0 + u;
    const T* xp = in + 2 * (n * n + n + am);
    x.v[4 * u] = xp[0];
    x.v[4 * u + 1] = xp[1];
    x.v[4 * u + 2] = xp[-4 * am];
    x.v[4 * u + 3] = xp[-4 * am + 1];
  }
}

// out_l^{+-am} = sum_{n=am}^{pt} Z[am][l][n] in_n^{+-am}, l = am+4q..am+4q+3;
// Zm: the tiles of order am (row blocks of n - am, tile column q).  Degrees
// pt+1..pt+3 of the input are zero, so the last row block needs no mask.
template <typename T>
__device__ __forceinline__ void grp_z_item(const T* __restrict__ Zm, int nq, int am, int q, int pt, const T* in,
                                           T (&r)[16]) {
  const T* Zq = Zm + 16 * q;
  const int nrb = (pt - am) / 4 + 1;
  const size_t stride = (size_t)nq * 16;
  V4<T> z0[4], z1[4];
  Rows<T> x0, x1;
  grp_tile(Zq, z0);
  grp_z_rows(x0, am, am, in);
  for (int rb = 0;;) {
    const int nx = min(rb + 1, nrb - 1);
    grp_tile(Zq + stride * (size_t)nx, z1);
    grp_z_rows(x1, am + 4 * nx, am, in);
    __builtin_amdgcn_sched_barrier(0);
    grp_rows_fma<T, false>(r, z0, x0);
    if (++rb >= nrb) break;
    const int ny = min(rb + 1, nrb - 1);
    grp_tile(Zq + stride * (size_t)ny, z0);
    grp_z_rows(x0, am + 4 * ny, am, in);
    __builtin_amdgcn_sched_barrier(0);
    grp_rows_fma<T, false>(r, z1, x1);
    if (++rb >= nrb) break;
  }
}

template <typename T>
__device__ __forceinline__ void grp_z_store(int am, int l0, int pt, const T (&r)[16], T* out) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int l = l0 + c;
    if (l > pt) break;
    out[2 * eidx(l, am)] = r[4 * c];
    out[2 * eidx(l, am) + 1] = r[4 * c + 1];
    if (am > 0) {
      out[2 * eidx(l, -am)] = r[4 * c + 2];
      out[2 * eidx(l, -am) + 1] = r[4 * c + 3];
    }
  }
}


// Complex tables (Helmholtz coaxial shifts): a tile is 16 real parts then
// 16 imaginary parts, both row-major [row][column].
// r[4c..4c+3] += (dr_c + i di_c) * ((xr + i xi), (yr + i yi))
template <typename T>
__device__ __forceinline__ void grp_fma_row_c(T (&r)[16], const V4<T>& dr, const V4<T>& di, T xr, T xi, T yr,
                                              T yi) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    r[4 * c] = fma(dr.v[c], xr, fma(-di.v[c], xi, r[4 * c]));
    r[4 * c + 1] = fma(dr.v[c], xi, fma(di.v[c], xr, r[4 * c + 1]));
    r[4 * c + 2] = fma(dr.v[c], yr, fma(-di.v[c], yi, r[4 * c + 2]));
    r[4 * c + 3] = fma(dr.v[c], yi, fma(di.v[c], yr, r[4 * c + 3]));
  }
}

template <typename T>
__device__ __forceinline__ void grp_tile_c(const T* __restrict__ t, V4<T> (&d)[8]) {
#pragma unroll
  for (int u = 0; u < 8; u++) d[u] = reinterpret_cast<const V4<T>*>(t)[u];
}

// out_l^{+-am} = sum_{n=am}^{pt} Z[am][l][n] in_n^{+-am} with complex Z,
// l = am+4q..am+4q+3; Zm: the complex tiles of order am, nq tile columns.
// One tile at a time (a complex tile is twice a real one in registers).
template <typename T>
__device__ __forceinline__ void grp_z_item_c(const T* __restrict__ Zm, int nq, int am, int q, int pt, const T* in,
                                             T (&r)[16]) {
  const T* Zq = Zm + 32 * q;
  const int nrb = (pt - am) / 4 + 1;
  const size_t stride = (size_t)nq * 32;
  for (int rb = 0; rb < nrb; rb++) {
    V4<T> z[8];
    Rows<T> x;
    grp_tile_c(Zq + stride * (size_t)rb, z);
    grp_z_rows(x, am + 4 * rb, am, in);
#pragma unroll
    for (int u = 0; u < 4; u++) grp_fma_row_c(r, z[u], z[4 + u], x.v[4 * u], x.v[4 * u + 1], x.v[4 * u + 2], x.v[4 * u + 3]);
  }
}

}  // namespace fmm3d
