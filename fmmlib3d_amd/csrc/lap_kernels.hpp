// Device-side data layout shared by the Laplace kernels and the host plan.
#pragma once
#include <hip/hip_runtime.h>

namespace fmm3d {

// Index of expansion coefficient (n,m), -n<=m<=n, in a packed expansion.
__host__ __device__ inline int eidx(int n, int m) { return n * n + n + m; }
// Index of (n,m), 0<=m<=n, in a half table.
__host__ __device__ inline int hidx(int n, int m) { return n * (n + 1) / 2 + m; }

// Tree + schedule as flat device arrays (all ints live in one buffer, all
// doubles in another; the pointers below point into them).
struct DevPlan {
  int nboxes, nlev, p, K;  // K = (p+1)^2
  int ns, nt;
  // boxes
  const int *level, *isrc, *nsrc, *itrg, *ntrg, *nchild, *child, *ijk;
  const double *center;   // 3 per box
  const double *inv_side; // 1/side per level (per scale index of a forest, see merge_forest)
  // lists (CSR over boxes)
  const int *l1_off, *l1_idx, *l2_off, *l2_idx, *l3_off, *l3_idx, *l4_off, *l4_idx;
  // work lists
  const int *p2m_boxes;  int n_p2m;
  const int *l2p_boxes;  int n_l2p;
  const int *l2p_items;  int n_l2pi;  // (position in l2p_boxes, kind, start) triples
  const int *m2p_boxes;  int n_m2p;   // leaves with list 3
  const int *p2l_boxes;  int n_p2l;   // boxes with list 4
  const int *m2l_boxes;  int n_m2l;   // targets with (pruned) list 2
  const int *m2m_boxes;  const int *m2m_lev_off;   // parents grouped by level
  const int *l2l_boxes;  const int *l2l_lev_off;
  const int *p2p_items;  int n_p2p;   // (leaf, kind, start) triples
  // permutations
  const int *src_perm, *trg_perm;
};

// Translation tables for one expansion order p (level independent because
// expansions carry rscale = 1/side).
struct DevTables {
  int p;
  // offset code (i+3)*49+(j+3)*7+(k+3) for source-minus-target in [-3,3]^3
  const int *off_rot;     // rotation class per offset code
  const int *off_dist;    // distance class per offset code
  const int *off_ptrunc;  // list-2 truncation order per offset code
  const double *off_phase;  // e^{i m phi}, m=0..p, per offset code (re,im)
  const double *rot;        // rotation blocks per class
  const long *rot_off;      // start of class c, degree n block: rot_off[c*(p+1)+n]
  const double *z_m2l;      // per distance class: [m][l][n] dense (p+1)^3
  const double *z_m2m;      // [m][n][l] dense (p+1)^3 : parent n from child l
  const double *z_l2l;      // [m][l][n] dense (p+1)^3 : child l from parent n
  const double *oct_phase;  // 8 octants x 2 (m2m, l2l) x (p+1) complex
  const int *oct_rot;       // 8 octants x 2 -> rotation class (m2m, l2l)
  const double *fac;        // derivative factors, see lap_kernels.hip
  const double *delta;      // d^n(pi/2) per degree, row-major (2n+1)^2
  const long *delta_off;    // start of degree n
  const double *rot_e;      // per rotation class: e^{i k theta}, k = 0..p
  // grouped M2L work items: rotation (n << 8 | q) by degree, count with n <= pt;
  // z-shift (|m| << 8 | q) per truncation order pt, start of pt's items
  const int *grp_ritems, *grp_rcount, *grp_zitems, *grp_zoff;
  // tiled tables of the grouped M2L: half rotation blocks [m][m' >= 0] per
  // (class, degree) at rot_half_off[c*(p+1)+n]; z-shift [n-m][l-m] per
  // distance class (zt_size values each), order m at zt_off[m]; both in 4x4
  // tiles [row block][column block][row][column], in FP64 and in FP32
  const double *rot_half, *z_m2l_t;
  const float *rot_half_f, *z_m2l_t_f;
  const long *rot_half_off, *zt_off;
  long zt_size;
  // MFMA-ordered M2L tables (FP64 and FP32): forward rotation blocks of
  // (class, degree) at mf_rot_off[c*(p+1)+n], back blocks mf_rot_back
  // further; z-shift blocks of order m at mf_z_off[m] within a distance
  // class of mf_z_size values
  const double *mf_rot, *mf_z;
  const float *mf_rot_f, *mf_z_f;
  const long *mf_rot_off, *mf_z_off;
  long mf_rot_back, mf_z_size;
};

}  // namespace fmm3d
