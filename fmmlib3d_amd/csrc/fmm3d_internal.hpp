// Internal interfaces between the C ABI layer and the FMM drivers.
#pragma once
#include "runtime.hpp"

namespace fmm3d {

struct Octree;

// One Laplace evaluation.  *_host pointers are host copies of the
// coordinates (the octree is built on the host); *_dev pointers are device
// buffers in the reference layouts (complex as interleaved doubles).
struct LapCall {
  int iprec = 1, ns = 0, nt = 0;
  int ifcharge = 0, ifdipole = 0, ifpot = 0, iffld = 0, ifpottarg = 0, iffldtarg = 0;
  const double *src_host = nullptr, *trg_host = nullptr;
  const double *src_dev = nullptr, *q_dev = nullptr, *ds_dev = nullptr, *dv_dev = nullptr,
               *trg_dev = nullptr;
  double *pot_dev = nullptr, *fld_dev = nullptr, *pottarg_dev = nullptr, *fldtarg_dev = nullptr;
  double zk_re = 0.0, zk_im = 0.0;  // Helmholtz wavenumber (unused for Laplace)
  // Laplace only: a finished tree with its lists (a forest of several
  // requests, merge_forest) used instead of building one from the points;
  // and whether the launches from the gather to the scatter are captured in
  // a HIP graph and launched as one (no stage timings then)
  const Octree* tree_in = nullptr;
  bool graph = false;
  // with tree_in: whether its lists are still to be built (a tree built for
  // this same call, handed on to another precision), and the device
  // permutations of a tree built on the device (else its host ones are used)
  bool tree_lists_todo = false;
  const int *sperm_in = nullptr, *tperm_in = nullptr;
  // Laplace, FP64 only: Hessians (pot_xx, pot_yy, pot_zz, pot_xy, pot_xz,
  // pot_yz; 6 complex per point) at the sources and at the targets
  int ifhess = 0, ifhesstarg = 0;
  double *hess_dev = nullptr, *hesstarg_dev = nullptr;
  // Laplace triangle FMM (FP64): the triangles' vertices (9 doubles each, in
  // the caller's order); src_dev then holds the evaluation points on them
  // (centroids), q_dev / ds_dev the single- and dipole-layer densities
  const double* tri_dev = nullptr;
  // Helmholtz triangle FMM: whether the single- / dipole-layer densities
  // take the Laplace part of the kernel (0: the difference kernel
  // e^{ikr}/r - 1/r, a density flag of 2 in the reference), and the degree
  // of the near-field rule (0: the one of iprec)
  int tri_lapc = 1, tri_lapd = 1, tri_nqtri = 0;
  // triangle FMMs: leave out the near field (the far field alone)
  bool far_only = false;
  // Laplace, FP64 only: quadrupole sources, quadstr (complex) and quadvec
  // (xx, yy, zz, xy, xz, yz) per source, in the caller's order
  int ifquad = 0;
  const double *qs_dev = nullptr, *qv_dev = nullptr;
};

// The precision mode a call runs in: the requested FMM3D_PREC_* mode while
// FP32 can meet the tolerance of iprec, FP64 throughout at iprec >= 3 (a
// tolerance of 0.5e-9 or tighter, src/lfmm3dpart.f:277-285, which FP32
// pair arithmetic and FP32 translations cannot reach).  At iprec 2 the
// Laplace driver also drops FP32 expansions (FMM3D_PREC_FAR32 bit) on a
// tree whose leaves lie on more than one level, once the tree is built.
inline int effective_precision(int iprec, int mode) { return iprec >= 3 ? 0 : mode; }

// the mode the last FMM call of this thread ran in
inline int& last_precision() {
  static thread_local int m = 0;
  return m;
}

// T: far-field arithmetic and expansion storage; TN: near-field arithmetic
template <typename T, typename TN>
int lap_solve(Context& ctx, const LapCall& c);

// precision mode bitmask (FMM3D_PREC_* in include/fmm3d_mi355x.h)
int lap_solve_mode(Context& ctx, const LapCall& c, int mode);
int lap_direct(Context& ctx, const LapCall& c);
// l3dtriadirect on the device (c.tri_dev set)
int tri_direct(Context& ctx, const LapCall& c);

// Helmholtz (hfmm3dparttarg / h3dpartdirect); mode 0 = FP64, 1 = FP32 near field
int helm_solve_mode(Context& ctx, const LapCall& c, int mode);
int helm_direct(Context& ctx, const LapCall& c);
// h3dtriadirect on the device (c.tri_dev set; c.tri_nqtri the rule's degree)
int htri_direct(Context& ctx, const LapCall& c);

}  // namespace fmm3d
