// Internal interfaces between the C ABI layer and the FMM drivers.
#pragma once
#include "runtime.hpp"

namespace fmm3d {

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
};

// T: far-field arithmetic and expansion storage; TN: near-field arithmetic
template <typename T, typename TN>
int lap_solve(Context& ctx, const LapCall& c);

// precision mode bitmask (FMM3D_PREC_* in include/fmm3d_mi355x.h)
int lap_solve_mode(Context& ctx, const LapCall& c, int mode);
int lap_direct(Context& ctx, const LapCall& c);

// Helmholtz (hfmm3dparttarg / h3dpartdirect); mode 0 = FP64, 1 = FP32 near field
int helm_solve_mode(Context& ctx, const LapCall& c, int mode);
int helm_direct(Context& ctx, const LapCall& c);

}  // namespace fmm3d
