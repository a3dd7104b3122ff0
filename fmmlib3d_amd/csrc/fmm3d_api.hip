// C ABI of libfmm3d_mi355x.so (declared in include/fmm3d_mi355x.h).
// Argument checking and ier mapping follow lfmm3dparttarg
// (src/lfmm3dpart.f:138-522); host arrays are uploaded once, the solve runs
// on the device and the results come back in the caller's order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/fmm3d_mi355x.h"
#include "fmm3d_internal.hpp"
#include "tree_gpu.hpp"
#include "tria_rules.hpp"
#include "octree.hpp"

using namespace fmm3d;

namespace {

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const AllocError& e) {
    return e.ier;
  } catch (const HipError&) {
    (void)hipGetLastError();
    return 64;
  } catch (const std::bad_alloc&) {
    return 8;
  }
}

// Host arrays of a few MB and more go through the context's pinned staging
// buffers (host copy and DMA overlapped); smaller ones are copied directly.
constexpr size_t kStagedMin = (size_t)4 << 20;

double* upload(Context& ctx, const char* name, const double* h, size_t n) {
  if (!h || n == 0) return nullptr;
  double* d = (double*)ctx.arena.get(name, n * sizeof(double));
  if (n * sizeof(double) >= kStagedMin && !ctx.arena.fixed())
    staged_upload(ctx, d, h, n * sizeof(double));
  else
    FMM_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
  return d;
}

void download(Context& ctx, double* h, const double* d, size_t n) {
  if (!h || !d || n == 0) return;
  if (n * sizeof(double) >= kStagedMin && !ctx.arena.fixed())
    staged_download(ctx, h, d, n * sizeof(double));
  else
    FMM_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
}

// kernel: nullptr = Laplace, else the Helmholtz wavenumber (re, im)
int run_solve(Context& ctx, const double* zk, const LapCall& c, int prec_mode) {
  return zk ? helm_solve_mode(ctx, c, prec_mode) : lap_solve_mode(ctx, c, prec_mode);
}

int check_args(const double* zk, int prec_mode, int iprec, int ns, int nt) {
  if (nt < 0) return 128;
  if (ns < 1) return 4;  // lfmm3dpart.f:309-330 / hfmm3dpart.f:307-330 after d3tstrcr's ier=128
  if (iprec < -2 || iprec > 6) return 32;
  if (zk) {
    if (zk[0] == 0.0 && zk[1] == 0.0) return 128;  // zk = 0: the Helmholtz scaling degenerates
    if (prec_mode != 0 && prec_mode != 1) return 128;
  } else if (prec_mode < 0 || prec_mode > 3) {
    return 128;
  }
  return 0;
}

// 128 when an array the flags ask for is NULL (the reference would fault;
// on the device path a NULL array would fault the GPU): source with ns > 0,
// charge with ifcharge = 1, dipstr/dipvec with ifdipole = 1, target with
// nt > 0, and each output whose flag is 1.
int check_ptrs(int ns, const double* source, int ifcharge, const double* charge, int ifdipole,
               const double* dipstr, const double* dipvec, int ifpot, const double* pot, int iffld,
               const double* fld, int nt, const double* target, int ifpottarg, const double* pottarg,
               int iffldtarg, const double* fldtarg) {
  if (ns > 0 && !source) return 128;
  if (ifcharge == 1 && ns > 0 && !charge) return 128;
  if (ifdipole == 1 && ns > 0 && (!dipstr || !dipvec)) return 128;
  if (ifpot == 1 && ns > 0 && !pot) return 128;
  if (iffld == 1 && ns > 0 && !fld) return 128;
  if (nt > 0) {
    if (!target) return 128;
    if (ifpottarg == 1 && !pottarg) return 128;
    if (iffldtarg == 1 && !fldtarg) return 128;
  }
  return 0;
}

int fmm_host(const double* zk, int prec_mode, int iprec, int ns, const double* source, int ifcharge,
             const double* charge, int ifdipole, const double* dipstr, const double* dipvec,
             int ifpot, double* pot, int iffld, double* fld, int nt, const double* target,
             int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg, int ifhess = 0,
             double* hess = nullptr, int ifhesstarg = 0, double* hesstarg = nullptr,
             const double* tria = nullptr, bool tri_mode = false, int ifquad = 0,
             const double* quadstr = nullptr, const double* quadvec = nullptr) {
  if (int e = check_args(zk, prec_mode, iprec, ns, nt)) return e;
  if (tri_mode && !tria) return 128;
  if (int e = check_ptrs(ns, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, nt,
                         target, ifpottarg, pottarg, iffldtarg, fldtarg))
    return e;
  if ((ifhess == 1 && !hess) || (nt > 0 && ifhesstarg == 1 && !hesstarg)) return 128;
  if (ifquad == 1 && (!quadstr || !quadvec)) return 128;
  Context& ctx = default_context();
  std::lock_guard<std::mutex> lock(ctx.mu);
  return guarded([&] {
    LapCall c;
    c.iprec = iprec; c.ns = ns; c.nt = nt;
    if (zk) { c.zk_re = zk[0]; c.zk_im = zk[1]; }
    c.ifcharge = ifcharge == 1; c.ifdipole = ifdipole == 1;
    c.ifpot = ifpot == 1; c.iffld = iffld == 1;
    c.ifpottarg = nt > 0 && ifpottarg == 1; c.iffldtarg = nt > 0 && iffldtarg == 1;
    c.src_host = source;
    c.trg_host = nt > 0 ? target : nullptr;
    c.src_dev = upload(ctx, "in_src", source, 3 * (size_t)ns);
    c.q_dev = c.ifcharge ? upload(ctx, "in_q", charge, 2 * (size_t)ns) : nullptr;
    c.ds_dev = c.ifdipole ? upload(ctx, "in_ds", dipstr, 2 * (size_t)ns) : nullptr;
    // (triangles: the dipole layer is along each triangle's normal, dipvec is not read)
    c.dv_dev = c.ifdipole && !tri_mode ? upload(ctx, "in_dv", dipvec, 3 * (size_t)ns) : nullptr;
    c.trg_dev = nt > 0 ? upload(ctx, "in_trg", target, 3 * (size_t)nt) : nullptr;
    if (tri_mode) c.tri_dev = upload(ctx, "in_tri", tria, 9 * (size_t)ns);
    c.ifquad = ifquad == 1;
    if (c.ifquad) {
      c.qs_dev = upload(ctx, "in_qs", quadstr, 2 * (size_t)ns);
      c.qv_dev = upload(ctx, "in_qv", quadvec, 6 * (size_t)ns);
    }
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("out_pot", 2 * sizeof(double) * ns);
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("out_fld", 6 * sizeof(double) * ns);
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("out_pt", 2 * sizeof(double) * nt);
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("out_ft", 6 * sizeof(double) * nt);
    c.ifhess = ifhess == 1;
    c.ifhesstarg = nt > 0 && ifhesstarg == 1;
    if (c.ifhess) c.hess_dev = (double*)ctx.arena.get("out_hess", 12 * sizeof(double) * ns);
    if (c.ifhesstarg) c.hesstarg_dev = (double*)ctx.arena.get("out_ht", 12 * sizeof(double) * nt);
    int ier = run_solve(ctx, zk, c, prec_mode);
    if (ier) return ier;
    if (c.ifpot) download(ctx, pot, c.pot_dev, 2 * (size_t)ns);
    if (c.iffld) download(ctx, fld, c.fld_dev, 6 * (size_t)ns);
    if (c.ifpottarg) download(ctx, pottarg, c.pottarg_dev, 2 * (size_t)nt);
    if (c.iffldtarg) download(ctx, fldtarg, c.fldtarg_dev, 6 * (size_t)nt);
    if (c.ifhess) download(ctx, hess, c.hess_dev, 12 * (size_t)ns);
    if (c.ifhesstarg) download(ctx, hesstarg, c.hesstarg_dev, 12 * (size_t)nt);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  });
}

}  // namespace

extern "C" {

int fmm3d_lfmm3dparttarg(int iprec, int nsource, const double* source, int ifcharge,
                         const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                         int ntarget, const double* target, int ifpottarg, double* pottarg,
                         int iffldtarg, double* fldtarg) {
  return fmm_host(nullptr, FMM3D_PREC_FP64, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr,
                  dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg,
                  fldtarg);
}

int fmm3d_lfmm3dparttarg_hess(int iprec, int nsource, const double* source, int ifcharge,
                              const double* charge, int ifdipole, const double* dipstr,
                              const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                              int ifhess, double* hess, int ntarget, const double* target,
                              int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg,
                              int ifhesstarg, double* hesstarg) {
  return fmm_host(nullptr, 0, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot,
                  iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg, ifhess, hess, ifhesstarg,
                  hesstarg);
}

int fmm3d_lfmm3dparttarg_qp(int iprec, int nsource, const double* source, int ifcharge,
                            const double* charge, int ifdipole, const double* dipstr,
                            const double* dipvec, int ifquad, const double* quadstr, const double* quadvec,
                            int ifpot, double* pot, int iffld, double* fld, int ntarget, const double* target,
                            int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  return fmm_host(nullptr, 0, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot,
                  iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg, 0, nullptr, 0, nullptr,
                  nullptr, false, ifquad, quadstr, quadvec);
}

int fmm3d_lfmm3dparttarg_ex(int prec_mode, int iprec, int nsource, const double* source,
                            int ifcharge, const double* charge, int ifdipole, const double* dipstr,
                            const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                            int ntarget, const double* target, int ifpottarg, double* pottarg,
                            int iffldtarg, double* fldtarg) {
  return fmm_host(nullptr, prec_mode, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec,
                  ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

int fmm3d_lfmm3dpartself(int iprec, int nsource, const double* source, int ifcharge,
                         const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld) {
  return fmm_host(nullptr, FMM3D_PREC_FP64, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr,
                  dipvec, ifpot, pot, iffld, fld, 0, nullptr, 0, nullptr, 0, nullptr);
}

}  // extern "C"

namespace {

// The device-resident call on a given context; the caller holds ctx.mu.
int fmm_dev_on(Context& ctx, const double* zk, int prec_mode, int iprec, int ns, const double* source_dev,
               int ifcharge, const double* charge_dev, int ifdipole, const double* dipstr_dev,
               const double* dipvec_dev, int ifpot, double* pot_dev, int iffld, double* fld_dev, int nt,
               const double* target_dev, int ifpottarg, double* pottarg_dev, int iffldtarg,
               double* fldtarg_dev, void* stream) {
  hipStream_t saved = ctx.stream;
  ctx.stream = (hipStream_t)stream;
  int ier = guarded([&] {
    // the octree is built on the device from the device coordinates; the
    // host builder (FMM3D_TREE=host) needs them back in pinned buffers
    double* hs = nullptr;
    double* ht = nullptr;
    if (!device_tree_enabled()) {
      hs = (double*)ctx.arena.get_host("coord_src", 3 * sizeof(double) * (size_t)ns);
      ht = (double*)ctx.arena.get_host("coord_trg", 3 * sizeof(double) * (size_t)std::max(nt, 1));
      FMM_HIP(hipMemcpyAsync(hs, source_dev, 3 * sizeof(double) * (size_t)ns, hipMemcpyDeviceToHost, ctx.stream));
      if (nt > 0)
        FMM_HIP(hipMemcpyAsync(ht, target_dev, 3 * sizeof(double) * (size_t)nt, hipMemcpyDeviceToHost, ctx.stream));
      FMM_HIP(hipStreamSynchronize(ctx.stream));
    }
    LapCall c;
    c.iprec = iprec; c.ns = ns; c.nt = nt;
    c.ifcharge = ifcharge == 1; c.ifdipole = ifdipole == 1;
    c.ifpot = ifpot == 1; c.iffld = iffld == 1;
    c.ifpottarg = nt > 0 && ifpottarg == 1; c.iffldtarg = nt > 0 && iffldtarg == 1;
    c.src_host = hs; c.trg_host = nt > 0 ? ht : nullptr;  // null with the device builder
    c.src_dev = source_dev; c.q_dev = charge_dev; c.ds_dev = dipstr_dev; c.dv_dev = dipvec_dev;
    c.trg_dev = target_dev;
    c.pot_dev = pot_dev; c.fld_dev = fld_dev; c.pottarg_dev = pottarg_dev; c.fldtarg_dev = fldtarg_dev;
    if (zk) { c.zk_re = zk[0]; c.zk_im = zk[1]; }
    return run_solve(ctx, zk, c, prec_mode);
  });
  ctx.stream = saved;
  return ier;
}

int fmm_dev(const double* zk, int prec_mode, int iprec, int ns, const double* source_dev,
            int ifcharge, const double* charge_dev, int ifdipole, const double* dipstr_dev,
            const double* dipvec_dev, int ifpot, double* pot_dev, int iffld, double* fld_dev, int nt,
            const double* target_dev, int ifpottarg, double* pottarg_dev, int iffldtarg,
            double* fldtarg_dev, void* stream) {
  if (int e = check_args(zk, prec_mode, iprec, ns, nt)) return e;
  if (int e = check_ptrs(ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev, dipvec_dev, ifpot, pot_dev,
                         iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev, iffldtarg, fldtarg_dev))
    return e;
  Context& ctx = default_context();
  std::lock_guard<std::mutex> lock(ctx.mu);
  return fmm_dev_on(ctx, zk, prec_mode, iprec, ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev,
                    dipvec_dev, ifpot, pot_dev, iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev,
                    iffldtarg, fldtarg_dev, stream);
}

// ---- allocation-free variant: every device buffer of the call is a slice
// of `scratch`, a file-scope device array of FMM3D_NOALLOC_SCRATCH_BYTES.
// Host code never uses the array's name as an address: the device address
// comes from hipGetSymbolAddress, once per device with that device current,
// and is kept per device in NoAllocDevice::ptr.
__device__ __align__(256) unsigned char scratch[FMM3D_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

struct NoAllocDevice {
  std::once_flag once;
  hipError_t status = hipErrorNotInitialized;
  void* ptr = nullptr;  // this device's copy of `scratch`
  Context ctx;          // arena in fixed mode over ptr; near field on the caller's stream
};

NoAllocDevice& noalloc_device(int dev) {
  static NoAllocDevice devs[kMaxDevices];
  NoAllocDevice& d = devs[dev];
  std::call_once(d.once, [&] {
    d.status = hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch));
    if (d.status != hipSuccess) {
      (void)hipGetLastError();
      d.ptr = nullptr;
      return;
    }
    d.ctx.device = dev;
    d.ctx.near_on_aux = false;
    d.ctx.arena.use_fixed(d.ptr, FMM3D_NOALLOC_SCRATCH_BYTES);
  });
  return d;
}

// Bytes per point that every call takes from the region whatever its tree
// (sorted positions and strengths, far and near sums, octree keys): a
// lower bound, checked before any device work; the exact bound is the
// region itself (AllocError kIerScratch from the arena).
constexpr size_t kNoAllocBytesPerPoint = 256;

}  // namespace

extern "C" {

int fmm3d_lfmm3dparttarg_dev(int prec_mode, int iprec, int ns, const double* source_dev,
                             int ifcharge, const double* charge_dev, int ifdipole,
                             const double* dipstr_dev, const double* dipvec_dev, int ifpot,
                             double* pot_dev, int iffld, double* fld_dev, int nt,
                             const double* target_dev, int ifpottarg, double* pottarg_dev,
                             int iffldtarg, double* fldtarg_dev, void* stream) {
  return fmm_dev(nullptr, prec_mode, iprec, ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev,
                 dipvec_dev, ifpot, pot_dev, iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev,
                 iffldtarg, fldtarg_dev, stream);
}

int fmm3d_lfmm3dparttarg_dev_noalloc(int prec_mode, int iprec, int ns, const double* source_dev,
                                     int ifcharge, const double* charge_dev, int ifdipole,
                                     const double* dipstr_dev, const double* dipvec_dev, int ifpot,
                                     double* pot_dev, int iffld, double* fld_dev, int nt,
                                     const double* target_dev, int ifpottarg, double* pottarg_dev,
                                     int iffldtarg, double* fldtarg_dev, void* stream) {
  if (int e = check_args(nullptr, prec_mode, iprec, ns, nt)) return e;
  if (int e = check_ptrs(ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev, dipvec_dev, ifpot, pot_dev,
                         iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev, iffldtarg, fldtarg_dev))
    return e;
  if (((size_t)ns + (size_t)nt) * kNoAllocBytesPerPoint > FMM3D_NOALLOC_SCRATCH_BYTES) return kIerScratch;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return 64;
  }
  NoAllocDevice& d = noalloc_device(dev);
  if (d.status != hipSuccess) return 64;
  std::lock_guard<std::mutex> lock(d.ctx.mu);
  // the region keeps the previous call's contents: forget its slices and
  // the tables it held (they are uploaded again into this call's slices)
  d.ctx.arena.reset_fixed();
  d.ctx.lap_tabs.clear();
  return fmm_dev_on(d.ctx, nullptr, prec_mode, iprec, ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev,
                    dipvec_dev, ifpot, pot_dev, iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev,
                    iffldtarg, fldtarg_dev, stream);
}

int fmm3d_hfmm3dparttarg_dev(int prec_mode, int iprec, const double* zk, int ns, const double* source_dev,
                             int ifcharge, const double* charge_dev, int ifdipole,
                             const double* dipstr_dev, const double* dipvec_dev, int ifpot,
                             double* pot_dev, int iffld, double* fld_dev, int nt,
                             const double* target_dev, int ifpottarg, double* pottarg_dev,
                             int iffldtarg, double* fldtarg_dev, void* stream) {
  if (!zk) return 128;
  return fmm_dev(zk, prec_mode, iprec, ns, source_dev, ifcharge, charge_dev, ifdipole, dipstr_dev,
                 dipvec_dev, ifpot, pot_dev, iffld, fld_dev, nt, target_dev, ifpottarg, pottarg_dev,
                 iffldtarg, fldtarg_dev, stream);
}

int fmm3d_hfmm3dparttarg(int iprec, const double* zk, int nsource, const double* source, int ifcharge,
                         const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                         int ntarget, const double* target, int ifpottarg, double* pottarg,
                         int iffldtarg, double* fldtarg) {
  if (!zk) return 128;
  return fmm_host(zk, FMM3D_PREC_FP64, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec,
                  ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

int fmm3d_hfmm3dparttarg_ex(int prec_mode, int iprec, const double* zk, int nsource, const double* source,
                            int ifcharge, const double* charge, int ifdipole, const double* dipstr,
                            const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                            int ntarget, const double* target, int ifpottarg, double* pottarg,
                            int iffldtarg, double* fldtarg) {
  if (!zk) return 128;
  return fmm_host(zk, prec_mode, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec,
                  ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

int fmm3d_hfmm3dpartself(int iprec, const double* zk, int nsource, const double* source, int ifcharge,
                         const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld) {
  if (!zk) return 128;
  return fmm_host(zk, FMM3D_PREC_FP64, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec,
                  ifpot, pot, iffld, fld, 0, nullptr, 0, nullptr, 0, nullptr);
}

}  // extern "C"

namespace {

void direct_host(const double* zk, int nsource, const double* source, int ifcharge, const double* charge,
                 int ifdipole, const double* dipstr, const double* dipvec, int ifpot, double* pot,
                 int iffld, double* fld, int ntarget, const double* target, int ifpottarg,
                 double* pottarg, int iffldtarg, double* fldtarg, const double* tria = nullptr,
                 bool tri_mode = false) {
  if (nsource < 0 || ntarget < 0) return;
  if (tri_mode && nsource > 0 && !tria) return;
  if (check_ptrs(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget,
                 target, ifpottarg, pottarg, iffldtarg, fldtarg))
    return;  // no error channel: a call with a missing array computes nothing
  Context& ctx = default_context();
  std::lock_guard<std::mutex> lock(ctx.mu);
  (void)guarded([&] {
    LapCall c;
    c.ns = nsource; c.nt = ntarget;
    if (zk) { c.zk_re = zk[0]; c.zk_im = zk[1]; }
    c.ifcharge = ifcharge == 1; c.ifdipole = ifdipole == 1;
    c.ifpot = ifpot == 1; c.iffld = iffld == 1;
    c.ifpottarg = ifpottarg == 1; c.iffldtarg = iffldtarg == 1;
    c.src_dev = upload(ctx, "in_src", source, 3 * (size_t)nsource);
    c.q_dev = c.ifcharge ? upload(ctx, "in_q", charge, 2 * (size_t)nsource) : nullptr;
    c.ds_dev = c.ifdipole ? upload(ctx, "in_ds", dipstr, 2 * (size_t)nsource) : nullptr;
    c.dv_dev = c.ifdipole && !tri_mode ? upload(ctx, "in_dv", dipvec, 3 * (size_t)nsource) : nullptr;
    c.trg_dev = ntarget > 0 ? upload(ctx, "in_trg", target, 3 * (size_t)ntarget) : nullptr;
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("out_pot", 2 * sizeof(double) * std::max(nsource, 1));
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("out_fld", 6 * sizeof(double) * std::max(nsource, 1));
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("out_pt", 2 * sizeof(double) * std::max(ntarget, 1));
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("out_ft", 6 * sizeof(double) * std::max(ntarget, 1));
    if (tri_mode) c.tri_dev = upload(ctx, "in_tri", tria, 9 * (size_t)nsource);
    int ier = tri_mode ? tri_direct(ctx, c) : zk ? helm_direct(ctx, c) : lap_direct(ctx, c);
    if (ier) return ier;
    if (c.ifpot) download(ctx, pot, c.pot_dev, 2 * (size_t)nsource);
    if (c.iffld) download(ctx, fld, c.fld_dev, 6 * (size_t)nsource);
    if (c.ifpottarg) download(ctx, pottarg, c.pottarg_dev, 2 * (size_t)ntarget);
    if (c.iffldtarg) download(ctx, fldtarg, c.fldtarg_dev, 6 * (size_t)ntarget);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  });
}

// out -= sub, n doubles
__global__ void k_subtract(size_t n, double* __restrict__ out, const double* __restrict__ sub) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] -= sub[i];
}

// The Helmholtz triangle FMM (hfmm3dtriatarg, src/hfmm3dtria.f:133-647):
// density flags 0, 1, or 2 (that density's Laplace part subtracted: the
// near field with the difference kernel, the far field as the Helmholtz far
// field minus the Laplace one, as the reference's second, far-only Laplace
// pass does, src/hfmm3dtria.f:552-640).
int htri_host(const double* zk, int iprec, int ns, const double* tria, const double* source, int ifcharge,
              const double* charge, int ifdipole, const double* dipstr, int ifpot, double* pot, int iffld,
              double* fld, int nt, const double* target, int ifpottarg, double* pottarg, int iffldtarg,
              double* fldtarg) {
  if (!zk) return 128;
  if (ifcharge < 0 || ifcharge > 2 || ifdipole < 0 || ifdipole > 2) return 128;
  if (int e = check_args(zk, 0, iprec, ns, nt)) return e;
  if (!tria) return 128;
  const int ifc = ifcharge ? 1 : 0, ifd = ifdipole ? 1 : 0;
  // the dipole orientation is the triangle normal: no dipvec is read
  if (int e = check_ptrs(ns, source, ifc, charge, 0, nullptr, nullptr, ifpot, pot, iffld, fld, nt, target,
                         ifpottarg, pottarg, iffldtarg, fldtarg))
    return e;
  if (ifd && !dipstr) return 128;
  Context& ctx = default_context();
  std::lock_guard<std::mutex> lock(ctx.mu);
  return guarded([&] {
    LapCall c;
    c.iprec = iprec; c.ns = ns; c.nt = nt;
    c.zk_re = zk[0]; c.zk_im = zk[1];
    c.ifcharge = ifc; c.ifdipole = ifd;
    c.ifpot = ifpot == 1; c.iffld = iffld == 1;
    c.ifpottarg = nt > 0 && ifpottarg == 1; c.iffldtarg = nt > 0 && iffldtarg == 1;
    c.tri_lapc = ifcharge != 2; c.tri_lapd = ifdipole != 2;
    c.src_host = source;
    c.trg_host = nt > 0 ? target : nullptr;
    c.src_dev = upload(ctx, "in_src", source, 3 * (size_t)ns);
    c.q_dev = ifc ? upload(ctx, "in_q", charge, 2 * (size_t)ns) : nullptr;
    c.ds_dev = ifd ? upload(ctx, "in_ds", dipstr, 2 * (size_t)ns) : nullptr;
    c.trg_dev = nt > 0 ? upload(ctx, "in_trg", target, 3 * (size_t)nt) : nullptr;
    c.tri_dev = upload(ctx, "in_tri", tria, 9 * (size_t)ns);
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("out_pot", 2 * sizeof(double) * ns);
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("out_fld", 6 * sizeof(double) * ns);
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("out_pt", 2 * sizeof(double) * nt);
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("out_ft", 6 * sizeof(double) * nt);
    int ier = helm_solve_mode(ctx, c, 0);
    if (ier) return ier;
    if (ifcharge == 2 || ifdipole == 2) {
      LapCall cl = c;
      cl.zk_re = cl.zk_im = 0.0;
      cl.ifcharge = ifcharge == 2; cl.ifdipole = ifdipole == 2;
      cl.far_only = true;
      if (c.ifpot) cl.pot_dev = (double*)ctx.arena.get("out_pot_lap", 2 * sizeof(double) * ns);
      if (c.iffld) cl.fld_dev = (double*)ctx.arena.get("out_fld_lap", 6 * sizeof(double) * ns);
      if (c.ifpottarg) cl.pottarg_dev = (double*)ctx.arena.get("out_pt_lap", 2 * sizeof(double) * nt);
      if (c.iffldtarg) cl.fldtarg_dev = (double*)ctx.arena.get("out_ft_lap", 6 * sizeof(double) * nt);
      ier = lap_solve_mode(ctx, cl, 0);
      if (ier) return ier;
      auto sub = [&](double* out, const double* l, size_t n) {
        if (out && n)
          hipLaunchKernelGGL(k_subtract, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx.stream, n, out, l);
      };
      sub(c.pot_dev, cl.pot_dev, 2 * (size_t)ns);
      sub(c.fld_dev, cl.fld_dev, 6 * (size_t)ns);
      sub(c.pottarg_dev, cl.pottarg_dev, 2 * (size_t)nt);
      sub(c.fldtarg_dev, cl.fldtarg_dev, 6 * (size_t)nt);
      FMM_HIP(hipGetLastError());
    }
    if (c.ifpot) download(ctx, pot, c.pot_dev, 2 * (size_t)ns);
    if (c.iffld) download(ctx, fld, c.fld_dev, 6 * (size_t)ns);
    if (c.ifpottarg) download(ctx, pottarg, c.pottarg_dev, 2 * (size_t)nt);
    if (c.iffldtarg) download(ctx, fldtarg, c.fldtarg_dev, 6 * (size_t)nt);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  });
}

// h3dtriadirect (src/hfmm3dtria.f:1456): no error channel, as the
// reference; a call with a missing array or an unsupported rule degree
// computes nothing.
void htri_direct_host(int nqtri, const double* zk, int ns, const double* tria, const double* source, int ifcharge,
                      const double* charge, int ifdipole, const double* dipstr, int ifpot, double* pot, int iffld,
                      double* fld, int nt, const double* target, int ifpottarg, double* pottarg, int iffldtarg,
                      double* fldtarg) {
  if (!zk || ns < 0 || nt < 0 || tri_sym_rule(nqtri).empty()) return;
  if (ifcharge < 0 || ifcharge > 2 || ifdipole < 0 || ifdipole > 2) return;
  if (ns > 0 && !tria) return;
  const int ifc = ifcharge ? 1 : 0, ifd = ifdipole ? 1 : 0;
  if (check_ptrs(ns, source, ifc, charge, 0, nullptr, nullptr, ifpot, pot, iffld, fld, nt, target, ifpottarg,
                 pottarg, iffldtarg, fldtarg))
    return;
  if (ifd && ns > 0 && !dipstr) return;
  Context& ctx = default_context();
  std::lock_guard<std::mutex> lock(ctx.mu);
  (void)guarded([&] {
    LapCall c;
    c.ns = ns; c.nt = nt;
    c.zk_re = zk[0]; c.zk_im = zk[1];
    c.ifcharge = ifc; c.ifdipole = ifd;
    c.ifpot = ifpot == 1; c.iffld = iffld == 1;
    c.ifpottarg = ifpottarg == 1; c.iffldtarg = iffldtarg == 1;
    c.tri_lapc = ifcharge != 2; c.tri_lapd = ifdipole != 2;
    c.tri_nqtri = nqtri;
    c.src_dev = upload(ctx, "in_src", source, 3 * (size_t)ns);
    c.q_dev = ifc ? upload(ctx, "in_q", charge, 2 * (size_t)ns) : nullptr;
    c.ds_dev = ifd ? upload(ctx, "in_ds", dipstr, 2 * (size_t)ns) : nullptr;
    c.trg_dev = nt > 0 ? upload(ctx, "in_trg", target, 3 * (size_t)nt) : nullptr;
    c.tri_dev = upload(ctx, "in_tri", tria, 9 * (size_t)ns);
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("out_pot", 2 * sizeof(double) * std::max(ns, 1));
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("out_fld", 6 * sizeof(double) * std::max(ns, 1));
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("out_pt", 2 * sizeof(double) * std::max(nt, 1));
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("out_ft", 6 * sizeof(double) * std::max(nt, 1));
    int ier = htri_direct(ctx, c);
    if (ier) return ier;
    if (c.ifpot) download(ctx, pot, c.pot_dev, 2 * (size_t)ns);
    if (c.iffld) download(ctx, fld, c.fld_dev, 6 * (size_t)ns);
    if (c.ifpottarg) download(ctx, pottarg, c.pottarg_dev, 2 * (size_t)nt);
    if (c.iffldtarg) download(ctx, fldtarg, c.fldtarg_dev, 6 * (size_t)nt);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  });
}

}  // namespace

extern "C" {

void fmm3d_l3dpartdirect(int nsource, const double* source, int ifcharge, const double* charge,
                         int ifdipole, const double* dipstr, const double* dipvec, int ifpot,
                         double* pot, int iffld, double* fld, int ntarget, const double* target,
                         int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  direct_host(nullptr, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld,
              fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

void fmm3d_h3dpartdirect(const double* zk, int nsource, const double* source, int ifcharge,
                         const double* charge, int ifdipole, const double* dipstr, const double* dipvec,
                         int ifpot, double* pot, int iffld, double* fld, int ntarget, const double* target,
                         int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  if (!zk) return;
  direct_host(zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld,
              fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

int fmm3d_lfmm3dtriatarg(int iprec, int nsource, const double* triaflat, const double* trianorm,
                         const double* source, int ifcharge, const double* charge, int ifdipole,
                         const double* dipstr, const double* dipvec, int ifpot, double* pot, int iffld, double* fld,
                         int ntarget, const double* target, int ifpottarg, double* pottarg, int iffldtarg,
                         double* fldtarg) {
  (void)trianorm;  // the dipole layer is along the normal of the vertex order, as in the reference
  return fmm_host(nullptr, FMM3D_PREC_FP64, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec,
                  ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg, 0, nullptr, 0,
                  nullptr, triaflat, true);
}

int fmm3d_lfmm3dtriaself(int iprec, int nsource, const double* triaflat, const double* trianorm,
                         const double* source, int ifcharge, const double* charge, int ifdipole,
                         const double* dipstr, const double* dipvec, int ifpot, double* pot, int iffld,
                         double* fld) {
  return fmm3d_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                              dipvec, ifpot, pot, iffld, fld, 0, nullptr, 0, nullptr, 0, nullptr);
}

int fmm3d_lfmm3dtriampftarg(int iprec, int nsource, const double* triaflat, const double* trianorm,
                            const double* source, int ifcharge, const double* charge, int ifdipole,
                            const double* dipstr, const double* dipvec, int ntarget, const double* target,
                            int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  return fmm3d_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                              dipvec, 0, nullptr, 0, nullptr, ntarget, target, ifpottarg, pottarg, iffldtarg,
                              fldtarg);
}

void fmm3d_l3dtriadirect(int nsource, const double* triaflat, const double* trianorm, const double* source,
                         int ifcharge, const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld, int ntarget,
                         const double* target, int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  (void)trianorm;
  direct_host(nullptr, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld,
              ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg, triaflat, true);
}

int fmm3d_hfmm3dtriatarg(int iprec, const double* zk, int nsource, const double* triaflat,
                         const double* trianorm, const double* source, int ifcharge, const double* charge,
                         int ifdipole, const double* dipstr, const double* dipvec, int ifpot, double* pot, int iffld,
                         double* fld, int ntarget, const double* target, int ifpottarg, double* pottarg,
                         int iffldtarg, double* fldtarg) {
  (void)trianorm;  // the dipole layer is along the normal of the vertex order, as in the reference
  (void)dipvec;
  return htri_host(zk, iprec, nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ifpot, pot, iffld, fld,
                   ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

int fmm3d_hfmm3dtriaself(int iprec, const double* zk, int nsource, const double* triaflat,
                         const double* trianorm, const double* source, int ifcharge, const double* charge,
                         int ifdipole, const double* dipstr, const double* dipvec, int ifpot, double* pot, int iffld,
                         double* fld) {
  return fmm3d_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                              dipvec, ifpot, pot, iffld, fld, 0, nullptr, 0, nullptr, 0, nullptr);
}

int fmm3d_hfmm3dtriampftarg(int iprec, const double* zk, int nsource, const double* triaflat,
                            const double* trianorm, const double* source, int ifcharge, const double* charge,
                            int ifdipole, const double* dipstr, const double* dipvec, int ntarget,
                            const double* target, int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  return fmm3d_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                              dipvec, 0, nullptr, 0, nullptr, ntarget, target, ifpottarg, pottarg, iffldtarg,
                              fldtarg);
}

void fmm3d_h3dtriadirect(int nqtri, const double* zk, int nsource, const double* triaflat, const double* trianorm,
                         const double* source, int ifcharge, const double* charge, int ifdipole, const double* dipstr,
                         const double* dipvec, int ifpot, double* pot, int iffld, double* fld, int ntarget,
                         const double* target, int ifpottarg, double* pottarg, int iffldtarg, double* fldtarg) {
  (void)trianorm;
  (void)dipvec;
  htri_direct_host(nqtri, zk, nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ifpot, pot, iffld, fld,
                   ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg);
}

void fmm3d_last_timings(double* ms, int n) {
  const auto& t = last_timings();
  for (int i = 0; i < n; i++) ms[i] = i < (int)t.size() ? t[i] : 0.0;
}

int fmm3d_last_precision(void) { return last_precision(); }

const char* fmm3d_version(void) { return "fmm3d-mi355x 0.2 (gfx950)"; }

}  // extern "C"

extern "C" int fmm3d_d3tstrcr(int nsource, const double* source, int ntarget, const double* target,
                              int nbox, int* nboxes, int* nlev, double* center, double* size,
                              int* nlist, int maxboxes, int* boxes20, int* isource, int* itarget,
                              int* list_counts, int maxlist, int* lists) {
  if (ntarget < 0) return 128;
  Octree t;
  int ier = build_octree(source, nsource, target, ntarget, nbox, t);
  if (ier) return ier;
  *nboxes = t.nboxes;
  *nlev = t.nlev;
  for (int d = 0; d < 3; d++) center[d] = t.center[d];
  *size = t.size;
  long tot = 0;
  for (int k = 0; k < 4; k++) tot += (long)t.list[k].idx.size();
  *nlist = (int)tot;
  if (maxboxes == 0) return 0;
  if (maxboxes < t.nboxes || maxlist < tot) return 32;
  for (int b = 0; b < t.nboxes; b++) {
    int* o = boxes20 + 20 * (size_t)b;
    std::memset(o, 0, sizeof(int) * 20);
    o[0] = t.level[b];
    o[1] = t.ijk[3 * b]; o[2] = t.ijk[3 * b + 1]; o[3] = t.ijk[3 * b + 2];
    o[4] = t.parent[b] + 1;
    for (int c = 0; c < t.nchild[b]; c++) o[5 + c] = t.child[8 * b + c] + 1;
    o[13] = t.isrc[b] + 1; o[14] = t.nsrc[b]; o[15] = t.itrg[b] + 1; o[16] = t.ntrg[b];
  }
  for (int i = 0; i < nsource; i++) isource[i] = t.src_perm[i] + 1;
  for (int i = 0; i < ntarget; i++) itarget[i] = t.trg_perm[i] + 1;
  long o = 0;
  for (int k = 0; k < 4; k++)
    for (int b = 0; b < t.nboxes; b++) {
      list_counts[k * (size_t)maxboxes + b] = t.list[k].count(b);
      for (int q = t.list[k].off[b]; q < t.list[k].off[b + 1]; q++) lists[o++] = t.list[k].idx[q] + 1;
    }
  return 0;
}
