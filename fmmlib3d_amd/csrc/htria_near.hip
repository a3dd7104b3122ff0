// Launchers of the Helmholtz triangle near-field kernels (k_htri_p2p,
// k_htri_direct), in a translation unit of their own.  It is compiled with
// machine loop-invariant code motion off (-mllvm -disable-machine-licm):
// otherwise the constants of the inlined FP64 sqrt, sincos, log and atan2
// (about sixty) are hoisted into vector registers ahead of the pair loop,
// and those kernels need 211 VGPRs instead of 256 plus 10 AGPRs (two waves
// per SIMD instead of one); measured on 327,680 triangles + 100,000
// targets at zk = 10, iprec 2: near field 65.8 -> 40.1 ms, same results.
#include "htria_kernels.hpp"
#include "plan.hpp"

namespace fmm3d {

void launch_htri_p2p(const DevPlan& P, const int* items, int n, int ifcharge, int ifdipole, const double* geo,
                     const double* tstr, const double* cpos, const double* tpos, int kind, const HtriParams& hp,
                     double* pot, double* fld, hipStream_t s) {
  if (n == 0 || !(ifcharge || ifdipole)) return;
  DevPlan Q = P;
  Q.p2p_items = items;
  Q.n_p2p = n;
  const dim3 g(item_groups(n)), bl(kP2PBlock * kItemWaves);
  if (ifcharge && ifdipole)
    hipLaunchKernelGGL((k_htri_p2p<true, true>), g, bl, 0, s, Q, geo, tstr, cpos, tpos, kind, hp, pot, fld);
  else if (ifcharge)
    hipLaunchKernelGGL((k_htri_p2p<true, false>), g, bl, 0, s, Q, geo, tstr, cpos, tpos, kind, hp, pot, fld);
  else
    hipLaunchKernelGGL((k_htri_p2p<false, true>), g, bl, 0, s, Q, geo, tstr, cpos, tpos, kind, hp, pot, fld);
}

void launch_htri_direct(int n, const double* geo, const double* tstr, int npts, const double* pts, int centroids,
                        const HtriParams& hp, int ifcharge, int ifdipole, double* pot, double* fld, hipStream_t s) {
  if (npts <= 0 || !(ifcharge || ifdipole)) return;
  const dim3 g(blocks(npts, 256)), bl(256);
  if (ifcharge && ifdipole)
    hipLaunchKernelGGL((k_htri_direct<true, true>), g, bl, 0, s, n, geo, tstr, npts, pts, centroids, hp, pot, fld);
  else if (ifcharge)
    hipLaunchKernelGGL((k_htri_direct<true, false>), g, bl, 0, s, n, geo, tstr, npts, pts, centroids, hp, pot, fld);
  else
    hipLaunchKernelGGL((k_htri_direct<false, true>), g, bl, 0, s, n, geo, tstr, npts, pts, centroids, hp, pot, fld);
}

}  // namespace fmm3d
