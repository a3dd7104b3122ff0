// Host-side pieces of the Helmholtz FMM: spherical Bessel / Hankel series,
// expansion orders (h3dterms), and the per-level translation tables.
#pragma once
#include <complex>
#include <vector>

#include "tables.hpp"

namespace fmm3d {

using cplx = std::complex<double>;

// h3dall (src/helmrouts3d.f:690): hs[i] = h_i(z) sc^i, i = 0..n
void sph_hankel_scaled(int n, cplx z, double sc, cplx* hs);
// jfuns3d (src/cdjseval3d.f:15): fj[i] = j_i(z) / sc^i, i = 0..n; 0 or 8
int sph_bessel_scaled(int n, cplx z, double sc, cplx* fj);

// h3dterms (src/h3dterms.f:28): expansion order for a box of side `size`
int helmholtz_nterms(double size, cplx zk, double eps);
// h3dterms_list2 (src/h3dterms.f:192): 7^3 table by offset code
std::vector<int> helmholtz_itable(double size, cplx zk, double eps);

// One coaxial shift operator, packed by m: for m = 0..min(pin,pout) a
// (pout+1-m) x (pin+1-m) complex block (rows l = m..pout, columns n = m..pin).
struct CoaxDesc {
  long base = 0;   // offset in complex entries into HelmHostTables::z
  int pin = 0, pout = 0;
};

struct HelmHostTables {
  cplx zk;
  double size = 0, eps = 0;
  int nlev = -1, pmax = 0;
  HostTables geo;                    // factors, rotations, phases at pmax
  std::vector<int> plev;             // expansion order per level (0 below level 2)
  std::vector<double> sclev;         // rscale per level
  std::vector<int> ptrunc;           // (nlev+1) x 343 list-2 orders
  std::vector<CoaxDesc> m2l;         // (nlev+1) x ndist
  std::vector<CoaxDesc> m2m, l2l;    // per parent level
  std::vector<double> z;             // all coaxial blocks, complex as (re, im)
};

// Builds every table the Helmholtz FMM needs for a tree of root side `size`
// and depth `nlev`.  Returns false when an order exceeds what the device
// kernels support (kMaxHelmOrder).
bool build_helm_tables(cplx zk, double size, int nlev, double eps, HelmHostTables& H);

constexpr int kMaxHelmOrder = 120;

// offset (complex entries) of the m-block inside a packed coaxial operator
inline long coax_block(int m, int pin, int pout) {
  const long A = pout + 1, B = pin + 1, M = m;
  return M * A * B - (A + B) * M * (M - 1) / 2 + (M - 1) * M * (2 * M - 1) / 6;
}

}  // namespace fmm3d
