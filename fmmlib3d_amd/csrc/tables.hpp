// Host-built translation tables (see tables.cpp) and expansion orders.
#pragma once
#include <vector>

namespace fmm3d {

struct HostTables {
  int p = -1, nrot = 0, ndist = 0;
  std::vector<double> fac;                         // FacLayout(p)
  std::vector<int> off_rot, off_dist, off_ptrunc;  // 343 offset codes
  std::vector<double> off_phase;                   // 343 x (p+1) complex
  std::vector<int> oct_rot;                        // 8 octants x {M2M, L2L}
  std::vector<double> oct_phase;                   // 16 x (p+1) complex
  std::vector<double> rot;                         // rotation blocks
  std::vector<long> rot_off;                       // nrot x (p+1)
  std::vector<double> z_m2l, z_m2m, z_l2l;         // dense [m][out][in]
  std::vector<double> dist_value;                  // |offset| of each distance class
  // Factored rotations: d^n(theta) = P Delta^n E(theta) Delta^nT P*, with
  // Delta^n = d^n(pi/2), P = diag(i^m), E = diag(e^{i m theta}).
  std::vector<double> delta;                       // per degree (2n+1)^2, row a col b = d^n_ab(pi/2)
  std::vector<long> delta_off;                     // p+1
  std::vector<double> rot_e;                       // nrot x (p+1) complex: e^{i k theta_class}
  // grouped M2L work items (see build_geometry_tables)
  std::vector<int> grp_ritems, grp_rcount;         // (n << 8 | q); count of items with n <= pt
  std::vector<int> grp_zitems, grp_zoff;           // (|m| << 8 | q) per pt; start of pt's items
  std::vector<double> rot_half;                    // per class and degree: 4x4 tiles of [m][m' >= 0]
  std::vector<long> rot_half_off;                  // nrot x (p+1)
  std::vector<double> z_m2l_t;                     // per distance class and order m: 4x4 tiles of [n-m][l-m]
  std::vector<long> zt_off;                        // start of order m within a class, p+2 entries
  long zt_size = 0;                                // doubles per distance class
};

// l3dterms (src/l3dterms.f:34): expansion order for tolerance eps
int laplace_nterms(double eps);
// l3dterms_list2 (src/l3dterms.f:136): 7^3 table indexed by offset code
std::vector<int> laplace_itable(double eps);
void build_tables(int p, double eps, HostTables& H);
// the kernel-independent part (factors, rotation classes and blocks, phases,
// distance classes) that build_tables starts from
void build_geometry_tables(int p, HostTables& H);

}  // namespace fmm3d
