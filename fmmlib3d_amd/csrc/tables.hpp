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
  // M2L on the matrix cores (16x16x4 MFMA operand order, see
  // build_mfma_tables): per rotation class and degree n the forward and back
  // rotation rows m' = 0..n in blocks (row tile of 16, column block of 4),
  // each block the 64 A-operand values in lane order; per distance class and
  // order m the z-shift rows l = m..p likewise
  std::vector<double> mf_rot;                      // forward blocks, then back blocks at + mf_rot_back
  std::vector<long> mf_rot_off;                    // nrot x (p+1)
  long mf_rot_back = 0;
  std::vector<double> mf_z;                        // ndist x mf_z_size
  std::vector<long> mf_z_off;                      // start of order m within a class, p+1 entries
  long mf_z_size = 0;
};

// Blocks of a degree-n rotation in the MFMA tables: row tiles and column
// blocks of 4 (rows m' = 0..n, columns = the 2n+1 input orders).
inline int mf_rot_tiles(int n) { return n / 16 + 1; }
inline int mf_rot_kblocks(int n) { return (2 * n) / 4 + 1; }
// z-shift of order m at order p: row tiles of l - m and column blocks of n - m
inline int mf_z_tiles(int p, int m) { return (p - m) / 16 + 1; }
inline int mf_z_kblocks(int p, int m) { return (p - m) / 4 + 1; }

// l3dterms (src/l3dterms.f:34): expansion order for tolerance eps
int laplace_nterms(double eps);
// l3dterms_list2 (src/l3dterms.f:136): 7^3 table indexed by offset code
std::vector<int> laplace_itable(double eps);
void build_tables(int p, double eps, HostTables& H);
// the factor table alone (FacLayout(p): degrees up to p+1), as built for order p
void build_fac_table(int p, std::vector<double>& fac);
// the kernel-independent part (factors, rotation classes and blocks, phases,
// distance classes) that build_tables starts from
void build_geometry_tables(int p, HostTables& H);

}  // namespace fmm3d
