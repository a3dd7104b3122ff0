// Host-built, level-independent translation tables for one expansion order p.
//
// Rotations: Wigner small-d matrices by coupling one spin-1/2 at a time
// (each half step averages the previous matrix with cos/sin(beta/2)
// weights, stable to high degree; the reference uses its own recursion,
// src/rotviarecur3.f:277).  Stored per rotation class and degree as
// D[m][m'] = e_m e_m' d^n_{m'm}(theta), e_k = (-1)^k for k > 0 else 1, which is
// the rotation of coefficients in the reference's convention.
// z-shifts: closed forms of the reference's l3dmpmpzshift, l3dmploczshiftstab
// and l3dlocloczshift (src/l3dtrans.f:351, :990, :1798) in scaled units.
#include "tables.hpp"

#include <cmath>
#include <map>

namespace fmm3d {

static std::vector<std::vector<double>> wigner_d(int p, double beta) {
  const double a = std::cos(0.5 * beta), s = std::sin(0.5 * beta);
  std::vector<std::vector<double>> out(p + 1);
  std::vector<double> cur(1, 1.0), nxt;
  out[0] = cur;
  for (int tj = 0; tj < 2 * p; tj++) {
    const int tj2 = tj + 1, dim = tj2 + 1;
    const double j = 0.5 * tj2;
    nxt.assign((size_t)dim * dim, 0.0);
    auto prev = [&](int a2, int b2) -> double {
      if (std::abs(a2) > tj || std::abs(b2) > tj) return 0.0;
      return cur[(size_t)((a2 + tj) / 2) * (tj + 1) + (b2 + tj) / 2];
    };
    for (int ip = 0; ip < dim; ip++) {
      const int mp2 = 2 * ip - tj2;
      const double mp = 0.5 * mp2;
      for (int iq = 0; iq < dim; iq++) {
        const int m2 = 2 * iq - tj2;
        const double m = 0.5 * m2;
        double v = std::sqrt((j + mp) * (j + m)) * a * prev(mp2 - 1, m2 - 1);
        v -= std::sqrt((j + mp) * (j - m)) * s * prev(mp2 - 1, m2 + 1);
        v += std::sqrt((j - mp) * (j + m)) * s * prev(mp2 + 1, m2 - 1);
        v += std::sqrt((j - mp) * (j - m)) * a * prev(mp2 + 1, m2 + 1);
        nxt[(size_t)ip * dim + iq] = v / (2.0 * j);
      }
    }
    cur.swap(nxt);
    if (tj2 % 2 == 0) out[tj2 / 2] = cur;
  }
  return out;
}

// l3dterms.f:34-77 decay rule
static int terms_decay(double z1, double z2, double eps) {
  double jfun = z2, hfun = 1.0 / (z1 * z1);
  for (int j = 2; j <= 1000; j++) {
    hfun /= z1;
    jfun *= z2;
    if (jfun * hfun < eps) return j;
  }
  return 1;
}

int laplace_nterms(double eps) { return terms_decay(1.5, std::sqrt(3.0) / 2.0, eps); }

// l3dterms.f:136-248 list-2 table, indexed by code (i+3)*49+(j+3)*7+(k+3)
std::vector<int> laplace_itable(double eps) {
  int nt[2][4][4];
  for (int ii = 2; ii <= 3; ii++)
    for (int jj = 0; jj <= 3; jj++)
      for (int kk = 0; kk <= 3; kk++) {
        double dx = ii, dy = jj, dz = kk;
        if (dx > 0) dx -= 0.5;
        if (dy > 0) dy -= 0.5;
        if (dz > 0) dz -= 0.5;
        nt[ii - 2][jj][kk] = terms_decay(std::sqrt(dx * dx + dy * dy + dz * dz), std::sqrt(3.0) / 2.0, eps);
      }
  std::vector<int> t(343, 0);
  for (int i = -3; i <= 3; i++)
    for (int j = -3; j <= 3; j++)
      for (int k = -3; k <= 3; k++) {
        int ai = std::abs(i), aj = std::abs(j), ak = std::abs(k), v = 0;
        if (ai > 1) v = nt[ai - 2][aj][ak];
        else if (aj > 1) v = nt[aj - 2][ai][ak];
        if (ai <= 1 && aj <= 1 && ak > 1) v = (ai >= aj) ? nt[ak - 2][ai][aj] : nt[ak - 2][aj][ai];
        t[((i + 3) * 7 + (j + 3)) * 7 + (k + 3)] = v;
      }
  return t;
}

// sqrt(C(a,b)) from a Pascal triangle (every entry is a sum of two rounded
// entries, so the relative error stays at a few ulps up to a = 2p+2).
struct SqrtBinom {
  int N;
  std::vector<double> t;
  explicit SqrtBinom(int nmax) : N(nmax + 1), t((size_t)(nmax + 1) * (nmax + 1), 0.0) {
    std::vector<double> c((size_t)N * N, 0.0);
    for (int a = 0; a < N; a++) {
      c[(size_t)a * N] = 1.0;
      for (int b = 1; b <= a; b++) c[(size_t)a * N + b] = c[(size_t)(a - 1) * N + b - 1] + (b < a ? c[(size_t)(a - 1) * N + b] : 0.0);
    }
    for (size_t i = 0; i < c.size(); i++) t[i] = std::sqrt(c[i]);
  }
  double operator()(int a, int b) const { return (b < 0 || b > a) ? 0.0 : t[(size_t)a * N + b]; }
};

// A-operand order of a 16x16x4 MFMA: lane l holds A[l % 16][l / 16] of a
// 16 x 4 block (row tile rt, column block kb of the matrix).
//
// Rotation of degree n, forward: out_{m'} = sum_m F[m'][m] in_m with
// F[m'][m] = rot block (row m + n, column m' + n); only rows m' >= 0 are
// stored, the negative orders follow from F[-m'][-m] = F[m'][m] (the kernel
// applies the same rows to the reversed input).  Back: the transpose,
// out_m = sum_{m'} F[m'][m] in_{m'}, which has the same symmetry.
// z-shift of order m: out_l = sum_n Z[m][l][n] in_n for l, n = m..p.
static void build_mfma_tables(int p, HostTables& H) {
  const int P1 = p + 1;
  H.mf_rot_off.assign((size_t)H.nrot * P1, 0);
  long tot = 0;
  for (int c = 0; c < H.nrot; c++)
    for (int n = 0; n <= p; n++) {
      H.mf_rot_off[(size_t)c * P1 + n] = tot;
      tot += 64L * mf_rot_tiles(n) * mf_rot_kblocks(n);
    }
  H.mf_rot_back = tot;
  H.mf_rot.assign(2 * (size_t)tot, 0.0);
  for (int c = 0; c < H.nrot; c++)
    for (int n = 0; n <= p; n++) {
      const int dim = 2 * n + 1, nkb = mf_rot_kblocks(n);
      const double* D = &H.rot[H.rot_off[(size_t)c * P1 + n]];
      double* fw = &H.mf_rot[H.mf_rot_off[(size_t)c * P1 + n]];
      double* bk = fw + tot;
      for (int rt = 0; rt < mf_rot_tiles(n); rt++)
        for (int kb = 0; kb < nkb; kb++)
          for (int l = 0; l < 64; l++) {
            const int r = 16 * rt + l % 16, kap = 4 * kb + l / 16;
            if (r > n || kap >= dim) continue;
            const size_t at = ((size_t)rt * nkb + kb) * 64 + l;
            fw[at] = D[(size_t)kap * dim + (r + n)];   // F[r][kap - n]
            bk[at] = D[(size_t)(r + n) * dim + kap];   // F[kap - n][r]
          }
    }
  H.mf_z_off.assign(P1, 0);
  long zs = 0;
  for (int m = 0; m <= p; m++) {
    H.mf_z_off[m] = zs;
    zs += 64L * mf_z_tiles(p, m) * mf_z_kblocks(p, m);
  }
  H.mf_z_size = zs;
  H.mf_z.assign((size_t)zs * H.ndist, 0.0);
  const size_t Z = (size_t)P1 * P1 * P1;
  for (int dc = 0; dc < H.ndist; dc++)
    for (int m = 0; m <= p; m++) {
      const int nkb = mf_z_kblocks(p, m);
      double* zb = &H.mf_z[(size_t)zs * dc + H.mf_z_off[m]];
      for (int lt = 0; lt < mf_z_tiles(p, m); lt++)
        for (int kb = 0; kb < nkb; kb++)
          for (int l = 0; l < 64; l++) {
            const int lo = m + 16 * lt + l % 16, ni = m + 4 * kb + l / 16;
            if (lo > p || ni > p) continue;
            zb[((size_t)lt * nkb + kb) * 64 + l] = H.z_m2l[Z * dc + ((size_t)m * P1 + lo) * P1 + ni];
          }
    }
}

void build_tables(int p, double eps, HostTables& H) {
  build_geometry_tables(p, H);
  const int P1 = p + 1;
  const std::vector<int> itab = laplace_itable(eps);
  std::vector<double> dists;
  for (int code = 0; code < 343; code++) H.off_ptrunc[code] = 0;
  for (int i = -3; i <= 3; i++)
    for (int j = -3; j <= 3; j++)
      for (int k = -3; k <= 3; k++) {
        int code = ((i + 3) * 7 + (j + 3)) * 7 + (k + 3);
        if (std::max(std::abs(i), std::max(std::abs(j), std::abs(k))) < 2) continue;
        H.off_ptrunc[code] = std::min(p, itab[code]);
      }
  for (int dc = 0; dc < H.ndist; dc++) dists.push_back(H.dist_value[dc]);
  // z-shift tables, dense [m][out][in]
  const SqrtBinom sqrt_binom(2 * p + 2);
  const size_t Z = (size_t)P1 * P1 * P1;
  H.z_m2l.assign(Z * H.ndist, 0.0);
  for (int dc = 0; dc < H.ndist; dc++) {
    double rho = dists[dc];
    double* z = &H.z_m2l[Z * dc];
    for (int m = 0; m <= p; m++)
      for (int l = m; l <= p; l++)
        for (int n = m; n <= p; n++) {
          double c = sqrt_binom(n + l, n - m) * sqrt_binom(n + l, n + m) * std::pow(rho, -(n + l + 1));
          if ((n + m) & 1) c = -c;
          z[((size_t)m * P1 + l) * P1 + n] = c;
        }
  }
  // the same coefficients for the grouped M2L in 4x4 tiles per order m:
  // [row block of n - m][column block of l - m][row][column], zero past p
  H.zt_off.assign(P1 + 1, 0);
  for (int m = 0; m <= p; m++) {
    const long nb4 = (p - m) / 4 + 1;
    H.zt_off[m + 1] = H.zt_off[m] + 16 * nb4 * nb4;
  }
  H.zt_size = H.zt_off[P1];
  H.z_m2l_t.assign((size_t)H.zt_size * H.ndist, 0.0);
  for (int dc = 0; dc < H.ndist; dc++)
    for (int m = 0; m <= p; m++) {
      const int nb4 = (p - m) / 4 + 1;
      double* zb = &H.z_m2l_t[(size_t)H.zt_size * dc + H.zt_off[m]];
      for (int n = m; n <= p; n++)
        for (int l = m; l <= p; l++) {
          const int rr = n - m, cc = l - m;
          zb[(((size_t)(rr / 4) * nb4 + cc / 4) * 4 + rr % 4) * 4 + cc % 4] =
              H.z_m2l[Z * dc + ((size_t)m * P1 + l) * P1 + n];
        }
    }
  build_mfma_tables(p, H);
  H.z_m2m.assign(Z, 0.0);
  H.z_l2l.assign(Z, 0.0);
  const double r = std::sqrt(3.0) / 4.0;
  for (int m = 0; m <= p; m++)
    for (int n = m; n <= p; n++)
      for (int l = m; l <= n; l++) {
        double c = sqrt_binom(n + m, l + m) * sqrt_binom(n - m, l - m) * std::ldexp(1.0, -l);
        // parent degree n from child degree l (M2M), child degree l from parent degree n (L2L)
        H.z_m2m[((size_t)m * P1 + n) * P1 + l] = c * std::pow(-r, n - l);
        H.z_l2l[((size_t)m * P1 + l) * P1 + n] = c * std::pow(r, n - l);
      }
}

// Factor table, rotation classes, rotation blocks and phases for order p:
// the kernel-independent half of the translation tables.
void build_fac_table(int p, std::vector<double>& fac) {
  const int P = p + 2, Hh = P * (P + 1) / 2, E = P * P;
  fac.assign((size_t)2 * (P + 1) + 2 * Hh + 6 * E, 0.0);
  const int MM = 0, M1 = MM + P + 1, RA = M1 + P + 1, RB = RA + Hh, AR = RB + Hh, BR = AR + E,
            CR = BR + E, AS = CR + E, BS = AS + E, CS = BS + E;
  for (int m = 1; m <= P; m++) fac[MM + m] = std::sqrt((2.0 * m - 1.0) / (2.0 * m));
  for (int m = 0; m <= P; m++) fac[M1 + m] = std::sqrt(2.0 * m + 1.0);
  for (int n = 0; n < P; n++)
    for (int m = 0; m <= n; m++) {
      int h = n * (n + 1) / 2 + m;
      fac[RA + h] = (n >= m + 2) ? std::sqrt((n + m - 1.0) * (n - m - 1.0)) : 0.0;
      fac[RB + h] = (n > m) ? 1.0 / std::sqrt((double)(n - m) * (n + m)) : 0.0;
    }
  for (int n = 0; n < P; n++)
    for (int m = -n; m <= n; m++) {
      int e = n * n + n + m;
      fac[AR + e] = std::sqrt((double)(n - m) * (n + m));
      fac[BR + e] = std::sqrt(std::max(0.0, (double)(n - m) * (n - m - 1))) * (m >= 0 ? 1.0 : -1.0);
      fac[CR + e] = std::sqrt(std::max(0.0, (double)(n + m) * (n + m - 1))) * (m >= 1 ? -1.0 : 1.0);
      fac[AS + e] = -std::sqrt((double)(n + 1 - m) * (n + 1 + m));
      fac[BS + e] = std::sqrt((double)(n + m + 1) * (n + m + 2)) * (m >= 0 ? 1.0 : -1.0);
      fac[CS + e] = std::sqrt((double)(n - m + 1) * (n - m + 2)) * (m >= 1 ? -1.0 : 1.0);
    }
}

void build_geometry_tables(int p, HostTables& H) {
  H = HostTables();
  H.p = p;
  const int P1 = p + 1;
  // --- factor table (layout mirrored by FacLayout in lap_kernels.hip)
  build_fac_table(p, H.fac);
  // --- rotation classes: one per distinct polar angle of the list-2 offsets
  //     and of the child-parent directions (+-1,+-1,+-1)
  std::map<long long, int> theta_class;
  std::vector<double> thetas;
  auto class_of = [&](double th) {
    long long key = std::llround(th * 1e12);
    auto it = theta_class.find(key);
    if (it != theta_class.end()) return it->second;
    int c = (int)thetas.size();
    theta_class[key] = c;
    thetas.push_back(th);
    return c;
  };
  std::map<long long, int> dist_class;
  std::vector<double> dists;
  H.off_rot.assign(343, 0);
  H.off_dist.assign(343, 0);
  H.off_ptrunc.assign(343, 0);
  H.off_phase.assign((size_t)343 * 2 * P1, 0.0);
  for (int i = -3; i <= 3; i++)
    for (int j = -3; j <= 3; j++)
      for (int k = -3; k <= 3; k++) {
        int code = ((i + 3) * 7 + (j + 3)) * 7 + (k + 3);
        if (std::max(std::abs(i), std::max(std::abs(j), std::abs(k))) < 2) continue;
        double rho = std::sqrt((double)(i * i + j * j + k * k));
        double th = std::atan2(std::sqrt((double)(i * i + j * j)), (double)k);
        double ph = (i == 0 && j == 0) ? 0.0 : std::atan2((double)j, (double)i);
        H.off_rot[code] = class_of(th);
        long long dk = (long long)(i * i + j * j + k * k);
        auto it = dist_class.find(dk);
        if (it == dist_class.end()) {
          dist_class[dk] = (int)dists.size();
          H.off_dist[code] = (int)dists.size();
          dists.push_back(rho);
        } else {
          H.off_dist[code] = it->second;
        }
        for (int m = 0; m <= p; m++) {
          H.off_phase[(size_t)code * 2 * P1 + 2 * m] = std::cos(m * ph);
          H.off_phase[(size_t)code * 2 * P1 + 2 * m + 1] = std::sin(m * ph);
        }
      }
  // child-parent directions: octant bits (x,y,z), child = parent + s h/4
  H.oct_rot.assign(16, 0);
  H.oct_phase.assign((size_t)16 * 2 * P1, 0.0);
  for (int o = 0; o < 8; o++) {
    double s[3] = {(o & 1) ? 1.0 : -1.0, (o & 2) ? 1.0 : -1.0, (o & 4) ? 1.0 : -1.0};
    for (int w = 0; w < 2; w++) {  // w=0: M2M (parent - child), w=1: L2L (child - parent)
      double sg = w == 0 ? -1.0 : 1.0;
      double v[3] = {sg * s[0], sg * s[1], sg * s[2]};
      double th = std::atan2(std::sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
      double ph = std::atan2(v[1], v[0]);
      H.oct_rot[o * 2 + w] = class_of(th);
      for (int m = 0; m <= p; m++) {
        H.oct_phase[(size_t)(o * 2 + w) * 2 * P1 + 2 * m] = std::cos(m * ph);
        H.oct_phase[(size_t)(o * 2 + w) * 2 * P1 + 2 * m + 1] = std::sin(m * ph);
      }
    }
  }
  // rotation blocks
  const int nrot = (int)thetas.size();
  H.rot_off.assign((size_t)nrot * P1, 0);
  long total = 0;
  for (int c = 0; c < nrot; c++)
    for (int n = 0; n <= p; n++) {
      H.rot_off[(size_t)c * P1 + n] = total;
      total += (long)(2 * n + 1) * (2 * n + 1);
    }
  H.rot.assign(total, 0.0);
  for (int c = 0; c < nrot; c++) {
    auto d = wigner_d(p, thetas[c]);
    for (int n = 0; n <= p; n++) {
      const int dim = 2 * n + 1;
      double* D = &H.rot[H.rot_off[(size_t)c * P1 + n]];
      for (int m = -n; m <= n; m++)
        for (int mp = -n; mp <= n; mp++) {
          double em = (m > 0 && (m & 1)) ? -1.0 : 1.0, emp = (mp > 0 && (mp & 1)) ? -1.0 : 1.0;
          D[(m + n) * dim + (mp + n)] = em * emp * d[n][(size_t)(mp + n) * dim + (m + n)];
        }
    }
  }
  H.nrot = nrot;
  // factored form: one Delta = d(pi/2) for every class plus a phase per class
  {
    auto dd = wigner_d(p, M_PI / 2.0);
    H.delta_off.assign(P1, 0);
    long tot = 0;
    for (int n = 0; n <= p; n++) {
      H.delta_off[n] = tot;
      tot += (long)(2 * n + 1) * (2 * n + 1);
    }
    H.delta.assign(tot, 0.0);
    for (int n = 0; n <= p; n++) {
      const int dim = 2 * n + 1;
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) H.delta[H.delta_off[n] + (long)a * dim + b] = dd[n][(size_t)a * dim + b];
    }
    H.rot_e.assign((size_t)nrot * 2 * P1, 0.0);
    for (int c = 0; c < nrot; c++)
      for (int k = 0; k <= p; k++) {
        H.rot_e[((size_t)c * P1 + k) * 2] = std::cos(k * thetas[c]);
        H.rot_e[((size_t)c * P1 + k) * 2 + 1] = std::sin(k * thetas[c]);
      }
  }
  H.ndist = (int)dists.size();
  H.dist_value = dists;
  // work items of the grouped M2L: a rotation item is (degree n, columns
  // 4q..4q+3), ordered by n so that the items of degrees <= pt are a prefix;
  // a z-shift item is (order |m|, degrees l = |m|+4q..+3), listed per pt
  H.grp_ritems.clear();
  H.grp_rcount.assign(P1, 0);
  for (int n = 0; n <= p; n++) {
    for (int q = 0; 4 * q <= n; q++) H.grp_ritems.push_back((n << 8) | q);
    H.grp_rcount[n] = (int)H.grp_ritems.size();
  }
  H.grp_zitems.clear();
  H.grp_zoff.assign(P1 + 1, 0);
  for (int pt = 0; pt <= p; pt++) {
    H.grp_zoff[pt] = (int)H.grp_zitems.size();
    for (int am = 0; am <= pt; am++)
      for (int q = 0; am + 4 * q <= pt; q++) H.grp_zitems.push_back((am << 8) | q);
  }
  H.grp_zoff[P1] = (int)H.grp_zitems.size();
  // rotation blocks for the grouped M2L: columns m' >= 0 only (the others
  // follow from D[-m][-m'] = D[m][m']), cut into 4x4 tiles [row block][column
  // block][row][column] so that one work item (four columns) reads one
  // contiguous 128-byte tile per four rows; rows past 2n and columns past n
  // are zero
  H.rot_half_off.assign((size_t)nrot * P1, 0);
  long th = 0;
  for (int c = 0; c < nrot; c++)
    for (int n = 0; n <= p; n++) {
      H.rot_half_off[(size_t)c * P1 + n] = th;
      th += 16L * (2 * n / 4 + 1) * (n / 4 + 1);
    }
  H.rot_half.assign(th, 0.0);
  for (int c = 0; c < nrot; c++)
    for (int n = 0; n <= p; n++) {
      const int dim = 2 * n + 1, nq = n / 4 + 1;
      const double* D = &H.rot[H.rot_off[(size_t)c * P1 + n]];
      double* Db = &H.rot_half[H.rot_half_off[(size_t)c * P1 + n]];
      for (int row = 0; row < dim; row++)
        for (int mp = 0; mp <= n; mp++)
          Db[(((size_t)(row / 4) * nq + mp / 4) * 4 + row % 4) * 4 + mp % 4] = D[row * dim + (mp + n)];
    }
}

}  // namespace fmm3d
