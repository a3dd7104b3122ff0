// Helmholtz translation tables.
//
// The expansions follow the reference's convention (src/helmrouts3d.f:3119,
// :3438): M_n^m = ik sum q j_n(kr)/sc^n Y_n^{-m}, L_n^m = ik sum q h_n(kr) sc^n
// Y_n^{-m}, Y_n^m = sqrt(2n+1) Yhat_n^m.  Every translation is rotate /
// coaxial shift / rotate back; the coaxial shift along the z axis by a
// distance d is, per order m, the matrix
//   T_{ln} = sum_s 4 pi i^{l+s-n} sigma^s f_s(kd) sqrt((2s+1)/4pi) G(n,s,l;m)
// (f = j for M2M and L2L, f = h for M2L; sigma = -1 when the rotation
// brought the source-minus-target direction onto +z, as for M2L here),
// G(n,s,l;m) the integral of three orthonormal spherical harmonics,
// computed exactly by Gauss-Legendre quadrature.  The reference computes the
// same operators by evaluating on a sphere and projecting
// (src/h3dtrans.f:211, :1229, :1998).  The powers of the scales are folded
// in with non-negative exponents only, so no intermediate overflows.
#include "helm_tables.hpp"

#include <cmath>
#include <map>
#include <memory>
#include <mutex>

namespace fmm3d {

void sph_hankel_scaled(int n, cplx z, double sc, cplx* hs) {
  if (std::abs(z) < 1.0e-15) {
    for (int i = 0; i <= n; i++) hs[i] = 0.0;
    return;
  }
  const cplx I(0.0, 1.0);
  const cplx h0 = std::exp(I * z) / (I * z);
  hs[0] = h0;
  if (n >= 1) hs[1] = h0 * (1.0 / z - I) * sc;
  const double sc2 = sc * sc;
  const cplx zinv = sc / z;
  for (int i = 1; i < n; i++) hs[i + 1] = zinv * (2.0 * i + 1.0) * hs[i] - sc2 * hs[i - 1];
}

int sph_bessel_scaled(int n, cplx z, double sc, cplx* fj) {
  const double up2 = 1.0e40, up2inv = 1.0e-40;
  if (std::abs(z) < 1.0e-200) {
    fj[0] = 1.0;
    for (int i = 1; i <= n; i++) fj[i] = 0.0;
    return 0;
  }
  const int nn = std::max(n, 1), lw = nn + 1000;
  std::vector<cplx> f(lw + 3);
  std::vector<char> isc(lw + 3, 0);
  const cplx zinv = 1.0 / z;
  f[nn] = 1.0;
  f[nn - 1] = 0.0;
  int ntop = 0;
  for (int i = nn; i <= lw; i++) {
    cplx t = (2.0 * i + 1.0) * zinv * f[i] - f[i - 1];
    f[i + 1] = t;
    if (std::norm(t) > up2) { ntop = i + 1; break; }
  }
  if (ntop == 0) return 8;
  f[ntop] = 0.0;
  f[ntop - 1] = 1.0;
  for (int i = ntop - 1; i >= 1; i--) {
    cplx t = (2.0 * i + 1.0) * zinv * f[i] - f[i + 1];
    f[i - 1] = t;
    if (std::norm(t) > up2) {
      f[i] *= up2inv;
      f[i - 1] *= up2inv;
      isc[i] = 1;
    }
  }
  double scinv = 1.0 / sc, sctot = 1.0;
  for (int i = 1; i <= ntop; i++) {
    sctot *= scinv;
    if (isc[i - 1]) sctot *= up2inv;
    f[i] *= sctot;
  }
  const cplx j0 = std::sin(z) * zinv, j1 = j0 * zinv - std::cos(z) * zinv;
  const cplx zs = (std::abs(j1) > std::abs(j0)) ? j1 / (f[1] * sc) : j0 / f[0];
  for (int i = 0; i <= n; i++) fj[i] = f[i] * zs;
  return 0;
}

// smallest j >= 2 with |j_j h_j| + |j_{j-1} h_{j-1}| < eps (|j_0 h_0| + |j_1 h_1|), as j+1
static int h_decay(cplx z1, cplx z2, double sc, double eps) {
  const int ntmax = 1000;
  std::vector<cplx> hf(ntmax + 1), jf(ntmax + 1);
  sph_hankel_scaled(ntmax, z1, sc, hf.data());
  if (sph_bessel_scaled(ntmax, z2, sc, jf.data()) != 0) return 1;
  const double x0 = std::abs(jf[0] * hf[0]) + std::abs(jf[1] * hf[1]);
  for (int j = 2; j <= ntmax; j++) {
    double x = std::abs(jf[j] * hf[j]) + std::abs(jf[j - 1] * hf[j - 1]);
    if (x < eps * x0) return j + 1;
  }
  return 1000;
}

int helmholtz_nterms(double size, cplx zk, double eps) {
  const double sc = std::min(std::abs(zk * size), 1.0);
  return h_decay(zk * size * 1.5, zk * size * std::sqrt(3.0) / 2.0, sc, eps);
}

std::vector<int> helmholtz_itable(double size, cplx zk, double eps) {
  const double sc = std::min(std::abs(zk * size), 1.0);
  int nt[2][4][4];
  for (int ii = 2; ii <= 3; ii++)
    for (int jj = 0; jj <= 3; jj++)
      for (int kk = 0; kk <= 3; kk++) {
        double dx = ii, dy = jj, dz = kk;
        if (dx > 0) dx -= 0.5;
        if (dy > 0) dy -= 0.5;
        if (dz > 0) dz -= 0.5;
        double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
        nt[ii - 2][jj][kk] = h_decay(zk * size * rr, zk * size * std::sqrt(3.0) / 2.0, sc, eps);
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

namespace {

// Gaunt integrals G(n,s,l;m) for 0 <= m <= min(n,l), n,l <= P,
// s = |n-l|, |n-l|+2, .., n+l, stored compactly; built once per P.
struct Gaunt {
  int P = -1;
  std::vector<long> off;   // [(m*(P+1)+n)*(P+1)+l]
  std::vector<double> g;
  long at(int m, int n, int l) const { return off[((size_t)m * (P + 1) + n) * (P + 1) + l]; }
};

void gauss_legendre(int q, std::vector<double>& x, std::vector<double>& w) {
  x.resize(q);
  w.resize(q);
  for (int i = 0; i < q; i++) {
    double t = std::cos(M_PI * (i + 0.75) / (q + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= q; k++) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = q * (t * p1 - p0) / (t * t - 1.0);
      double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    x[i] = t;
    w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

// normalised associated Legendre functions Theta_n^m(x) (orthonormal
// harmonics up to their e^{im phi}), 0 <= m <= n <= S
void theta_table(int S, double x, double* y) {
  const int N = S + 1;
  const double u = -std::sqrt((1.0 - x) * (1.0 + x));
  y[0] = 1.0;
  for (int m = 0; m <= S; m++) {
    if (m > 0) y[m * N + m] = y[(m - 1) * N + m - 1] * u * std::sqrt((2.0 * m - 1.0) / (2.0 * m));
    if (m < S) y[(m + 1) * N + m] = x * y[m * N + m] * std::sqrt(2.0 * m + 1.0);
    for (int n = m + 2; n <= S; n++)
      y[n * N + m] = ((2.0 * n - 1.0) * x * y[(n - 1) * N + m] -
                      std::sqrt((n + m - 1.0) * (n - m - 1.0)) * y[(n - 2) * N + m]) /
                     std::sqrt((double)(n - m) * (n + m));
  }
  for (int n = 0; n <= S; n++)
    for (int m = 0; m <= n; m++) y[n * N + m] *= std::sqrt((2.0 * n + 1.0) / (4.0 * M_PI));
}

std::shared_ptr<const Gaunt> gaunt_for(int P) {
  static std::mutex mu;
  static std::shared_ptr<const Gaunt> cached;
  std::lock_guard<std::mutex> lk(mu);
  if (cached && cached->P >= P) return cached;
  auto G = std::make_shared<Gaunt>();
  G->P = P;
  const int P1 = P + 1, S = 2 * P, Q = 2 * P + 2;
  G->off.assign((size_t)P1 * P1 * P1, 0);
  long tot = 0;
  for (int m = 0; m <= P; m++)
    for (int n = 0; n <= P; n++)
      for (int l = 0; l <= P; l++) {
        G->off[((size_t)m * P1 + n) * P1 + l] = tot;
        if (n >= m && l >= m) tot += std::min(n, l) + 1;
      }
  G->g.assign(std::max(tot, 1L), 0.0);
  std::vector<double> x, w;
  gauss_legendre(Q, x, w);
  const int N = S + 1;
  std::vector<double> th((size_t)Q * N * N);
  for (int q = 0; q < Q; q++) theta_table(S, x[q], &th[(size_t)q * N * N]);
#pragma omp parallel for schedule(dynamic, 1)
  for (int m = 0; m <= P; m++) {
    std::vector<double> prod(Q);
    for (int n = m; n <= P; n++)
      for (int l = m; l <= P; l++) {
        for (int q = 0; q < Q; q++)
          prod[q] = 2.0 * M_PI * w[q] * th[(size_t)q * N * N + n * N + m] * th[(size_t)q * N * N + l * N + m];
        long o = G->off[((size_t)m * P1 + n) * P1 + l];
        for (int s = std::abs(n - l), k = 0; s <= n + l; s += 2, k++) {
          double a = 0.0;
          for (int q = 0; q < Q; q++) a += prod[q] * th[(size_t)q * N * N + s * N];
          G->g[o + k] = a;
        }
      }
  }
  cached = G;
  return cached;
}

// mode 0: M2M (R|R), 1: M2L (S|R), 2: L2L (R|R); sigma = direction sign
void coax(const Gaunt& G, int mode, double sigma, cplx zk, double d, double sc_in, double sc_out,
          int pin, int pout, std::vector<double>& z, CoaxDesc& desc) {
  desc.pin = pin;
  desc.pout = pout;
  desc.base = (long)(z.size() / 2);
  const int S = pin + pout;
  std::vector<cplx> f(S + 2);
  const double base = (mode == 0) ? sc_out : sc_in;
  if (mode == 1) sph_hankel_scaled(S, zk * d, base, f.data());
  else sph_bessel_scaled(S, zk * d, base, f.data());
  const int mtop = std::min(pin, pout);
  const long total = coax_block(mtop + 1, pin, pout);
  // powers used below: base^e for 0 <= e <= S, ratio^e for 0 <= e <= max(pin, pout)
  std::vector<double> pb(S + 1, 1.0), pr(std::max(pin, pout) + 1, 1.0);
  const double ratio = (mode == 0) ? sc_in / sc_out : (mode == 2 ? sc_out / sc_in : 1.0);
  for (int e = 1; e <= S; e++) pb[e] = pb[e - 1] * base;
  for (int e = 1; e < (int)pr.size(); e++) pr[e] = pr[e - 1] * ratio;
  const size_t o0 = z.size();
  z.resize(o0 + 2 * (size_t)total, 0.0);
  for (int m = 0; m <= mtop; m++) {
    const long mb = coax_block(m, pin, pout);
    const int cols = pin + 1 - m;
    for (int l = m; l <= pout; l++)
      for (int n = m; n <= pin; n++) {
        const long go = G.at(m, n, l);
        cplx acc = 0.0;
        for (int s = std::abs(n - l), k = 0; s <= n + l; s += 2, k++) {
          double sg = (((l + s - n) / 2) & 1) ? -1.0 : 1.0;
          if (sigma < 0 && (s & 1)) sg = -sg;
          double c = 4.0 * M_PI * sg * std::sqrt((2.0 * s + 1.0) / (4.0 * M_PI)) * G.g[go + k];
          double pw;
          if (mode == 1) pw = pb[n + l - s];
          else if (mode == 0) pw = pb[s - l + n] * pr[n];
          else pw = pb[s - n + l] * pr[l];
          acc += c * pw * f[s];
        }
        const size_t e = o0 + 2 * (size_t)(mb + (long)(l - m) * cols + (n - m));
        z[e] = acc.real();
        z[e + 1] = acc.imag();
      }
  }
}

}  // namespace

bool build_helm_tables(cplx zk, double size, int nlev, double eps, HelmHostTables& H) {
  H = HelmHostTables();
  H.zk = zk;
  H.size = size;
  H.eps = eps;
  H.nlev = nlev;
  H.plev.assign(nlev + 1, 0);
  H.sclev.assign(nlev + 1, 1.0);
  std::vector<double> bsize(nlev + 1);
  int pmax = 1;
  for (int l = 0; l <= nlev; l++) {
    bsize[l] = size / std::pow(2.0, l);
    H.sclev[l] = std::min(std::abs(zk * bsize[l]), 1.0);
    if (l >= 2) {
      H.plev[l] = helmholtz_nterms(bsize[l], zk, eps);
      pmax = std::max(pmax, H.plev[l]);
    }
  }
  H.pmax = pmax;
  if (pmax > kMaxHelmOrder) return false;
  build_geometry_tables(pmax, H.geo);
  const int nd = H.geo.ndist;
  H.ptrunc.assign((size_t)(nlev + 1) * 343, 0);
  H.m2l.assign((size_t)(nlev + 1) * nd, CoaxDesc());
  H.m2m.assign(nlev + 1, CoaxDesc());
  H.l2l.assign(nlev + 1, CoaxDesc());
  auto G = gaunt_for(pmax);
  for (int l = 2; l <= nlev; l++) {
    const std::vector<int> it = helmholtz_itable(bsize[l], zk, eps);
    std::vector<int> need(nd, 0);
    for (int code = 0; code < 343; code++) {
      int i = code / 49 - 3, j = (code / 7) % 7 - 3, k = code % 7 - 3;
      if (std::max(std::abs(i), std::max(std::abs(j), std::abs(k))) < 2) continue;
      int pt = std::min(it[code], H.plev[l]);
      H.ptrunc[(size_t)l * 343 + code] = pt;
      int dc = H.geo.off_dist[code];
      need[dc] = std::max(need[dc], pt);
    }
    std::vector<std::vector<double>> part(nd);
    std::vector<CoaxDesc> desc(nd);
#pragma omp parallel for schedule(dynamic, 1)
    for (int dc = 0; dc < nd; dc++)
      if (need[dc] > 0)
        coax(*G, 1, -1.0, zk, bsize[l] * H.geo.dist_value[dc], H.sclev[l], H.sclev[l], need[dc], need[dc],
             part[dc], desc[dc]);
    for (int dc = 0; dc < nd; dc++) {
      if (need[dc] == 0) continue;
      desc[dc].base = (long)(H.z.size() / 2);
      H.z.insert(H.z.end(), part[dc].begin(), part[dc].end());
      H.m2l[(size_t)l * nd + dc] = desc[dc];
    }
  }
  for (int l = 2; l <= nlev - 1; l++) {
    const double d = bsize[l + 1] * std::sqrt(3.0) / 2.0;
    coax(*G, 0, 1.0, zk, d, H.sclev[l + 1], H.sclev[l], H.plev[l + 1], H.plev[l], H.z, H.m2m[l]);
    coax(*G, 2, 1.0, zk, d, H.sclev[l], H.sclev[l + 1], H.plev[l], H.plev[l + 1], H.z, H.l2l[l]);
  }
  return true;
}

}  // namespace fmm3d
