// Leaf sizes and quadrature rules of the triangle FMMs.
//
// Far field: each triangle becomes the nodes of a rule on it -- up to
// degree 6 the symmetric rules below, beyond that a product Gauss rule (the
// square [0,1]^2 collapsed onto the reference triangle, exact to degree
// 2q - 2).
//
// Near field of the Helmholtz triangle FMM: the smooth part
// (e^{ikr} - 1)/r of the kernel is integrated with the D3-symmetric rules
// of Xiao and Gimbutas that the reference takes from triasymq
// (src/triasymq.f:18) -- 1, 3, 6, 6, 7 and 12 nodes for degrees 1..6, the
// degree-3 rule being the degree-4 one.  Each rule is a set of orbits of
// the centroid (S3), of the barycentric points (a, a, 1-2a) (S21) and of
// (a, b, 1-a-b) (S111); their parameters solve the moment equations
//   sum_q w_q xi_q^i eta_q^j = i! j! / (i+j+2)!,   i + j <= degree,
// found here by Gauss-Newton from two-digit starting values.
#include "tria_rules.hpp"

#include <cmath>
#include <mutex>

namespace fmm3d {

int tri_nbox(int iprec, int ns, int nt) {
  static const int nb[] = {2, 5, 10, 20, 40, 80, 160, 233};
  return iprec == 6 ? ns + nt : nb[iprec + 2];
}

namespace {

// Gauss-Legendre nodes and weights on [0, 1], q points (Newton on P_q)
void gauss01(int q, std::vector<double>& x, std::vector<double>& w) {
  x.assign(q, 0.0);
  w.assign(q, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < q; i++) {
    double t = std::cos(pi * (i + 0.75) / (q + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= q; k++) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double pq = q == 1 ? t : p1, pm = q == 1 ? 1.0 : p0;
      dp = q * (t * pq - pm) / (t * t - 1.0);
      const double dt = pq / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - t);
    w[i] = 1.0 / ((1.0 - t * t) * dp * dp);  // 2/((1-t^2) P'^2) on [-1,1], halved
  }
}

enum Orbit { S3, S21, S111 };

// nodes of the orbits with parameters par (S3: w; S21: a, w; S111: a, b, w)
std::vector<double> orbit_nodes(const std::vector<Orbit>& orb, const std::vector<double>& par) {
  std::vector<double> r;
  size_t k = 0;
  auto put = [&](double l1, double l2, double w) { r.push_back(l1); r.push_back(l2); r.push_back(w); };
  for (Orbit o : orb) {
    if (o == S3) {
      put(1.0 / 3.0, 1.0 / 3.0, par[k]);
      k += 1;
    } else if (o == S21) {
      const double a = par[k], c = 1.0 - 2.0 * a, w = par[k + 1];
      put(a, c, w);
      put(c, a, w);
      put(a, a, w);
      k += 2;
    } else {
      const double a = par[k], b = par[k + 1], c = 1.0 - a - b, w = par[k + 2];
      put(b, c, w);
      put(c, b, w);
      put(a, c, w);
      put(c, a, w);
      put(a, b, w);
      put(b, a, w);
      k += 3;
    }
  }
  return r;
}

std::vector<double> moment_residual(const std::vector<Orbit>& orb, const std::vector<double>& par, int deg) {
  const std::vector<double> nd = orbit_nodes(orb, par);
  std::vector<double> r;
  std::vector<double> f(2 * deg + 3, 1.0);
  for (size_t i = 1; i < f.size(); i++) f[i] = f[i - 1] * (double)i;
  for (int i = 0; i <= deg; i++)
    for (int j = 0; i + j <= deg; j++) {
      double s = 0.0;
      for (size_t q = 0; q < nd.size(); q += 3) s += nd[q + 2] * std::pow(nd[q], i) * std::pow(nd[q + 1], j);
      r.push_back(s - f[i] * f[j] / f[i + j + 2]);
    }
  return r;
}

// Gauss-Newton on the (consistent) moment equations, Jacobian by central
// differences, normal equations solved by elimination with pivoting
std::vector<double> solve_rule(const std::vector<Orbit>& orb, std::vector<double> par, int deg) {
  const int np = (int)par.size();
  for (int it = 0; it < 40; it++) {
    const std::vector<double> r = moment_residual(orb, par, deg);
    const int ne = (int)r.size();
    std::vector<double> J((size_t)ne * np);
    for (int k = 0; k < np; k++) {
      std::vector<double> pp = par, pm = par;
      pp[k] += 1e-7;
      pm[k] -= 1e-7;
      const std::vector<double> rp = moment_residual(orb, pp, deg), rm = moment_residual(orb, pm, deg);
      for (int e = 0; e < ne; e++) J[(size_t)e * np + k] = (rp[e] - rm[e]) / 2e-7;
    }
    std::vector<double> A((size_t)np * np, 0.0), b(np, 0.0);
    for (int i = 0; i < np; i++) {
      for (int e = 0; e < ne; e++) b[i] -= J[(size_t)e * np + i] * r[e];
      for (int j = 0; j < np; j++)
        for (int e = 0; e < ne; e++) A[(size_t)i * np + j] += J[(size_t)e * np + i] * J[(size_t)e * np + j];
    }
    for (int c = 0; c < np; c++) {
      int p = c;
      for (int rr = c + 1; rr < np; rr++)
        if (std::fabs(A[(size_t)rr * np + c]) > std::fabs(A[(size_t)p * np + c])) p = rr;
      for (int k = 0; k < np; k++) std::swap(A[(size_t)c * np + k], A[(size_t)p * np + k]);
      std::swap(b[c], b[p]);
      for (int rr = c + 1; rr < np; rr++) {
        const double f = A[(size_t)rr * np + c] / A[(size_t)c * np + c];
        for (int k = c; k < np; k++) A[(size_t)rr * np + k] -= f * A[(size_t)c * np + k];
        b[rr] -= f * b[c];
      }
    }
    double step = 0.0;
    for (int c = np - 1; c >= 0; c--) {
      double s = b[c];
      for (int k = c + 1; k < np; k++) s -= A[(size_t)c * np + k] * b[k];
      b[c] = s / A[(size_t)c * np + c];
      par[c] += b[c];
      step = std::max(step, std::fabs(b[c]));
    }
    if (step < 1e-17) break;
  }
  return orbit_nodes(orb, par);
}

}  // namespace

std::vector<double> tri_rule(int iprec, bool helmholtz) {
  // Degree by iprec: 2 (iprec <= -1), 4 (iprec 0), 6 (iprec 1, 2) by the
  // symmetric rules (3, 6, 12 nodes); from iprec 3 on the product rule with
  // q = iprec + 2 (degrees 8, 10, 12 for iprec 3, 4, 5), one more for the
  // Helmholtz kernel, whose expansions of a triangle also carry its phase
  // e^{ik.y}.  The reference's own far-field rules are of order 2, 4, 6
  // (Laplace, src/lfmm3dtria.f:680-701) and 2, 4 (Helmholtz,
  // src/hfmm3dtria.f:730-751).  iprec 6 has no far field (every point in
  // one leaf): the smallest rule.
  if (iprec <= 2 || iprec >= 6) return tri_sym_rule(iprec <= -1 || iprec >= 6 ? 2 : iprec == 0 ? 4 : 6);
  int q = iprec + 2;
  if (helmholtz) q += 1;
  std::vector<double> x, w, r;
  gauss01(q, x, w);
  for (int i = 0; i < q; i++)
    for (int j = 0; j < q; j++) {
      const double u = x[i], v = x[j];
      r.push_back(u * (1.0 - v));
      r.push_back(u * v);
      r.push_back(w[i] * w[j] * u);
    }
  return r;
}

const std::vector<double>& tri_sym_rule(int degree) {
  static std::once_flag once;
  static std::vector<double> rules[7];
  std::call_once(once, [] {
    rules[1] = {1.0 / 3.0, 1.0 / 3.0, 0.5};
    rules[2] = orbit_nodes({S21}, {1.0 / 6.0, 1.0 / 6.0});
    rules[4] = solve_rule({S21, S21}, {0.09, 0.05, 0.45, 0.11}, 4);
    rules[3] = rules[4];
    rules[5] = solve_rule({S3, S21, S21}, {0.11, 0.10, 0.06, 0.47, 0.07}, 5);
    rules[6] = solve_rule({S21, S21, S111}, {0.48, 0.04, 0.22, 0.09, 0.02, 0.14, 0.02}, 6);
  });
  return rules[degree >= 1 && degree <= 6 ? degree : 0];
}

}  // namespace fmm3d
