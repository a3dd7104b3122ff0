// Host code of the library under AddressSanitizer and UndefinedBehaviour-
// Sanitizer (CPU only): the octree and its interaction lists
// (csrc/octree.cpp), the flattened plan (csrc/plan.hpp) and the translation
// and order tables (csrc/tables.cpp, csrc/helm_tables.cpp), on uniform,
// clustered, spherical, planar and degenerate inputs.  Besides the
// sanitizers' own checks it verifies the structural invariants of the tree
// and that every plan offset lies inside the flattened buffers.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <vector>

#include "helm_tables.hpp"
#include "octree.hpp"
#include "plan.hpp"
#include "tables.hpp"

using namespace fmm3d;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand() {  // splitmix64 -> [0, 1)
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (double)((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
}
static double nrand() { return std::sqrt(-2.0 * std::log(urand() + 1e-300)) * std::cos(6.283185307179586 * urand()); }

static std::vector<double> points(int n, int dist) {
  std::vector<double> x(3 * (size_t)n);
  double c[8][3];
  for (auto& cc : c)
    for (double& v : cc) v = 0.1 + 0.8 * urand();
  for (int i = 0; i < n; i++) {
    double* p = &x[3 * (size_t)i];
    if (dist == 0) {            // cube
      for (int d = 0; d < 3; d++) p[d] = urand();
    } else if (dist == 1) {     // 8 clusters
      const int k = (int)(8 * urand()) & 7;
      for (int d = 0; d < 3; d++) p[d] = std::fmin(1.0, std::fmax(0.0, c[k][d] + 0.02 * nrand()));
    } else if (dist == 2) {     // sphere
      double v[3] = {nrand(), nrand(), nrand()};
      const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-300;
      for (int d = 0; d < 3; d++) p[d] = v[d] / r;
    } else if (dist == 3) {     // plane
      p[0] = urand(); p[1] = urand(); p[2] = 0.25;
    } else {                    // every point at the same place
      p[0] = p[1] = p[2] = 0.5;
    }
  }
  return x;
}

static int fails = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      fails++;                                                          \
    }                                                                   \
  } while (0)

static void check_tree(const Octree& t, int ns, int nt) {
  std::vector<int> seen(ns, 0), seent(nt > 0 ? nt : 1, 0);
  for (int v : t.src_perm) { CHECK(v >= 0 && v < ns); if (v >= 0 && v < ns) seen[v]++; }
  for (int v : t.trg_perm) { CHECK(v >= 0 && v < nt); if (v >= 0 && v < nt) seent[v]++; }
  for (int i = 0; i < ns; i++) CHECK(seen[i] == 1);
  for (int i = 0; i < nt; i++) CHECK(seent[i] == 1);
  long leaf_src = 0, leaf_trg = 0;
  for (int b = 0; b < t.nboxes; b++) {
    CHECK(t.isrc[b] >= 0 && t.isrc[b] + t.nsrc[b] <= ns);
    CHECK(t.itrg[b] >= 0 && t.itrg[b] + t.ntrg[b] <= (nt > 0 ? nt : 0));
    if (t.is_leaf(b)) { leaf_src += t.nsrc[b]; leaf_trg += t.ntrg[b]; }
    for (int c = 0; c < t.nchild[b]; c++) {
      const int ch = t.child[8 * (size_t)b + c];
      CHECK(ch > b && ch < t.nboxes);
      if (ch > b && ch < t.nboxes) CHECK(t.parent[ch] == b && t.level[ch] == t.level[b] + 1);
    }
    for (int l = 0; l < 4; l++) {
      const Csr& L = t.list[l];
      for (int k = L.off[b]; k < L.off[b + 1]; k++) {
        CHECK(L.idx[k] >= 0 && L.idx[k] < t.nboxes);
        if (k > L.off[b]) CHECK(L.idx[k] > L.idx[k - 1]);
        if (l == 1 && L.idx[k] >= 0 && L.idx[k] < t.nboxes) CHECK(t.level[L.idx[k]] == t.level[b]);
      }
    }
  }
  CHECK(leaf_src == ns && leaf_trg == (nt > 0 ? nt : 0));
  // list 4 is the transpose of list 3
  long n3 = 0, n4 = 0;
  for (int b = 0; b < t.nboxes; b++) {
    n3 += t.list[2].count(b);
    n4 += t.list[3].count(b);
    for (int k = t.list[2].off[b]; k < t.list[2].off[b + 1]; k++) {
      const int j = t.list[2].idx[k];
      bool found = false;
      for (int q = t.list[3].off[j]; q < t.list[3].off[j + 1]; q++) found |= t.list[3].idx[q] == b;
      CHECK(found);
    }
  }
  CHECK(n3 == n4);
}

static void check_plan(const Octree& t) {
  Flags f{1, 1, 1, 1, t.nt > 0, t.nt > 0};
  for (int minlev = 0; minlev <= 2; minlev += 2) {
    Flat F;
    PlanOffsets O;
    build_plan(t, f, F, O, minlev);
    const size_t NI = F.I.size(), ND = F.D.size();
    const size_t nb = (size_t)t.nboxes;
    const size_t sizes[][2] = {{O.level, nb}, {O.isrc, nb}, {O.nsrc, nb}, {O.itrg, nb}, {O.ntrg, nb},
                               {O.nchild, nb}, {O.child, 8 * nb}, {O.ijk, 3 * nb}, {O.l1o, nb + 1},
                               {O.l2o, nb + 1}, {O.l3o, nb + 1}, {O.l4o, nb + 1}, {O.p2m, (size_t)O.n_p2m},
                               {O.l2p, (size_t)O.n_l2p}, {O.m2p, (size_t)O.n_m2p}, {O.p2l, (size_t)O.n_p2l},
                               {O.m2l, (size_t)O.n_m2l}, {O.p2p0, 3 * (size_t)O.n_p2p0},
                               {O.l2pi, 3 * (size_t)O.n_l2pi}, {O.sperm, (size_t)t.ns}, {O.tperm, (size_t)t.nt}};
    for (auto& s : sizes) CHECK(s[0] + s[1] <= NI);
    CHECK(O.center + 3 * nb <= ND);
    CHECK(O.inv_side + (size_t)t.nlev + 1 <= ND);
    for (int k = 0; k < O.n_p2p0; k++) {
      const int b = F.I[O.p2p0 + 3 * (size_t)k];
      CHECK(b >= -1 && b < t.nboxes);
    }
    for (auto& g : O.m2m_groups) CHECK(g.first >= 0 && g.first + g.second <= (int)(O.l2l - O.m2m + NI));
  }
}

int main() {
  const int cases[][3] = {  // ns, nt, dist
      {20000, 5000, 0}, {30000, 0, 1}, {25000, 2000, 2}, {12000, 3000, 3}, {3000, 100, 4}, {1, 0, 0}, {50, 50, 1}};
  const int nboxes[] = {40, 80, 160, 400};
  for (auto& c : cases)
    for (int nbox : nboxes) {
      std::vector<double> s = points(c[0], c[2]), tg = points(c[1], c[2]);
      Octree t;
      const int ier = build_octree(s.data(), c[0], c[1] ? tg.data() : nullptr, c[1], nbox, t);
      CHECK(ier == 0);
      if (ier) continue;
      check_tree(t, c[0], c[1]);
      check_plan(t);
      // rebuild into the same storage (reset keeps the capacity)
      t.reset();
      CHECK(build_octree(s.data(), c[0], c[1] ? tg.data() : nullptr, c[1], nbox, t) == 0);
    }
  // forests of several trees (fused small requests): the merged tree keeps
  // every tree's boxes, point ranges and lists, and its plan is in range
  {
    const int specs[][3] = {{9000, 2000, 0}, {15000, 0, 1}, {4000, 4000, 2}, {100, 0, 0}, {20000, 500, 3}};
    std::vector<Octree> trees(5);
    std::vector<const Octree*> tp;
    int so = 0, to = 0;
    std::vector<int> soff, toff;
    for (int r = 0; r < 5; r++) {
      std::vector<double> s = points(specs[r][0], specs[r][2]), tg = points(specs[r][1], specs[r][2]);
      CHECK(build_octree(s.data(), specs[r][0], specs[r][1] ? tg.data() : nullptr, specs[r][1], 80, trees[r]) == 0);
      tp.push_back(&trees[r]);
      soff.push_back(so);
      toff.push_back(to);
      so += specs[r][0];
      to += specs[r][1];
    }
    // merged twice into the same storage (a worker reuses its forest): the
    // second merge must not keep anything of the first
    Octree f;
    merge_forest(tp, f);
    merge_forest(tp, f);
    CHECK(f.ns == so && f.nt == to);
    check_tree(f, so, to);
    check_plan(f);
    CHECK(f.scale_id.size() == (size_t)f.nboxes);
    // each tree's boxes appear, level by level in their own order, with the
    // same contents (ids mapped, point ranges shifted)
    auto fid = [&](int r, int b) {
      const int l = trees[r].level[b];
      int id = f.ladd[l];
      for (int q = 0; q < r; q++)
        if (l <= trees[q].nlev) id += trees[q].ladd[l + 1] - trees[q].ladd[l];
      return id + (b - trees[r].ladd[l]);
    };
    for (int r = 0; r < 5; r++) {
      const Octree& t = trees[r];
      for (int b = 0; b < t.nboxes; b++) {
        const int g = fid(r, b);
        CHECK(g >= 0 && g < f.nboxes);
        if (g < 0 || g >= f.nboxes) continue;
        CHECK(f.nsrc[g] == t.nsrc[b] && f.ntrg[g] == t.ntrg[b]);
        CHECK(f.isrc[g] == t.isrc[b] + soff[r] && f.itrg[g] == t.itrg[b] + toff[r]);
        CHECK(f.nchild[g] == t.nchild[b] && f.level[g] == t.level[b]);
        CHECK(f.scale_inv[f.scale_id[g]] == 1.0 / t.side(t.level[b]));
        for (int c = 0; c < t.nchild[b]; c++) CHECK(f.child[8 * (size_t)g + c] == fid(r, t.child[8 * (size_t)b + c]));
        for (int k = 0; k < 4; k++) {
          CHECK(f.list[k].count(g) == t.list[k].count(b));
          if (f.list[k].count(g) != t.list[k].count(b)) continue;
          for (int q = 0; q < t.list[k].count(b); q++)
            CHECK(f.list[k].idx[f.list[k].off[g] + q] == fid(r, t.list[k].idx[t.list[k].off[b] + q]));
        }
      }
      for (int i = 0; i < t.ns; i++) CHECK(f.src_perm[soff[r] + i] == t.src_perm[i] + soff[r]);
      for (int i = 0; i < t.nt; i++) CHECK(f.trg_perm[toff[r] + i] == t.trg_perm[i] + toff[r]);
    }
    for (int l = 0; l <= f.nlev; l++)
      for (int b = f.ladd[l]; b < f.ladd[l + 1]; b++) CHECK(f.level[b] == l);
    long nb = 0;
    for (auto* t : tp) nb += t->nboxes;
    CHECK(nb == f.nboxes);
  }
  // tables of the Laplace orders of iprec -2..3 and the Helmholtz orders
  const double eps[] = {.5e0, .5e-1, .5e-2, .5e-3, .5e-6, .5e-9};
  for (double e : eps) {
    const int p = laplace_nterms(e);
    HostTables H;
    build_tables(p, e, H);
    CHECK(!H.mf_rot.empty() && !H.mf_z.empty());
    CHECK((long)H.mf_rot.size() == 2 * H.mf_rot_back);
    CHECK((long)H.mf_z.size() == H.mf_z_size * H.ndist);
  }
  // Helmholtz tables: orders per level, list-2 tables, coaxial operators
  const double zks[][2] = {{20.0, 0.0}, {1.0, 0.1}, {30.0, 0.0}, {4.0, 2.0}};
  for (auto& z : zks)
    for (double e : {.5e-3, .5e-6, .5e-9}) {
      HelmHostTables H;
      const bool ok = build_helm_tables(cplx(z[0], z[1]), 1.0, 5, e, H);
      if (!ok) continue;  // an order beyond kMaxHelmOrder: reported, not built
      CHECK((int)H.plev.size() == 6);
      for (int l = 2; l <= 5; l++) CHECK(H.plev[l] > 0 && H.plev[l] <= H.pmax);
      for (auto& d : H.m2l) CHECK(d.base >= 0 && 2 * (size_t)d.base <= H.z.size());
    }
  std::printf("host_check: %d failed checks\n", fails);
  return fails ? 1 : 0;
}
