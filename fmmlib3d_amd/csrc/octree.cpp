// Host octree builder; see octree.hpp.  The tree is identical, box for box, to
// the one d3tstrcr builds (src/d3tstrcr.f), but the construction is ours:
// up to four levels are resolved per pass over the points (octant digit
// codes, a histogram that decides every split, one stable counting sort),
// child ids are assigned breadth first, and the interaction lists are
// derived from integer box coordinates (colleagues from the parent's
// colleagues' children) instead of d3tstrcr's floating-point corner tests and
// linked-list store.
#include "octree.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fmm3d {

double Octree::side(int l) const { return size / std::ldexp(1.0, l); }

void Octree::reset() {
  nboxes = nlev = ns = nt = 0;
  center[0] = center[1] = center[2] = 0.0;
  size = 0.0;
  for (auto* v : {&level, &parent, &nchild, &child, &ijk, &isrc, &nsrc, &itrg, &ntrg, &ladd, &src_perm, &trg_perm})
    v->clear();
  bcenter.clear();
  for (auto& c : list) {
    c.off.clear();
    c.idx.clear();
  }
  scale_id.clear();
  scale_inv.clear();
}

void merge_forest(const std::vector<const Octree*>& trees, Octree& f) {
  f.reset();
  const int nt = (int)trees.size();
  int maxlev = 0, nb = 0;
  for (const Octree* t : trees) {
    maxlev = std::max(maxlev, t->nlev);
    nb += t->nboxes;
    f.ns += t->ns;
    f.nt += t->nt;
  }
  f.nboxes = nb;
  f.nlev = maxlev;
  // new ids: level-major, then tree, then the tree's own (breadth-first) order
  std::vector<std::vector<int>> nid(nt);
  std::vector<int> so(nt), to(nt), sbase(nt);
  for (int r = 0, s0 = 0, t0 = 0, b0 = 0; r < nt; r++) {
    nid[r].assign(trees[r]->nboxes, -1);
    so[r] = s0; to[r] = t0; sbase[r] = b0;
    s0 += trees[r]->ns; t0 += trees[r]->nt; b0 += trees[r]->nlev + 1;
  }
  f.ladd.assign(maxlev + 2, 0);
  std::vector<std::pair<int, int>> order;  // (tree, old box) by new id
  order.reserve(nb);
  for (int l = 0; l <= maxlev; l++) {
    f.ladd[l] = (int)order.size();
    for (int r = 0; r < nt; r++) {
      const Octree& t = *trees[r];
      if (l > t.nlev) continue;
      for (int b = t.ladd[l]; b < t.ladd[l + 1]; b++) {
        nid[r][b] = (int)order.size();
        order.push_back({r, b});
      }
    }
  }
  f.ladd[maxlev + 1] = nb;
  f.level.resize(nb); f.parent.resize(nb); f.nchild.resize(nb); f.child.resize(8 * (size_t)nb);
  f.ijk.resize(3 * (size_t)nb); f.isrc.resize(nb); f.nsrc.resize(nb); f.itrg.resize(nb); f.ntrg.resize(nb);
  f.bcenter.resize(3 * (size_t)nb); f.scale_id.resize(nb);
  for (int b = 0; b < nb; b++) {
    const int r = order[b].first, o = order[b].second;
    const Octree& t = *trees[r];
    f.level[b] = t.level[o];
    f.parent[b] = t.parent[o] < 0 ? t.parent[o] : nid[r][t.parent[o]];
    f.nchild[b] = t.nchild[o];
    for (int c = 0; c < 8; c++) {
      const int ch = t.child[8 * (size_t)o + c];
      f.child[8 * (size_t)b + c] = ch < 0 ? ch : nid[r][ch];
    }
    for (int d = 0; d < 3; d++) {
      f.ijk[3 * (size_t)b + d] = t.ijk[3 * (size_t)o + d];
      f.bcenter[3 * (size_t)b + d] = t.bcenter[3 * (size_t)o + d];
    }
    f.isrc[b] = t.isrc[o] + so[r];
    f.nsrc[b] = t.nsrc[o];
    f.itrg[b] = t.itrg[o] + to[r];
    f.ntrg[b] = t.ntrg[o];
    f.scale_id[b] = sbase[r] + t.level[o];
  }
  f.scale_inv.clear();
  for (const Octree* t : trees)
    for (int l = 0; l <= t->nlev; l++) f.scale_inv.push_back(1.0 / t->side(l));
  f.src_perm.clear();
  f.trg_perm.clear();
  for (int r = 0; r < nt; r++) {
    for (int v : trees[r]->src_perm) f.src_perm.push_back(v + so[r]);
    for (int v : trees[r]->trg_perm) f.trg_perm.push_back(v + to[r]);
  }
  for (int k = 0; k < 4; k++) {
    Csr& L = f.list[k];
    L.off.assign(nb + 1, 0);
    L.idx.clear();
    for (int b = 0; b < nb; b++) {
      const int r = order[b].first, o = order[b].second;
      const Csr& S = trees[r]->list[k];
      for (int q = S.off[o]; q < S.off[o + 1]; q++) L.idx.push_back(nid[r][S.idx[q]]);
      L.off[b + 1] = (int)L.idx.size();
    }
  }
}

namespace {

// Centre of box (i,j,k) at `level`: the same arithmetic as d3tcentf
// (d3tstrcr.f:1675-1700), kept free of FMA contraction so that ties at the
// splitting planes fall the same way.
#pragma clang fp contract(off)
inline double box_center_1d(double c0, double size, int level, int i) {
  double side = size / std::ldexp(1.0, level);
  double side2 = side / 2;
  double x0 = c0 - size / 2;
  return x0 + (i - 1) * side + side2;
}

// One point in the partition workspace: position, original index, and the
// octant digits below the box being split.
struct Pt {
  double x, y, z;
  int idx;
  uint32_t code;
};

// Persistent per-thread workspace (no page faults after the first call).
struct Work {
  std::vector<Pt> a, b;
};
thread_local Work g_src, g_trg;

// Stable counting sort of a[lo, lo+n) into b[lo, lo+n) by key(p) < nbins,
// split over `T` threads (chunk histograms, bin-major prefix, stable scatter).
template <typename Key>
void stable_bin_sort(const Pt* a, Pt* b, int lo, int n, int nbins, int T, const Key& key) {
  if (n <= 0) return;
  if (T <= 1 || n < (1 << 15)) {
    std::vector<int> off(nbins + 1, 0);
    for (int i = lo; i < lo + n; i++) off[key(a[i]) + 1]++;
    off[0] = lo;
    for (int k = 0; k < nbins; k++) off[k + 1] += off[k];
    for (int i = lo; i < lo + n; i++) b[off[key(a[i])]++] = a[i];
    return;
  }
  std::vector<int> h((size_t)T * nbins, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const int c0 = lo + (int)((long)n * t / T), c1 = lo + (int)((long)n * (t + 1) / T);
    int* ht = &h[(size_t)t * nbins];
    for (int i = c0; i < c1; i++) ht[key(a[i])]++;
#pragma omp barrier
#pragma omp single
    {
      int s = lo;
      for (int k = 0; k < nbins; k++)
        for (int tt = 0; tt < T; tt++) {
          int v = h[(size_t)tt * nbins + k];
          h[(size_t)tt * nbins + k] = s;
          s += v;
        }
    }
    for (int i = c0; i < c1; i++) b[ht[key(a[i])]++] = a[i];
  }
}

void push_box(Octree& t, int level, int i, int j, int k, int par, int is, int n_s, int it, int n_t) {
  t.level.push_back(level); t.parent.push_back(par); t.nchild.push_back(0);
  for (int q = 0; q < 8; q++) t.child.push_back(-1);
  t.ijk.push_back(i); t.ijk.push_back(j); t.ijk.push_back(k);
  t.isrc.push_back(is); t.nsrc.push_back(n_s); t.itrg.push_back(it); t.ntrg.push_back(n_t);
  t.bcenter.push_back(box_center_1d(t.center[0], t.size, level, i));
  t.bcenter.push_back(box_center_1d(t.center[1], t.size, level, j));
  t.bcenter.push_back(box_center_1d(t.center[2], t.size, level, k));
}

// closed-interval overlap of two boxes given by level and 1-based (i,j,k)
inline bool touching(const Octree& t, int a, int b) {
  int la = t.level[a], lb = t.level[b];
  int L = std::max(la, lb);
  for (int d = 0; d < 3; d++) {
    long a0 = (long)(t.ijk[3 * a + d] - 1) << (L - la), a1 = (long)t.ijk[3 * a + d] << (L - la);
    long b0 = (long)(t.ijk[3 * b + d] - 1) << (L - lb), b1 = (long)t.ijk[3 * b + d] << (L - lb);
    if (a0 > b1 || b0 > a1) return false;
  }
  return true;
}

void flatten(std::vector<std::vector<int>>& per, int nb, Csr& c) {
  c.off.assign(nb + 1, 0);
  for (int b = 0; b < nb; b++) c.off[b + 1] = c.off[b] + (int)per[b].size();
  c.idx.resize(c.off[nb]);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++) {
    std::sort(per[b].begin(), per[b].end());
    std::copy(per[b].begin(), per[b].end(), c.idx.begin() + c.off[b]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Box logic shared by the host and device builders
// ---------------------------------------------------------------------------

void tree_root(Octree& t, const double lo[3], const double hi[3], int ns, int nt, int nbox,
               std::vector<int>& frontier) {
  // bounding cube of sources and targets (d3tstrcr.f:1357-1395)
  double size = hi[0] - lo[0];
  if (hi[1] - lo[1] > size) size = hi[1] - lo[1];
  if (hi[2] - lo[2] > size) size = hi[2] - lo[2];
  for (int d = 0; d < 3; d++) t.center[d] = (lo[d] + hi[d]) / 2;
  t.size = size;
  t.ns = ns;
  t.nt = nt;
  push_box(t, 0, 1, 1, 1, -1, 0, ns, 0, nt);
  t.ladd = {0, 1};
  frontier.clear();
  if (ns > nbox || nt > nbox) frontier.push_back(0);
}

void tree_pass_setup(const Octree& t, int L, TreePass& ps) {
  const int nf = (int)ps.frontier.size();
  // D levels per pass, with nf * 8^D <= 2^18: the histograms, node ids and
  // split flags of a pass (copied back and scanned on the host) scale with it
  int D = 4;
  while (D > 1 && (long)nf << (3 * D) > (1L << 18)) D--;
  D = std::min(D, kTreeMaxLevel - L);
  ps.L = L;
  ps.D = ps.Dfull = D;
  ps.NB = 1 << (3 * D);
  ps.CT = (1 << D) * 2;  // sum_{d<D} 2^d < 2^D, padded
  ps.ctab.assign((size_t)nf * 3 * D * ps.CT, 0.0);
  for (int f = 0; f < nf; f++) {
    const int b = ps.frontier[f];
    for (int dim = 0; dim < 3; dim++)
      for (int d = 0; d < D; d++)
        for (int ii = 0; ii < (1 << d); ii++)
          ps.ctab[(((size_t)f * 3 + dim) * D + d) * ps.CT + ii] =
              box_center_1d(t.center[dim], t.size, L + d, (t.ijk[3 * b + dim] - 1) * (1 << d) + 1 + ii);
  }
  ps.hs.assign((size_t)nf * (ps.NB + 1), 0);
  ps.ht.assign((size_t)nf * (ps.NB + 1), 0);
}

int tree_pos(int d) {
  int o = 0;
  for (int e = 1; e < d; e++) o += 1 << (3 * e);
  return o;
}

void tree_pass_expand(Octree& t, int nbox, TreePass& ps, std::vector<int>& next) {
  // sub-boxes, level by level across all frontier boxes (breadth-first ids);
  // node (d, prefix) of frontier box f: id and whether it splits further
  const int nf = (int)ps.frontier.size();
  int D = ps.D;
  const int L = ps.L;
  ps.NS = ps.NB * 2;  // sum_{d<=D} 8^d < 2*8^D, stored at [pos(d) + prefix]
  const int NS = ps.NS;
  std::vector<int> node((size_t)nf * NS, -1);
  ps.splits.assign((size_t)nf * NS, 0);
  next.clear();
  for (int d = 1; d <= D; d++) {
    const int sh = 3 * (D - d), pd = tree_pos(d), pp = tree_pos(d - 1);
    for (int f = 0; f < nf; f++) {
      const int fb = ps.frontier[f];
      const int* hsf = &ps.hs[(size_t)f * (ps.NB + 1)];
      const int* htf = &ps.ht[(size_t)f * (ps.NB + 1)];
      const int nparents = d == 1 ? 1 : 1 << (3 * (d - 1));
      for (int P0 = 0; P0 < nparents; P0++) {
        int par;
        if (d == 1) par = fb;
        else {
          if (!ps.splits[(size_t)f * NS + pp + P0]) continue;
          par = node[(size_t)f * NS + pp + P0];
        }
        const int bi = t.ijk[3 * par], bj = t.ijk[3 * par + 1], bk = t.ijk[3 * par + 2];
        for (int o = 0; o < 8; o++) {
          const int P = P0 * 8 + o;
          const int s0 = hsf[P << sh], s1 = hsf[(P + 1) << sh];
          const int u0 = htf[P << sh], u1 = htf[(P + 1) << sh];
          const int a = s1 - s0, e = u1 - u0;
          if (a == 0 && e == 0) continue;
          const int id = (int)t.level.size();
          push_box(t, L + d, 2 * (bi - 1) + 1 + (o & 1), 2 * (bj - 1) + 1 + ((o >> 1) & 1),
                   2 * (bk - 1) + 1 + ((o >> 2) & 1), par, t.isrc[fb] + s0, a, t.itrg[fb] + u0, e);
          t.child[8 * par + t.nchild[par]++] = id;
          node[(size_t)f * NS + pd + P] = id;
          if (a > nbox || e > nbox) {
            if (d < D) ps.splits[(size_t)f * NS + pd + P] = 1;
            else next.push_back(id);
          }
        }
      }
    }
    t.ladd.push_back((int)t.level.size());
    t.nlev = L + d;
    bool more = false;
    for (int f = 0; f < nf && !more; f++)
      for (int P = 0; P < (1 << (3 * d)) && !more; P++) more = ps.splits[(size_t)f * NS + pd + P] != 0;
    if (!more && d < D) {
      // nothing below this depth: the codes are truncated here for every point
      D = d;
      break;
    }
  }
  ps.Dk = D;
  for (int d = 0; d <= D; d++) ps.posd[d] = tree_pos(d);
}

// ---------------------------------------------------------------------------
// Interaction lists from the box structure
// ---------------------------------------------------------------------------
namespace {
// per-box scratch lists kept between calls (no allocation once grown); the
// calling thread's copies, which the OpenMP workers reach through references
struct ListScratch {
  std::vector<std::vector<int>> coll, l1, l2, l3, l4;
  std::vector<std::vector<std::pair<int, int>>> mirror;
};
thread_local ListScratch g_lists;
}  // namespace

void tree_lists_near(Octree& t) {
  const int nb = t.nboxes;
  const int nthreads = omp_get_max_threads();
  ListScratch& S = g_lists;
  auto& coll = S.coll;
  auto& l1 = S.l1;
  auto& l3 = S.l3;
  auto& mirror = S.mirror;
  for (auto* v : {&S.coll, &S.l1, &S.l2, &S.l3, &S.l4})
    if ((int)v->size() < nb) v->resize(nb);
  if ((int)mirror.size() < nthreads) mirror.resize(nthreads);
  for (auto& m : mirror) m.clear();
  auto& l2s = S.l2;
  auto& l4s = S.l4;
  // one parallel region (a fork/join costs more than a level's work here)
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (int b = 0; b < nb; b++) {
      coll[b].clear(); l1[b].clear(); l2s[b].clear(); l3[b].clear(); l4s[b].clear();
    }
  // colleagues, level by level: the same-level boxes touching b are children
  // of b's parent or of the parent's colleagues
  for (int l = 1; l <= t.nlev; l++) {
#pragma omp for schedule(static)
    for (int b = t.ladd[l]; b < t.ladd[l + 1]; b++) {
      const int par = t.parent[b];
      auto visit = [&](int q) {
        for (int c = 0; c < t.nchild[q]; c++) {
          const int kid = t.child[8 * q + c];
          if (kid == b) continue;
          int d = 0;
          for (int x = 0; x < 3; x++) d = std::max(d, std::abs(t.ijk[3 * kid + x] - t.ijk[3 * b + x]));
          if (d <= 1) coll[b].push_back(kid);
        }
      };
      visit(par);
      for (int q : coll[par]) visit(q);
      std::sort(coll[b].begin(), coll[b].end());
    }
  }
  // lists 1 and 3 for leaves below the root (d3tlst31, d3tstrcr.f:912-1069);
  // the coarse-to-fine list-1 entries are mirrored afterwards.
#pragma omp for schedule(dynamic, 16)
  for (int b = 1; b < nb; b++) {
    if (t.nchild[b]) continue;
    int tid = omp_get_thread_num();
    thread_local std::vector<int> stack;
    stack.clear();
    for (int j0 : coll[b]) {
      stack.push_back(j0);
      while (!stack.empty()) {
        int j = stack.back();
        stack.pop_back();
        if (!touching(t, b, j)) { l3[b].push_back(j); continue; }
        if (t.nchild[j] == 0) {
          l1[b].push_back(j);
          if (t.level[j] != t.level[b]) mirror[tid].push_back({j, b});
          continue;
        }
        for (int c = 0; c < t.nchild[j]; c++) stack.push_back(t.child[8 * j + c]);
      }
    }
  }
  }  // parallel
  for (int th = 0; th < nthreads; th++)
    for (auto& pr : mirror[th]) l1[pr.first].push_back(pr.second);
  flatten(l1, nb, t.list[0]);
  flatten(l3, nb, t.list[2]);
}

void tree_lists_far(Octree& t) {
  const int nb = t.nboxes;
  ListScratch& S = g_lists;
  auto& coll = S.coll;
  auto& l2 = S.l2;
  auto& l4 = S.l4;
  // list 2: children of the parent's colleagues (and of the parent) that do
  // not touch the box (d3tstrcr.f:768-821)
#pragma omp parallel for schedule(dynamic, 64)
  for (int b = 1; b < nb; b++) {
    int p = t.parent[b];
    auto visit = [&](int q) {
      for (int c = 0; c < t.nchild[q]; c++) {
        int kid = t.child[8 * q + c];
        if (kid == b) continue;
        int d = 0;
        for (int x = 0; x < 3; x++) d = std::max(d, std::abs(t.ijk[3 * kid + x] - t.ijk[3 * b + x]));
        if (d > 1) l2[b].push_back(kid);
      }
    };
    visit(p);
    for (int q : coll[p]) visit(q);
  }
  // list 4 is the transpose of list 3
  for (int b = 0; b < nb; b++)
    for (int k = t.list[2].off[b]; k < t.list[2].off[b + 1]; k++) l4[t.list[2].idx[k]].push_back(b);
  flatten(l2, nb, t.list[1]);
  flatten(l4, nb, t.list[3]);
}

void tree_lists(Octree& t) {
  tree_lists_near(t);
  tree_lists_far(t);
}

// ---------------------------------------------------------------------------
// Host builder: the points live in host memory
// ---------------------------------------------------------------------------
int build_octree(const double* src, int ns, const double* trg, int nt, int nbox, Octree& t, bool lists) {
  t.reset();
  if (ns < 1) return 128;
  if (nt < 0) nt = 0;
  double lo[3] = {src[0], src[1], src[2]}, hi[3] = {src[0], src[1], src[2]};
  auto scan = [&](const double* p, int n) {
    double l0 = lo[0], l1 = lo[1], l2 = lo[2], h0 = hi[0], h1 = hi[1], h2 = hi[2];
#pragma omp parallel for reduction(min : l0, l1, l2) reduction(max : h0, h1, h2)
    for (int i = 0; i < n; i++) {
      l0 = std::min(l0, p[3 * i]); h0 = std::max(h0, p[3 * i]);
      l1 = std::min(l1, p[3 * i + 1]); h1 = std::max(h1, p[3 * i + 1]);
      l2 = std::min(l2, p[3 * i + 2]); h2 = std::max(h2, p[3 * i + 2]);
    }
    lo[0] = l0; lo[1] = l1; lo[2] = l2; hi[0] = h0; hi[1] = h1; hi[2] = h2;
  };
  scan(src, ns);
  if (nt > 0) scan(trg, nt);
  std::vector<int> frontier;
  tree_root(t, lo, hi, ns, nt, nbox, frontier);

  // workspace: positions with their original index
  Work& WS = g_src;
  Work& WT = g_trg;
  if ((int)WS.a.size() < ns) { WS.a.resize(ns); WS.b.resize(ns); }
  if ((int)WT.a.size() < nt) { WT.a.resize(nt); WT.b.resize(nt); }
  Pt* SA = WS.a.data();
  Pt* SB = WS.b.data();
  Pt* TA = WT.a.data();
  Pt* TB = WT.b.data();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ns; i++) SA[i] = Pt{src[3 * i], src[3 * i + 1], src[3 * i + 2], i, 0u};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nt; i++) TA[i] = Pt{trg[3 * i], trg[3 * i + 1], trg[3 * i + 2], i, 0u};

  const int nthreads = omp_get_max_threads();
  TreePass ps;
  std::vector<int> next;
  int L = 0;  // level of the frontier boxes
  // Each pass resolves up to D levels below the frontier at once: every point
  // gets the D octant digits of its path (the same centre arithmetic and
  // strict ">" test as one split per level), a histogram of the digit codes
  // gives every sub-box's counts and hence which sub-boxes split, and one
  // stable counting sort by the code truncated at the point's leaf puts the
  // points in the order D successive stable 8-way partitions would.
  while (!frontier.empty() && L < kTreeMaxLevel) {
    ps.frontier = frontier;
    tree_pass_setup(t, L, ps);
    const int nf = (int)frontier.size();
    const int D = ps.D, NB = ps.NB, CT = ps.CT;
    auto code_of = [&](const Pt& p, int f) -> uint32_t {
      const double* cx = &ps.ctab[(((size_t)f * 3 + 0) * D) * CT];
      const double* cy = &ps.ctab[(((size_t)f * 3 + 1) * D) * CT];
      const double* cz = &ps.ctab[(((size_t)f * 3 + 2) * D) * CT];
      uint32_t code = 0;
      int ii = 0, jj = 0, kk = 0;
      for (int d = 0; d < D; d++) {
        const int o = (p.x > cx[d * CT + ii]) | ((p.y > cy[d * CT + jj]) << 1) | ((p.z > cz[d * CT + kk]) << 2);
        code = (code << 3) | (uint32_t)o;
        ii = 2 * ii + (o & 1);
        jj = 2 * jj + ((o >> 1) & 1);
        kk = 2 * kk + ((o >> 2) & 1);
      }
      return code;
    };
    // digit codes and per-box histograms (exclusive prefix sums)
    auto codes_and_hist = [&](Pt* P, int lo, int n, int f, int* h, bool par) {
      if (par) {
        std::vector<int> part((size_t)nthreads * NB, 0);
#pragma omp parallel num_threads(nthreads)
        {
          const int th = omp_get_thread_num();
          const int c0 = lo + (int)((long)n * th / nthreads), c1 = lo + (int)((long)n * (th + 1) / nthreads);
          int* pt = &part[(size_t)th * NB];
          for (int i = c0; i < c1; i++) {
            P[i].code = code_of(P[i], f);
            pt[P[i].code]++;
          }
        }
        for (int th = 0; th < nthreads; th++)
          for (int k = 0; k < NB; k++) h[k + 1] += part[(size_t)th * NB + k];
      } else {
        for (int i = lo; i < lo + n; i++) {
          P[i].code = code_of(P[i], f);
          h[P[i].code + 1]++;
        }
      }
      for (int k = 0; k < NB; k++) h[k + 1] += h[k];
    };
    const bool few = nf < 2 * nthreads;
    if (few) {
      for (int f = 0; f < nf; f++) {
        const int b = frontier[f];
        codes_and_hist(SA, t.isrc[b], t.nsrc[b], f, &ps.hs[(size_t)f * (NB + 1)], t.nsrc[b] > (1 << 15));
        codes_and_hist(TA, t.itrg[b], t.ntrg[b], f, &ps.ht[(size_t)f * (NB + 1)], t.ntrg[b] > (1 << 15));
      }
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int f = 0; f < nf; f++) {
        const int b = frontier[f];
        codes_and_hist(SA, t.isrc[b], t.nsrc[b], f, &ps.hs[(size_t)f * (NB + 1)], false);
        codes_and_hist(TA, t.itrg[b], t.ntrg[b], f, &ps.ht[(size_t)f * (NB + 1)], false);
      }
    }
    tree_pass_expand(t, nbox, ps, next);
    // stable sort of each frontier box's points by the code truncated at
    // the depth where the point's box stops splitting
    const int NBk = 1 << (3 * ps.Dk);
    auto sort_box = [&](Pt* A, Pt* Bf, int lo, int n, int f, int T) {
      stable_bin_sort(A, Bf, lo, n, NBk, T, [&](const Pt& p) { return (int)tree_eff_code(ps, f, p.code); });
      std::memcpy(A + lo, Bf + lo, sizeof(Pt) * (size_t)n);
    };
    if (nf == 1 && frontier[0] == 0) {
      // the root: every point moves, so sort into the second buffer and swap
      stable_bin_sort(SA, SB, 0, ns, NBk, nthreads, [&](const Pt& p) { return (int)tree_eff_code(ps, 0, p.code); });
      stable_bin_sort(TA, TB, 0, nt, NBk, nthreads, [&](const Pt& p) { return (int)tree_eff_code(ps, 0, p.code); });
      std::swap(SA, SB);
      std::swap(TA, TB);
    } else if (few) {
      for (int f = 0; f < nf; f++) {
        const int b = frontier[f];
        sort_box(SA, SB, t.isrc[b], t.nsrc[b], f, nthreads);
        sort_box(TA, TB, t.itrg[b], t.ntrg[b], f, nthreads);
      }
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int f = 0; f < nf; f++) {
        const int b = frontier[f];
        sort_box(SA, SB, t.isrc[b], t.nsrc[b], f, 1);
        sort_box(TA, TB, t.itrg[b], t.ntrg[b], f, 1);
      }
    }
    frontier.swap(next);
    L += ps.Dk;
  }
  t.nboxes = (int)t.level.size();
  t.src_perm.resize(ns);
  t.trg_perm.resize(nt);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ns; i++) t.src_perm[i] = SA[i].idx;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nt; i++) t.trg_perm[i] = TA[i].idx;
  if (lists) tree_lists(t);
  return 0;
}

}  // namespace fmm3d
