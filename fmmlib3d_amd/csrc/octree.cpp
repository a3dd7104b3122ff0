// Host octree builder; see octree.hpp.  The tree is identical, box for box, to
// the one d3tstrcr builds (src/d3tstrcr.f), but the construction is ours:
// every level is split in one parallel pass over the boxes that exceed nbox
// (each box does a stable 8-way counting partition of its own point range,
// large boxes split their range across threads), child ids are assigned by a
// prefix over the level, and the interaction lists are derived from integer
// box coordinates and a per-level hash of (i,j,k) instead of d3tstrcr's
// floating-point corner tests and linked-list store.
#include "octree.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace fmm3d {

double Octree::side(int l) const { return size / std::ldexp(1.0, l); }

namespace {

struct PointSet {
  std::vector<double> x, y, z;
  std::vector<int> perm;
  std::vector<double> tx, ty, tz;
  std::vector<int> tperm;
  std::vector<uint8_t> oct;
  void init(const double* p, int n) {
    x.resize(n); y.resize(n); z.resize(n); perm.resize(n);
    tx.resize(n); ty.resize(n); tz.resize(n); tperm.resize(n); oct.resize(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      x[i] = p[3 * i]; y[i] = p[3 * i + 1]; z[i] = p[3 * i + 2]; perm[i] = i;
    }
  }
};

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

inline uint8_t octant(double x, double y, double z, const double* c) {
  return (uint8_t)((x > c[0]) | ((y > c[1]) << 1) | ((z > c[2]) << 2));
}

// Stable 8-way partition of [a, a+n) by octant; counts returned in cnt.
void partition_serial(PointSet& P, int a, int n, const double* c, int* cnt) {
  for (int k = 0; k < 8; k++) cnt[k] = 0;
  for (int i = a; i < a + n; i++) {
    uint8_t o = octant(P.x[i], P.y[i], P.z[i], c);
    P.oct[i] = o;
    cnt[o]++;
  }
  int off[8], s = a;
  for (int k = 0; k < 8; k++) { off[k] = s; s += cnt[k]; }
  for (int i = a; i < a + n; i++) {
    int d = off[P.oct[i]]++;
    P.tx[d] = P.x[i]; P.ty[d] = P.y[i]; P.tz[d] = P.z[i]; P.tperm[d] = P.perm[i];
  }
  std::memcpy(&P.x[a], &P.tx[a], sizeof(double) * n);
  std::memcpy(&P.y[a], &P.ty[a], sizeof(double) * n);
  std::memcpy(&P.z[a], &P.tz[a], sizeof(double) * n);
  std::memcpy(&P.perm[a], &P.tperm[a], sizeof(int) * n);
}

// Same result as partition_serial, with the range split across threads:
// per-chunk counts, an (octant, chunk)-major prefix, then a stable scatter.
void partition_parallel(PointSet& P, int a, int n, const double* c, int* cnt) {
  int T = omp_get_max_threads();
  std::vector<int> counts(8 * T, 0);
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    int lo = a + (int)((long)n * t / T), hi = a + (int)((long)n * (t + 1) / T);
    int* ct = &counts[8 * t];
    for (int i = lo; i < hi; i++) {
      uint8_t o = octant(P.x[i], P.y[i], P.z[i], c);
      P.oct[i] = o;
      ct[o]++;
    }
#pragma omp barrier
#pragma omp single
    {
      int s = a;
      for (int k = 0; k < 8; k++)
        for (int tt = 0; tt < T; tt++) {
          int v = counts[8 * tt + k];
          counts[8 * tt + k] = s;
          s += v;
        }
    }
    int off[8];
    for (int k = 0; k < 8; k++) off[k] = ct[k];
    for (int i = lo; i < hi; i++) {
      int d = off[P.oct[i]]++;
      P.tx[d] = P.x[i]; P.ty[d] = P.y[i]; P.tz[d] = P.z[i]; P.tperm[d] = P.perm[i];
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (int i = a; i < a + n; i++) {
      P.x[i] = P.tx[i]; P.y[i] = P.ty[i]; P.z[i] = P.tz[i]; P.perm[i] = P.tperm[i];
    }
  }
  for (int k = 0; k < 8; k++) cnt[k] = 0;
  for (int i = a; i < a + n; i++) cnt[P.oct[i]]++;
}

inline uint64_t ijk_key(int i, int j, int k) {
  return ((uint64_t)(uint32_t)i << 42) | ((uint64_t)(uint32_t)j << 21) | (uint64_t)(uint32_t)k;
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

void flatten(std::vector<std::vector<int>>& per, Csr& c) {
  int nb = (int)per.size();
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

int build_octree(const double* src, int ns, const double* trg, int nt, int nbox, Octree& t) {
  t = Octree();
  if (ns < 1) return 128;
  if (nt < 0) nt = 0;
  // bounding cube of sources and targets (d3tstrcr.f:1357-1395)
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
  double size = hi[0] - lo[0];
  if (hi[1] - lo[1] > size) size = hi[1] - lo[1];
  if (hi[2] - lo[2] > size) size = hi[2] - lo[2];
  for (int d = 0; d < 3; d++) t.center[d] = (lo[d] + hi[d]) / 2;
  t.size = size;

  PointSet S, T;
  S.init(src, ns);
  T.init(trg, nt);

  auto push_box = [&](int level, int i, int j, int k, int par, int is, int n_s, int it, int n_t) {
    t.level.push_back(level); t.parent.push_back(par); t.nchild.push_back(0);
    for (int q = 0; q < 8; q++) t.child.push_back(-1);
    t.ijk.push_back(i); t.ijk.push_back(j); t.ijk.push_back(k);
    t.isrc.push_back(is); t.nsrc.push_back(n_s); t.itrg.push_back(it); t.ntrg.push_back(n_t);
    t.bcenter.push_back(box_center_1d(t.center[0], size, level, i));
    t.bcenter.push_back(box_center_1d(t.center[1], size, level, j));
    t.bcenter.push_back(box_center_1d(t.center[2], size, level, k));
  };
  push_box(0, 1, 1, 1, -1, 0, ns, 0, nt);
  t.ladd = {0, 1};
  int nthreads = omp_get_max_threads();
  for (int level = 0; level < 100; level++) {
    int first = t.ladd[level], last = t.ladd[level + 1];
    std::vector<int> split;
    for (int b = first; b < last; b++)
      if (t.nsrc[b] > nbox || t.ntrg[b] > nbox) split.push_back(b);
    if (split.empty()) break;
    int ns_split = (int)split.size();
    std::vector<int> cs(8 * ns_split), ct(8 * ns_split);
    auto do_box = [&](int q, bool par) {
      int b = split[q];
      const double* c = &t.bcenter[3 * b];
      if (par) {
        partition_parallel(S, t.isrc[b], t.nsrc[b], c, &cs[8 * q]);
        partition_parallel(T, t.itrg[b], t.ntrg[b], c, &ct[8 * q]);
      } else {
        partition_serial(S, t.isrc[b], t.nsrc[b], c, &cs[8 * q]);
        partition_serial(T, t.itrg[b], t.ntrg[b], c, &ct[8 * q]);
      }
    };
    if (ns_split < nthreads) {
      for (int q = 0; q < ns_split; q++) {
        bool big = (t.nsrc[split[q]] + t.ntrg[split[q]]) > (1 << 16);
        do_box(q, big);
      }
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int q = 0; q < ns_split; q++) do_box(q, false);
    }
    // children in (box, octant) order; empty ones dropped (d3tstrcr.f:1529)
    for (int q = 0; q < ns_split; q++) {
      int b = split[q];
      int os = t.isrc[b], ot = t.itrg[b];
      int bi = t.ijk[3 * b], bj = t.ijk[3 * b + 1], bk = t.ijk[3 * b + 2];
      for (int o = 0; o < 8; o++) {
        int a = cs[8 * q + o], e = ct[8 * q + o];
        if (a == 0 && e == 0) continue;
        int id = (int)t.level.size();
        push_box(level + 1, 2 * (bi - 1) + 1 + (o & 1), 2 * (bj - 1) + 1 + ((o >> 1) & 1),
                 2 * (bk - 1) + 1 + ((o >> 2) & 1), b, os, a, ot, e);
        t.child[8 * b + t.nchild[b]++] = id;
        os += a;
        ot += e;
      }
    }
    t.ladd.push_back((int)t.level.size());
    t.nlev = level + 1;
  }
  t.nboxes = (int)t.level.size();
  t.src_perm.swap(S.perm);
  t.trg_perm.swap(T.perm);
  const int nb = t.nboxes;

  // colleagues through a per-level hash of (i,j,k)
  std::vector<std::unordered_map<uint64_t, int>> H(t.nlev + 1);
  for (int l = 0; l <= t.nlev; l++) {
    H[l].reserve(2 * (t.ladd[l + 1] - t.ladd[l]));
    for (int b = t.ladd[l]; b < t.ladd[l + 1]; b++)
      H[l][ijk_key(t.ijk[3 * b], t.ijk[3 * b + 1], t.ijk[3 * b + 2])] = b;
  }
  std::vector<std::vector<int>> coll(nb);
#pragma omp parallel for schedule(dynamic, 64)
  for (int b = 1; b < nb; b++) {
    int l = t.level[b];
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++)
        for (int dk = -1; dk <= 1; dk++) {
          if (!di && !dj && !dk) continue;
          auto it = H[l].find(ijk_key(t.ijk[3 * b] + di, t.ijk[3 * b + 1] + dj, t.ijk[3 * b + 2] + dk));
          if (it != H[l].end()) coll[b].push_back(it->second);
        }
  }
  // list 2: children of the parent's colleagues (and of the parent) that do
  // not touch the box (d3tstrcr.f:768-821)
  std::vector<std::vector<int>> l2(nb), l1(nb), l3(nb), l4(nb);
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
  // lists 1 and 3 for leaves below the root (d3tlst31, d3tstrcr.f:912-1069);
  // the coarse-to-fine list-1 entries are mirrored afterwards.
  std::vector<std::vector<std::pair<int, int>>> mirror(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int b = 1; b < nb; b++) {
    if (t.nchild[b]) continue;
    int tid = omp_get_thread_num();
    std::vector<int> stack;
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
  for (auto& v : mirror)
    for (auto& pr : v) l1[pr.first].push_back(pr.second);
  for (int b = 0; b < nb; b++)
    for (int j : l3[b]) l4[j].push_back(b);
  flatten(l1, t.list[0]);
  flatten(l2, t.list[1]);
  flatten(l3, t.list[2]);
  flatten(l4, t.list[3]);
  return 0;
}

}  // namespace fmm3d
