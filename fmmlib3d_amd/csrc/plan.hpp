// Flattened tree + schedule shared by the Laplace and Helmholtz drivers:
// one int buffer and one double buffer uploaded per call, work lists per
// FMM step, plus small launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "lap_kernels.hpp"
#include "octree.hpp"
#include "runtime.hpp"

namespace fmm3d {

// Evaluation points per P2P / L2P work item and threads per workgroup: one
// wave per item, so a leaf of n points keeps ceil(n/64) waves busy (with
// 256-point items a 244-point leaf left 12 lanes idle and a 300-point leaf
// 212)
constexpr int kP2PBlock = 64;
// items per workgroup (one per wave): the workgroups stay 256 threads wide,
// like those of the far-field kernels that run beside the near field
constexpr int kItemWaves = 4;
inline unsigned item_groups(int n) { return (unsigned)((n + kItemWaves - 1) / kItemWaves); }

// iprec -> tolerance and leaf size (src/lfmm3dpart.f:277-300)
inline bool iprec_params(int iprec, int ns, int nt, double& eps, int& nbox) {
  static const double e[] = {.5e0, .5e-1, .5e-2, .5e-3, .5e-6, .5e-9, .5e-12, .5e-15, 0.0};
  static const int nb[] = {40, 50, 80, 160, 400, 800, 1200, 1400, 0};
  if (iprec < -2 || iprec > 6) return false;
  eps = e[iprec + 2];
  nbox = (iprec == 6) ? ns + nt : nb[iprec + 2];
  return true;
}

struct Flat {
  std::vector<int> I;
  std::vector<double> D;
  size_t i(const std::vector<int>& v) {
    size_t o = I.size();
    I.insert(I.end(), v.begin(), v.end());
    while (I.size() % 16) I.push_back(0);
    return o;
  }
  size_t d(const std::vector<double>& v) {
    size_t o = D.size();
    D.insert(D.end(), v.begin(), v.end());
    while (D.size() % 8) D.push_back(0.0);
    return o;
  }
};

struct PlanOffsets {
  size_t level, isrc, nsrc, itrg, ntrg, nchild, child, ijk, l1o, l1i, l2o, l2i, l3o, l3i, l4o, l4i;
  size_t p2m, l2p, m2p, p2l, m2l, m2m, l2l, p2p0, p2p1, sperm, tperm, l2pi;
  size_t center, inv_side;
  int n_p2m, n_l2p, n_m2p, n_p2l, n_m2l, n_p2p0, n_p2p1, n_l2pi;
  std::vector<std::pair<int, int>> m2m_groups, l2l_groups;  // (first, count) in launch order
  std::vector<int> m2m_group_level, l2l_group_level;         // parent level of each group
};

struct Flags {
  int ifcharge, ifdipole, ifpot, iffld, ifpottarg, iffldtarg;
};

// Part A of the plan: the boxes, the P2M list and the M2M groups, all that
// the first FMM steps need.  p2m_min_level: leaves above this level get no
// multipole (2 for Helmholtz, src/hfmm3dpart.f:681 forms them from level 2
// down only).  Needs the tree's boxes only, not its lists.
inline void build_plan_boxes(const Octree& t, const Flags& f, Flat& F, PlanOffsets& O, int p2m_min_level) {
  (void)f;
  const int nb = t.nboxes;
  // the Laplace kernels read a box's level only to find its scale 1/side: a
  // forest (Laplace only) uploads its per-tree scale index instead
  O.level = F.i(t.scale_id.empty() ? t.level : t.scale_id); O.isrc = F.i(t.isrc); O.nsrc = F.i(t.nsrc);
  O.itrg = F.i(t.itrg); O.ntrg = F.i(t.ntrg); O.nchild = F.i(t.nchild);
  O.child = F.i(t.child); O.ijk = F.i(t.ijk);
  std::vector<int> p2m;
  for (int b = 0; b < nb; b++)
    if (t.nchild[b] == 0 && t.nsrc[b] > 0 && t.level[b] >= p2m_min_level) p2m.push_back(b);
  O.p2m = F.i(p2m); O.n_p2m = (int)p2m.size();
  // M2M parents, levels nlev-1 .. 2 (lfmm3drouts.f:195-258)
  std::vector<int> m2m;
  O.m2m_groups.clear();
  O.m2m_group_level.clear();
  for (int l = t.nlev - 1; l >= 2; l--) {
    int first = (int)m2m.size();
    for (int b = t.ladd[l]; b < t.ladd[l + 1]; b++)
      if (t.nchild[b] > 0 && t.nsrc[b] > 0) m2m.push_back(b);
    if ((int)m2m.size() > first) {
      O.m2m_groups.push_back({first, (int)m2m.size() - first});
      O.m2m_group_level.push_back(l);
    }
  }
  O.m2m = F.i(m2m);
  O.sperm = F.i(t.src_perm);
  O.tperm = F.i(t.trg_perm);
  O.center = F.d(t.bcenter);
  if (t.scale_inv.empty()) {
    std::vector<double> inv(t.nlev + 1);
    for (int l = 0; l <= t.nlev; l++) inv[l] = 1.0 / t.side(l);
    O.inv_side = F.d(inv);
  } else {
    O.inv_side = F.d(t.scale_inv);
  }
}

// Work items (box, kind, first point) of kP2PBlock points for one box.  A
// box's items go into one workgroup (kItemWaves items) when they fit, so
// its waves share the source data in cache: the list is padded with null
// items (box -1, skipped by the kernels) up to the next workgroup when the
// box's items would straddle one.
inline void push_items(std::vector<int>& it, int b, int kind, int npts) {
  const int c = (npts + kP2PBlock - 1) / kP2PBlock;
  if (c == 0) return;
  const int used = (int)(it.size() / 3) % kItemWaves;
  if (used > 0 && used + c > kItemWaves)
    for (int k = used; k < kItemWaves; k++) { it.push_back(-1); it.push_back(kind); it.push_back(0); }
  for (int q = 0; q < npts; q += kP2PBlock) { it.push_back(b); it.push_back(kind); it.push_back(q); }
}

// Part B1: list 1 and the P2P work items (needs tree_lists_near).
inline void build_plan_near(const Octree& t, const Flags& f, Flat& F, PlanOffsets& O) {
  const int nb = t.nboxes;
  const bool want_src = f.ifpot || f.iffld, want_trg = f.ifpottarg || f.iffldtarg;
  O.l1o = F.i(t.list[0].off); O.l1i = F.i(t.list[0].idx);
  std::vector<int> p2p0, p2p1;
  for (int b = 0; b < nb; b++) {
    if (t.nchild[b] != 0) continue;
    if (want_src) push_items(p2p0, b, 0, t.nsrc[b]);
    if (want_trg) push_items(p2p1, b, 1, t.ntrg[b]);
  }
  O.p2p0 = F.i(p2p0); O.n_p2p0 = (int)p2p0.size() / 3;
  O.p2p1 = F.i(p2p1); O.n_p2p1 = (int)p2p1.size() / 3;
}

// Part B2: lists 2-4 and the work lists of the far-field steps (needs all lists).
inline void build_plan_far(const Octree& t, const Flags& f, Flat& F, PlanOffsets& O) {
  const int nb = t.nboxes;
  const bool want_src = f.ifpot || f.iffld, want_trg = f.ifpottarg || f.iffldtarg;
  // list 2 pruned as lfmm3drouts.f:385-387: sourceless sources, and
  // targetless targets when no source outputs are wanted
  {
    std::vector<int> off(nb + 1, 0), idx;
    idx.reserve(t.list[1].idx.size());
    for (int b = 0; b < nb; b++) {
      off[b] = (int)idx.size();
      bool skip_b = t.level[b] < 2 || (!want_src && t.ntrg[b] == 0);
      if (!skip_b)
        for (int k = t.list[1].off[b]; k < t.list[1].off[b + 1]; k++) {
          int j = t.list[1].idx[k];
          if (t.nsrc[j] > 0) idx.push_back(j);
        }
    }
    off[nb] = (int)idx.size();
    O.l2o = F.i(off); O.l2i = F.i(idx);
    std::vector<int> m2l;
    for (int b = 0; b < nb; b++)
      if (off[b + 1] > off[b]) m2l.push_back(b);
    O.m2l = F.i(m2l); O.n_m2l = (int)m2l.size();
  }
  O.l3o = F.i(t.list[2].off); O.l3i = F.i(t.list[2].idx);
  O.l4o = F.i(t.list[3].off); O.l4i = F.i(t.list[3].idx);
  std::vector<int> l2p, m2p, p2l;
  for (int b = 0; b < nb; b++) {
    const bool leaf = t.nchild[b] == 0;
    const bool has_pts = (want_src && t.nsrc[b] > 0) || (want_trg && t.ntrg[b] > 0);
    if (leaf && t.level[b] >= 2 && has_pts) l2p.push_back(b);
    if (t.list[2].count(b) > 0 && has_pts) m2p.push_back(b);
    if (t.list[3].count(b) > 0) {
      bool any = false;
      for (int k = t.list[3].off[b]; k < t.list[3].off[b + 1]; k++) any |= t.nsrc[t.list[3].idx[k]] > 0;
      if (any) p2l.push_back(b);
    }
  }
  O.l2p = F.i(l2p); O.n_l2p = (int)l2p.size();
  {
    // L2P work items (position in the L2P list, kind, first point): one
    // workgroup per kP2PBlock evaluation points of one kind
    std::vector<int> it;
    for (int i = 0; i < (int)l2p.size(); i++) {
      const int b = l2p[i];
      if (want_src) push_items(it, i, 0, t.nsrc[b]);
      if (want_trg) push_items(it, i, 1, t.ntrg[b]);
    }
    O.l2pi = F.i(it); O.n_l2pi = (int)it.size() / 3;
  }
  O.m2p = F.i(m2p); O.n_m2p = (int)m2p.size();
  O.p2l = F.i(p2l); O.n_p2l = (int)p2l.size();
  // L2L parents, levels 2 .. nlev-1 (lfmm3drouts.f:498-558)
  std::vector<int> l2l;
  O.l2l_groups.clear();
  O.l2l_group_level.clear();
  for (int l = 2; l <= t.nlev - 1; l++) {
    int first = (int)l2l.size();
    for (int b = t.ladd[l]; b < t.ladd[l + 1]; b++)
      if (t.nchild[b] > 0) l2l.push_back(b);
    if ((int)l2l.size() > first) {
      O.l2l_groups.push_back({first, (int)l2l.size() - first});
      O.l2l_group_level.push_back(l);
    }
  }
  O.l2l = F.i(l2l);
}

inline void build_plan_lists(const Octree& t, const Flags& f, Flat& F, PlanOffsets& O) {
  build_plan_near(t, f, F, O);
  build_plan_far(t, f, F, O);
}

// All parts into one flat buffer.
inline void build_plan(const Octree& t, const Flags& f, Flat& F, PlanOffsets& O, int p2m_min_level) {
  build_plan_boxes(t, f, F, O, p2m_min_level);
  build_plan_lists(t, f, F, O);
}

// Device plan over part A at (IA, DA) and part B at IB (IB = IA when one
// flat holds both; IB may be null until part B is uploaded, see
// devplan_lists).
inline void devplan_near(DevPlan& P, const PlanOffsets& O, const int* I) {
  P.l1_off = I + O.l1o; P.l1_idx = I + O.l1i;
  P.p2p_items = I + O.p2p0; P.n_p2p = O.n_p2p0;
}
inline void devplan_far(DevPlan& P, const PlanOffsets& O, const int* I) {
  P.l2_off = I + O.l2o; P.l2_idx = I + O.l2i;
  P.l3_off = I + O.l3o; P.l3_idx = I + O.l3i; P.l4_off = I + O.l4o; P.l4_idx = I + O.l4i;
  P.l2p_boxes = I + O.l2p; P.n_l2p = O.n_l2p;
  P.l2p_items = I + O.l2pi; P.n_l2pi = O.n_l2pi;
  P.m2p_boxes = I + O.m2p; P.n_m2p = O.n_m2p;
  P.p2l_boxes = I + O.p2l; P.n_p2l = O.n_p2l;
  P.m2l_boxes = I + O.m2l; P.n_m2l = O.n_m2l;
  P.l2l_boxes = I + O.l2l; P.l2l_lev_off = nullptr;
}
inline void devplan_lists(DevPlan& P, const PlanOffsets& O, const int* I) {
  devplan_near(P, O, I);
  devplan_far(P, O, I);
}

inline DevPlan make_devplan(const Octree& t, int p, const PlanOffsets& O, const int* I, const double* D,
                            const int* IB = nullptr) {
  DevPlan P{};
  P.nboxes = t.nboxes; P.nlev = t.nlev; P.p = p; P.K = (p + 1) * (p + 1);
  P.ns = t.ns; P.nt = t.nt;
  P.level = I + O.level; P.isrc = I + O.isrc; P.nsrc = I + O.nsrc; P.itrg = I + O.itrg;
  P.ntrg = I + O.ntrg; P.nchild = I + O.nchild; P.child = I + O.child; P.ijk = I + O.ijk;
  P.center = D + O.center; P.inv_side = D + O.inv_side;
  P.p2m_boxes = I + O.p2m; P.n_p2m = O.n_p2m;
  P.m2m_boxes = I + O.m2m; P.m2m_lev_off = nullptr;
  P.src_perm = I + O.sperm; P.trg_perm = I + O.tperm;
  if (IB) devplan_lists(P, O, IB);
  return P;
}

inline size_t max_lds_bytes(Context& ctx) {
  static size_t v = 0;
  if (!v) {
    hipDeviceProp_t prop;
    FMM_HIP(hipGetDeviceProperties(&prop, ctx.device));
    v = prop.sharedMemPerBlock;
  }
  return v;
}

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace fmm3d
