// Laplace particle FMM driver: device octree -> box plan -> the eight FMM
// steps of lfmm3dparttargmain (src/lfmm3dpart.f:528-1169) as HIP kernels,
// the interaction lists built on the host while the first kernels run, the
// near field on a second stream -> far + near scattered back to the
// caller's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fmm3d_internal.hpp"
#include "lap_kernels.hip"
#include "m2l_mfma.hpp"
#include "hess_kernels.hpp"
#include "qp_kernels.hpp"
#include "tria_kernels.hpp"
#include "tria_rules.hpp"
#include "octree.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "tables.hpp"
#include "tree_gpu.hpp"

namespace fmm3d {

namespace {

// Translation tables of one order, flattened for upload: ints, doubles, longs
// and the FP32 copies of the grouped-M2L tables, with the offsets of each part.
struct LapTabSet {
  std::vector<int> I;
  std::vector<double> D;
  std::vector<long> L;
  std::vector<float> F;
  size_t off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t off_l = 0, off_re = 0;
  size_t off_i[4] = {0, 0, 0, 0};  // grouped M2L items: ritems, rcount, zitems, zoff
  size_t off_g[4] = {0, 0, 0, 0};  // grouped M2L tables: rot_half, z_m2l_t (doubles); their offsets (longs)
  size_t off_f = 0;
  long zt_size = 0;
  size_t off_m[4] = {0, 0, 0, 0};    // MFMA tables: mf_rot, mf_z (doubles); mf_rot_off, mf_z_off (longs)
  size_t off_mf[2] = {0, 0};         // their FP32 copies
  long mf_rot_back = 0, mf_z_size = 0;
};

// The host tables of order p, built once per process (tens to a hundred ms
// at p = 26..39) and shared by every context.
std::shared_ptr<const LapTabSet> lap_host_tables(int p, double eps) {
  static std::mutex mu;
  static std::map<std::pair<int, double>, std::shared_ptr<const LapTabSet>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({p, eps});
  if (it != cache.end()) return it->second;
  HostTables H;
  build_tables(p, eps, H);
  auto T = std::make_shared<LapTabSet>();
  // ints: off_rot | off_dist | off_ptrunc | oct_rot | grouped items ; doubles:
  // fac | off_phase | oct_phase | rot | z_m2l | z_m2m | z_l2l | delta | rot_e |
  // rot_half | z_m2l_t ; longs: rot_off | delta_off | rot_half_off | zt_off
  auto& Ii = T->I;
  auto& Dd = T->D;
  Ii.insert(Ii.end(), H.off_rot.begin(), H.off_rot.end());
  Ii.insert(Ii.end(), H.off_dist.begin(), H.off_dist.end());
  Ii.insert(Ii.end(), H.off_ptrunc.begin(), H.off_ptrunc.end());
  Ii.insert(Ii.end(), H.oct_rot.begin(), H.oct_rot.end());
  const std::vector<int>* gi[4] = {&H.grp_ritems, &H.grp_rcount, &H.grp_zitems, &H.grp_zoff};
  for (int q = 0; q < 4; q++) {
    T->off_i[q] = Ii.size();
    Ii.insert(Ii.end(), gi[q]->begin(), gi[q]->end());
  }
  const std::vector<double>* dparts[9] = {&H.fac, &H.off_phase, &H.oct_phase, &H.rot, &H.z_m2l,
                                          &H.z_m2m, &H.z_l2l, &H.delta, &H.rot_e};
  size_t o[9];
  for (int q = 0; q < 9; q++) {
    o[q] = Dd.size();
    Dd.insert(Dd.end(), dparts[q]->begin(), dparts[q]->end());
  }
  // 32-byte aligned starts for the padded tables of the grouped M2L
  while (Dd.size() % 4) Dd.push_back(0.0);
  T->off_g[0] = Dd.size();
  Dd.insert(Dd.end(), H.rot_half.begin(), H.rot_half.end());
  while (Dd.size() % 4) Dd.push_back(0.0);
  T->off_g[1] = Dd.size();
  Dd.insert(Dd.end(), H.z_m2l_t.begin(), H.z_m2l_t.end());
  T->L = H.rot_off;
  T->off_l = T->L.size();
  T->L.insert(T->L.end(), H.delta_off.begin(), H.delta_off.end());
  T->off_g[2] = T->L.size();
  T->L.insert(T->L.end(), H.rot_half_off.begin(), H.rot_half_off.end());
  T->off_g[3] = T->L.size();
  T->L.insert(T->L.end(), H.zt_off.begin(), H.zt_off.end());
  T->zt_size = H.zt_size;
  // MFMA-ordered M2L tables (64-value blocks, 512-byte aligned starts)
  while (Dd.size() % 64) Dd.push_back(0.0);
  T->off_m[0] = Dd.size();
  Dd.insert(Dd.end(), H.mf_rot.begin(), H.mf_rot.end());
  while (Dd.size() % 64) Dd.push_back(0.0);
  T->off_m[1] = Dd.size();
  Dd.insert(Dd.end(), H.mf_z.begin(), H.mf_z.end());
  T->off_m[2] = T->L.size();
  T->L.insert(T->L.end(), H.mf_rot_off.begin(), H.mf_rot_off.end());
  T->off_m[3] = T->L.size();
  T->L.insert(T->L.end(), H.mf_z_off.begin(), H.mf_z_off.end());
  T->mf_rot_back = H.mf_rot_back;
  T->mf_z_size = H.mf_z_size;
  // FP32 copies of the grouped-M2L tables for the FP32 far field
  T->F.assign(H.rot_half.begin(), H.rot_half.end());
  while (T->F.size() % 4) T->F.push_back(0.0f);
  T->off_f = T->F.size();
  T->F.insert(T->F.end(), H.z_m2l_t.begin(), H.z_m2l_t.end());
  while (T->F.size() % 64) T->F.push_back(0.0f);
  T->off_mf[0] = T->F.size();
  T->F.insert(T->F.end(), H.mf_rot.begin(), H.mf_rot.end());
  while (T->F.size() % 64) T->F.push_back(0.0f);
  T->off_mf[1] = T->F.size();
  T->F.insert(T->F.end(), H.mf_z.begin(), H.mf_z.end());
  for (int q = 0; q < 7; q++) T->off[q] = o[q];
  T->off[7] = o[7];
  T->off_re = o[8];
  cache[{p, eps}] = T;
  return T;
}

// The device tables of (order p, tolerance eps) for this context: uploaded
// on first use of the pair and kept (each pair under its own arena names).
DevTables ensure_tables(Context& ctx, int p, double eps) {
  const auto hs = lap_host_tables(p, eps);
  const LapTabSet& H = *hs;
  char key[48];
  std::snprintf(key, sizeof key, ".%d.%a", p, eps);
  const std::string sfx = key;
  int* di = (int*)ctx.arena.get("tab_i" + sfx, H.I.size() * sizeof(int));
  double* dd = (double*)ctx.arena.get("tab_d" + sfx, H.D.size() * sizeof(double));
  long* dl = (long*)ctx.arena.get("tab_l" + sfx, H.L.size() * sizeof(long));
  float* df = (float*)ctx.arena.get("tab_f" + sfx, H.F.size() * sizeof(float));
  if (!ctx.lap_tabs.count({p, eps})) {
    FMM_HIP(hipMemcpyAsync(di, H.I.data(), H.I.size() * sizeof(int), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipMemcpyAsync(dd, H.D.data(), H.D.size() * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipMemcpyAsync(dl, H.L.data(), H.L.size() * sizeof(long), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipMemcpyAsync(df, H.F.data(), H.F.size() * sizeof(float), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    ctx.lap_tabs.insert({p, eps});
  }
  DevTables T;
  T.p = p;
  T.off_rot = di; T.off_dist = di + 343; T.off_ptrunc = di + 686; T.oct_rot = di + 1029;
  T.fac = dd + H.off[0]; T.off_phase = dd + H.off[1]; T.oct_phase = dd + H.off[2];
  T.rot = dd + H.off[3]; T.z_m2l = dd + H.off[4]; T.z_m2m = dd + H.off[5];
  T.z_l2l = dd + H.off[6];
  T.rot_off = dl;
  T.delta = dd + H.off[7];
  T.delta_off = dl + H.off_l;
  T.rot_e = dd + H.off_re;
  T.grp_ritems = di + H.off_i[0]; T.grp_rcount = di + H.off_i[1];
  T.grp_zitems = di + H.off_i[2]; T.grp_zoff = di + H.off_i[3];
  T.rot_half = dd + H.off_g[0]; T.z_m2l_t = dd + H.off_g[1];
  T.rot_half_off = dl + H.off_g[2]; T.zt_off = dl + H.off_g[3];
  T.zt_size = H.zt_size;
  T.rot_half_f = df;
  T.z_m2l_t_f = df + H.off_f;
  T.mf_rot = dd + H.off_m[0]; T.mf_z = dd + H.off_m[1];
  T.mf_rot_f = df + H.off_mf[0]; T.mf_z_f = df + H.off_mf[1];
  T.mf_rot_off = dl + H.off_m[2]; T.mf_z_off = dl + H.off_m[3];
  T.mf_rot_back = H.mf_rot_back; T.mf_z_size = H.mf_z_size;
  return T;
}

template <typename TN>
void launch_p2p(const DevPlan& P, const int* items, int n, int kind, const double* spos,
                const TN* str, const double* str64, const double* tpos, int ifcharge, int ifdipole, bool fld,
                double* pot, double* fo, hipStream_t s) {
  if (n == 0) return;
  DevPlan Q = P;
  Q.p2p_items = items;
  Q.n_p2p = n;
  dim3 g(item_groups(n)), b(kP2PBlock * kItemWaves);
#define P2P_CASE(C, Dp, F)                                                                        \
  if (ifcharge == C && ifdipole == Dp && fld == F) {                                             \
    hipLaunchKernelGGL((k_p2p<TN, C, Dp, F>), g, b, 0, s, Q, spos, str, str64, tpos, kind, pot, fo); \
    return;                                                                                       \
  }
  P2P_CASE(1, 0, false) P2P_CASE(1, 0, true) P2P_CASE(0, 1, false) P2P_CASE(0, 1, true)
  P2P_CASE(1, 1, false) P2P_CASE(1, 1, true)
#undef P2P_CASE
}

// Factor table of order p + 1 for the Hessian kernels (degrees up to p + 2),
// built once per process and order, uploaded once per context and order.
const double* hess_factors(Context& ctx, int p) {
  static std::mutex mu;
  static std::map<int, std::vector<double>> cache;
  const std::vector<double>* h;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(p);
    if (it == cache.end()) {
      std::vector<double> v;
      build_fac_table(p + 1, v);
      it = cache.emplace(p, std::move(v)).first;
    }
    h = &it->second;
  }
  const std::string name = "hess_fac." + std::to_string(p);
  double* d = (double*)ctx.arena.get(name, h->size() * sizeof(double));
  FMM_HIP(hipMemcpyAsync(d, h->data(), h->size() * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
  return d;
}

}  // namespace

template <typename T, typename TN>
int lap_solve(Context& ctx, const LapCall& c) {
  auto t0 = std::chrono::steady_clock::now();
  double eps;
  int nbox;
  if (!iprec_params(c.iprec, c.ns, c.nt, eps, nbox)) return 32;
  const bool tri = c.tri_dev != nullptr;
  if (tri) {
    if (!(std::is_same<T, double>::value && std::is_same<TN, double>::value) || c.ifhess || c.ifhesstarg)
      return 128;
    nbox = tri_nbox(c.iprec, c.ns, c.nt);
  }
  const bool qp = c.ifquad != 0;
  if (qp && (tri || c.ifhess || c.ifhesstarg || !(std::is_same<T, double>::value && std::is_same<TN, double>::value)))
    return 128;
  const int ns = c.ns, nt = c.nt;
  const Flags f{c.ifcharge ? 1 : 0, c.ifdipole ? 1 : 0, c.ifpot ? 1 : 0, c.iffld ? 1 : 0,
                nt > 0 && c.ifpottarg ? 1 : 0, nt > 0 && c.iffldtarg ? 1 : 0};
  thread_local Octree own_tree;  // storage reused across calls
  Octree& tree = c.tree_in ? const_cast<Octree&>(*c.tree_in) : own_tree;
  const bool prebuilt = c.tree_in != nullptr;
  const int* dsperm = nullptr;
  const int* dtperm = nullptr;
  // the interaction lists are built after the first kernels are queued
  if (prebuilt) {
    if (tree.ns != ns || tree.nt != nt) return 128;
    dsperm = c.sperm_in;
    dtperm = c.tperm_in;
  } else if (device_tree_enabled()) {
    if (build_octree_device(ctx, ctx.stream, c.src_dev, ns, c.trg_dev, nt, nbox, tree, &dsperm, &dtperm, false) != 0)
      return 4;
  } else if (build_octree(c.src_host, ns, c.trg_host, nt, nbox, tree, false) != 0) {
    return 4;
  }
  const bool lists_done = prebuilt && !c.tree_lists_todo;
  // FP32 expansions at iprec 2 (tolerance 0.5e-6) only where every leaf lies
  // on one level.  Boxes of different sizes (lists 3 and 4, adaptive trees of
  // clustered points) evaluate expansions near their radius of convergence,
  // where the FP32 rounding of the large high-degree terms reached 6e-7 of
  // the potential on clustered inputs; such a tree runs the far field in
  // FP64 (the same tree, handed on).
  if (std::is_same<T, float>::value && c.iprec == 2 && !tree.leaves_on_one_level()) {
    LapCall c2 = c;
    c2.tree_in = &tree;
    c2.tree_lists_todo = !lists_done;
    c2.sperm_in = dsperm;
    c2.tperm_in = dtperm;
    last_precision() &= ~2;
    return lap_solve<double, TN>(ctx, c2);
  }
  auto t_tree = std::chrono::steady_clock::now();
  const int p = laplace_nterms(eps);
  const int K = (p + 1) * (p + 1);
  if (sizeof(T) * 4 * (size_t)K > max_lds_bytes(ctx)) return 16;
  thread_local Flat FA, FB;  // plan parts A (boxes) and B (lists), storage reused
  FA.I.clear(); FA.D.clear(); FB.I.clear(); FB.D.clear();
  PlanOffsets O;
  build_plan_boxes(tree, f, FA, O, 0);
  auto t1 = std::chrono::steady_clock::now();
  hipStream_t s = ctx.stream;
  // stage timers (event records), not while the launches are captured
  const bool timed = !c.graph;
  Timer tm;
  tm.start(s, 11);
  auto mark = [&](int i) {
    if (timed) tm.mark(i);
  };
  // ---- device buffers; plan uploads go through pinned staging buffers so
  // that they do not wait for the kernels already queued
  auto upload_ints = [&](const char* name, const std::vector<int>& v, hipStream_t st) {
    int* h = (int*)ctx.arena.get_host(std::string(name) + "_h", v.size() * sizeof(int));
    std::memcpy(h, v.data(), v.size() * sizeof(int));
    int* d = (int*)ctx.arena.get(name, v.size() * sizeof(int));
    FMM_HIP(hipMemcpyAsync(d, h, v.size() * sizeof(int), hipMemcpyHostToDevice, st));
    return d;
  };
  // the near field (P2P) runs on a second stream into its own sums, which
  // the final scatter adds to the far-field sums
  hipStream_t s2 = near_stream(ctx);
  int* dI = upload_ints("plan_i", FA.I, s);
  double* hD = (double*)ctx.arena.get_host("plan_d_h", FA.D.size() * sizeof(double));
  std::memcpy(hD, FA.D.data(), FA.D.size() * sizeof(double));
  double* dD = (double*)ctx.arena.get("plan_d", FA.D.size() * sizeof(double));
  FMM_HIP(hipMemcpyAsync(dD, hD, FA.D.size() * sizeof(double), hipMemcpyHostToDevice, s));
  DevTables Tb = ensure_tables(ctx, p, eps);
  DevPlan P = make_devplan(tree, p, O, dI, dD);
  // expansion storage for every box: allocated before any kernel is queued
  const size_t nb = tree.nboxes;
  constexpr bool near32 = sizeof(TN) < sizeof(double);
  double* spos = (double*)ctx.arena.get("spos", sizeof(double) * 4 * (size_t)ns);
  double* str64 = (double*)ctx.arena.get("str64", sizeof(double) * 8 * (size_t)ns);
  float* str32 = near32 ? (float*)ctx.arena.get("str32", sizeof(float) * 8 * (size_t)ns) : nullptr;
  double* tpos = (double*)ctx.arena.get("tpos", sizeof(double) * 4 * (size_t)std::max(nt, 1));
  T* mp = (T*)ctx.arena.get("mp", sizeof(T) * 2 * K * nb, 16);
  T* lo = (T*)ctx.arena.get("lo", sizeof(T) * 2 * K * nb, 16);
  double* pot_s = (double*)ctx.arena.get("pot_s", sizeof(double) * 2 * (size_t)ns);
  double* fld_s = (double*)ctx.arena.get("fld_s", sizeof(double) * 6 * (size_t)ns);
  double* pot_t = (double*)ctx.arena.get("pot_t", sizeof(double) * 2 * (size_t)std::max(nt, 1));
  double* fld_t = (double*)ctx.arena.get("fld_t", sizeof(double) * 6 * (size_t)std::max(nt, 1));
  double* pot_sn = (double*)ctx.arena.get("pot_sn", sizeof(double) * 2 * (size_t)ns);
  double* fld_sn = (double*)ctx.arena.get("fld_sn", sizeof(double) * 6 * (size_t)ns);
  double* pot_tn = (double*)ctx.arena.get("pot_tn", sizeof(double) * 2 * (size_t)std::max(nt, 1));
  double* fld_tn = (double*)ctx.arena.get("fld_tn", sizeof(double) * 6 * (size_t)std::max(nt, 1));
  const TN* strn = near32 ? (const TN*)(const void*)str32 : (const TN*)(const void*)str64;
  const int* sperm = dsperm ? dsperm : dI + O.sperm;
  const int* tperm = dsperm ? dtperm : dI + O.tperm;
  P.src_perm = sperm;
  P.trg_perm = tperm;
  // triangle FMM: frames and densities in sorted order, and the quadrature
  // points that stand for the triangles in P2M / P2L (PF: the plan with each
  // box's source range counted in points)
  const double* fpos = spos;
  const double* fstr = str64;
  double *tgeo = nullptr, *tstr = nullptr;
  int* tri_rng = nullptr;
  std::vector<double> rule;
  if (tri) rule = tri_rule(c.iprec, false);
  const int nq = (int)rule.size() / 3;
  // graph mode: every launch and copy from here to the scatter, on both
  // streams (the near-field stream joins through the events), is captured
  // and then launched as one graph; buffers are zeroed by kernels, not
  // memset nodes
  hipGraph_t graph = nullptr;
  struct CaptureGuard {  // ends a capture that an error left open
    hipStream_t s;
    bool on = false;
    ~CaptureGuard() {
      if (!on) return;
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
    }
  } capture{s};
  if (c.graph) {
    FMM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    capture.on = true;
  }
  // (triangles: the point strengths are not used -- k_tri_prep forms the
  // triangle densities and quadrature points -- and no dipole vectors are read)
  hipLaunchKernelGGL(k_gather_sources, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, sperm, c.src_dev,
                     c.q_dev, c.ds_dev, c.dv_dev, tri ? 0 : f.ifcharge, tri ? 0 : f.ifdipole, spos, str64, str32);
  if (nt > 0)
    hipLaunchKernelGGL(k_gather_targets, dim3(blocks(nt, 256)), dim3(256), 0, s, nt, tperm, c.trg_dev, tpos);
  double* qstr = nullptr;
  if (qp) {
    qstr = (double*)ctx.arena.get("qp_str", sizeof(double) * 8 * (size_t)ns);
    hipLaunchKernelGGL(k_gather_qp, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, sperm, c.qs_dev, c.qv_dev, qstr);
  }
  auto zero = [&](auto* a, size_t n) {
    using E = std::remove_pointer_t<decltype(a)>;
    hipLaunchKernelGGL((k_fill_zero<E>), dim3(std::min<size_t>(blocks(n, 256), 4096)), dim3(256), 0, s, n, a);
  };
  zero(mp, 2 * (size_t)K * nb);
  zero(lo, 2 * (size_t)K * nb);
  zero(pot_s, 2 * (size_t)ns);
  zero(fld_s, 6 * (size_t)ns);
  zero(pot_sn, 2 * (size_t)ns);
  zero(fld_sn, 6 * (size_t)ns);
  if (nt > 0) {
    zero(pot_t, 2 * (size_t)nt);
    zero(fld_t, 6 * (size_t)nt);
    zero(pot_tn, 2 * (size_t)nt);
    zero(fld_tn, 6 * (size_t)nt);
  }
  if (tri) {
    tgeo = (double*)ctx.arena.get("tri_geo", sizeof(double) * kTriGeo * (size_t)ns);
    tstr = (double*)ctx.arena.get("tri_str", sizeof(double) * 4 * (size_t)ns);
    double* npos = (double*)ctx.arena.get("tri_npos", sizeof(double) * 4 * (size_t)ns * nq);
    double* nstr = (double*)ctx.arena.get("tri_nstr", sizeof(double) * 8 * (size_t)ns * nq);
    double* hr = (double*)ctx.arena.get_host("tri_rule_h", sizeof(double) * rule.size());
    std::memcpy(hr, rule.data(), sizeof(double) * rule.size());
    double* dr = (double*)ctx.arena.get("tri_rule", sizeof(double) * rule.size());
    FMM_HIP(hipMemcpyAsync(dr, hr, sizeof(double) * rule.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tri_prep, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, sperm, c.tri_dev, c.q_dev,
                       c.ds_dev, f.ifcharge, f.ifdipole, nq, dr, tgeo, tstr, npos, nstr);
    tri_rng = (int*)ctx.arena.get("tri_ranges", sizeof(int) * 2 * nb);
    hipLaunchKernelGGL(k_scale_ranges, dim3(blocks(nb, 256)), dim3(256), 0, s, (int)nb, P.isrc, P.nsrc, nq, tri_rng,
                       tri_rng + nb);
    fpos = npos;
    fstr = nstr;
  }
  FMM_HIP(hipEventRecord(ctx.aux_ev[0], s));  // inputs gathered, sums zeroed
  mark(1);
  const double* fac = Tb.fac;
  // ---- STEP 1: P2M, STEP 2: P2L (list 4)
  // k_form2 (per-thread sums) when its half table fits I1 (or 2 I1) entries
  // per thread, else k_form (per-coefficient contraction)
  auto form = [&](auto irreg_tag, int nbox_form, T* dst) {
    constexpr bool IR = decltype(irreg_tag)::value;
    constexpr int NT = form2_threads<IR>(), I1 = 512 / NT;
    if (nbox_form == 0 || (!f.ifcharge && !f.ifdipole)) return;
    const int pt = IR ? p + 1 : p;
    const int H = (pt + 1) * (pt + 2) / 2;
    const size_t P2 = (size_t)(pt + 1) * (pt + 1);
    const int HS = form2_row_stride<T>(H);
    DevPlan PF = P;  // the current plan (P2L needs list 4), source ranges in points for triangles
    if (tri) {
      PF.isrc = tri_rng;
      PF.nsrc = tri_rng + nb;
    }
    int ts = (int)(65536 / (2 * sizeof(T) * (size_t)HS));
    ts = std::max(1, std::min(32, ts));
    const size_t lds2 = form2_fac_bytes(p) + 2 * sizeof(T) * std::max((size_t)ts * HS, 4 * P2) + (kFormScStride + 8) * sizeof(T) * ts;
#define FORM2(C, D, I)                                                                                 \
  if (f.ifcharge == C && f.ifdipole == D)                                                              \
    hipLaunchKernelGGL((k_form2<T, IR, C, D, I>), dim3(nbox_form), dim3(NT), lds2, s, PF, fac, fpos, fstr, ts, dst);
    const bool lds2_fits = lds2 <= max_lds_bytes(ctx);
    if (H <= I1 * NT && lds2_fits) {
      FORM2(1, 0, I1) FORM2(0, 1, I1) FORM2(1, 1, I1)
    } else if (H <= 2 * I1 * NT && lds2_fits) {
      FORM2(1, 0, 2 * I1) FORM2(0, 1, 2 * I1) FORM2(1, 1, 2 * I1)
    } else {
#undef FORM2
      // tile of ts sources: 2K accumulators + ts (2H + 11) table, positions
      // and strengths, sized in signed arithmetic against the device limit
      const long avail = (long)(max_lds_bytes(ctx) / sizeof(T)) - 2L * K;
      const long ts1 = avail / (2L * H + 11);
      if (ts1 < 1) throw AllocError("P2M/P2L tile does not fit LDS", 16);
      const int ts = (int)std::min<long>(32, ts1);
      const size_t lds = sizeof(T) * (2 * (size_t)K + (size_t)ts * (2 * H + 11));
      hipLaunchKernelGGL((k_form<T, IR>), dim3(nbox_form), dim3(256), lds, s, PF, fac, fpos, fstr,
                         f.ifcharge, f.ifdipole, ts, dst);
    }
  };
  form(std::false_type{}, P.n_p2m, mp);
  // quadrupoles: their multipoles (degrees >= 2) and, below, their list-4
  // locals, added to those of the charges and dipoles (FP64)
  const double* qfac = qp ? hess_factors(ctx, p) : nullptr;
  auto qp_form = [&](auto irreg_tag, int nbox_form, T* dst) {
    constexpr bool IR = decltype(irreg_tag)::value;
    if (!qp || nbox_form == 0 || (!IR && p < 2)) return;
    const int pt = IR ? p + 2 : p - 2;
    // tiles within 64 KB (two workgroups per CU) where one source fits, else
    // within the device's LDS
    const long per = (long)(qp_form_lds_doubles(p, pt, 1) - 2 * (size_t)K);
    const long lim = (long)(max_lds_bytes(ctx) / sizeof(double)) - 2L * K;
    const long small = std::min<long>(65536 / sizeof(double), lim + 2L * K) - 2L * K;
    const long avail = small >= per ? small : lim;
    const int ts = (int)std::min<long>(32, avail / per);
    if (ts < 1) throw AllocError("quadrupole P2M/P2L tile does not fit LDS", 16);
    const size_t lds = sizeof(double) * qp_form_lds_doubles(p, pt, ts);
    hipLaunchKernelGGL((k_qp_form<IR>), dim3(nbox_form), dim3(256), lds, s, P, qfac, spos, qstr, ts,
                       reinterpret_cast<double*>(dst));
  };
  qp_form(std::false_type{}, P.n_p2m, mp);
  mark(2);
  // ---- STEP 3: M2M, bottom-up
  const size_t lds_ps = sizeof(T) * 4 * K;
  for (auto& g : O.m2m_groups)
    hipLaunchKernelGGL((k_m2m<T>), dim3(g.second), dim3(256), lds_ps, s, P, Tb, g.first, mp);
  mark(3);
  // ---- while the device runs P2M and M2M: list 1 and the P2P plan (part B1)
  auto t_lists0 = std::chrono::steady_clock::now();
  if (!lists_done) tree_lists_near(tree);
  build_plan_near(tree, f, FB, O);
  // the P2P work items are addressed explicitly: they must lie in part B1
  if (O.p2p0 + 3 * (size_t)O.n_p2p0 > FB.I.size() || O.p2p1 + 3 * (size_t)O.n_p2p1 > FB.I.size())
    throw HipError("plan: P2P work items outside the near-field part of the plan");
  auto t_lists1 = std::chrono::steady_clock::now();
  // ---- STEP 8: P2P (self + list 1) on the second stream, once the inputs
  // are gathered; its plan part (B1) is uploaded on that stream
  FMM_HIP(hipStreamWaitEvent(s2, ctx.aux_ev[0], 0));
  Timer tn;
  if (timed) tn.start(s2, 2);
  const int* dIB = upload_ints("plan_ib", FB.I, s2);  // part B1: offsets in O from build_plan_near
  devplan_near(P, O, dIB);
  const int* dIB_near = dIB;
  if (tri && c.far_only) {
    // the far field alone
  } else if (tri && (f.ifcharge || f.ifdipole)) {
    auto tri_near = [&](const int* items, int n, int kind, double* po, double* fo) {
      if (n == 0) return;
      DevPlan Q = P;
      Q.p2p_items = items;
      Q.n_p2p = n;
      dim3 g(item_groups(n)), bl(kP2PBlock * kItemWaves);
      if (f.ifcharge && f.ifdipole)
        hipLaunchKernelGGL((k_tri_p2p<true, true>), g, bl, 0, s2, Q, tgeo, tstr, spos, tpos, kind, po, fo);
      else if (f.ifcharge)
        hipLaunchKernelGGL((k_tri_p2p<true, false>), g, bl, 0, s2, Q, tgeo, tstr, spos, tpos, kind, po, fo);
      else
        hipLaunchKernelGGL((k_tri_p2p<false, true>), g, bl, 0, s2, Q, tgeo, tstr, spos, tpos, kind, po, fo);
    };
    tri_near(dIB + O.p2p0, O.n_p2p0, 0, pot_sn, f.iffld ? fld_sn : nullptr);
    tri_near(dIB + O.p2p1, O.n_p2p1, 1, pot_tn, f.iffldtarg ? fld_tn : nullptr);
  } else if (f.ifcharge || f.ifdipole) {
    launch_p2p<TN>(P, dIB + O.p2p0, O.n_p2p0, 0, spos, strn, str64, tpos, f.ifcharge, f.ifdipole,
                   f.iffld != 0, pot_sn, fld_sn, s2);
    launch_p2p<TN>(P, dIB + O.p2p1, O.n_p2p1, 1, spos, strn, str64, tpos, f.ifcharge, f.ifdipole,
                   f.iffldtarg != 0, pot_tn, fld_tn, s2);
  }
  if (qp) {
    // after the charges' near field on the same stream: both add into the sums
    auto qp_near = [&](const int* items, int n, int kind, double* po, double* fo) {
      if (n == 0) return;
      DevPlan Q = P;
      Q.p2p_items = items;
      Q.n_p2p = n;
      hipLaunchKernelGGL(k_p2p_qp, dim3(item_groups(n)), dim3(kP2PBlock * kItemWaves), 0, s2, Q, spos, qstr, tpos,
                         kind, po, fo);
    };
    qp_near(dIB + O.p2p0, O.n_p2p0, 0, pot_sn, f.iffld ? fld_sn : nullptr);
    qp_near(dIB + O.p2p1, O.n_p2p1, 1, pot_tn, f.iffldtarg ? fld_tn : nullptr);
  }
  if (timed) tn.mark(1);
  FMM_HIP(hipEventRecord(ctx.aux_ev[1], s2));  // near field done
  mark(4);
  // ---- while the device runs P2P: lists 2-4 and the far-field plan (part B2)
  auto t_lists2 = std::chrono::steady_clock::now();
  if (!lists_done) tree_lists_far(tree);
  thread_local Flat FC;
  FC.I.clear(); FC.D.clear();
  build_plan_far(tree, f, FC, O);
  devplan_far(P, O, upload_ints("plan_ic", FC.I, s));
  auto t_lists3 = std::chrono::steady_clock::now();
  form(std::true_type{}, P.n_p2l, lo);
  qp_form(std::true_type{}, P.n_p2l, lo);
  mark(5);
  // ---- STEP 4: M2L
  if (P.n_m2l) {
    // Default: pairs grouped by translation class (k_m2l_grp).
    // FMM3D_M2L_ROTATION=pairs selects one pair at a time with per-class
    // rotation blocks, =factored the factored rotations, =mfma the grouped
    // kernel on the matrix cores.
    static const int m2l_form = [] {
      const char* v = std::getenv("FMM3D_M2L_ROTATION");
      if (v && std::strcmp(v, "factored") == 0) return 2;
      if (v && std::strcmp(v, "pairs") == 0) return 1;
      if (v && std::strcmp(v, "mfma") == 0) return 3;
      return 0;
    }();
    // FMM3D_M2L_ROTATION=mfma: the grouped M2L on the matrix cores
    // (k_m2l_mfma) wherever its LDS fits and p <= 47 (three row tiles);
    // FMM3D_M2L_MFMA_THREADS=512 runs eight waves per workgroup.
    static const int mf_nt = [] {
      const char* v = std::getenv("FMM3D_M2L_MFMA_THREADS");
      return (v && std::strcmp(v, "512") == 0) ? 512 : 256;
    }();
    const size_t mf_lds = m2l_mfma_lds<T>(p);
    if (m2l_form == 3 && p <= 47 && mf_lds <= max_lds_bytes(ctx)) {
      const int mrt = p / 16 + 1;
#define M2L_MFMA(NT_, R_) \
  hipLaunchKernelGGL((k_m2l_mfma<T, NT_, R_>), dim3(P.n_m2l), dim3(NT_), mf_lds, s, P, Tb, mp, lo)
      if (mf_nt == 512) {
        if (mrt == 1) M2L_MFMA(512, 1); else if (mrt == 2) M2L_MFMA(512, 2); else M2L_MFMA(512, 3);
      } else {
        if (mrt == 1) M2L_MFMA(256, 1); else if (mrt == 2) M2L_MFMA(256, 2); else M2L_MFMA(256, 3);
      }
#undef M2L_MFMA
    } else {
      const bool lds = sizeof(T) * 6 * (size_t)K <= max_lds_bytes(ctx);
      const size_t sh = lds ? sizeof(T) * 6 * K : lds_ps;
      // launch shape of k_m2l_grp: (largest group, threads, items per thread);
      // the work items of one stage must fit, the group's expansions must fit LDS
      static const int m2l_shape = [] {
        const char* v = std::getenv("FMM3D_M2L_SHAPE");
        if (v && std::strcmp(v, "g2t256i1") == 0) return 1;
        if (v && std::strcmp(v, "g4t512i1") == 0) return 2;
        return 0;
      }();
      const int gcap = m2l_shape == 1 ? 2 : kM2LGroup;
      const int nt = m2l_shape == 2 ? 512 : 256, items = m2l_shape == 0 ? 2 : 1;
      int nI = 0;
      for (int n = 0; n <= p; n++) nI += n / 4 + 1;
      int gmax = 0;
      // LDS: six int arrays, the local, a guard of 8 values, g members of (p+4)^2
      auto m2l_lds = [&](int g) {
        return (6 * kM2LChunk + (size_t)grp_kn_ints(p)) * sizeof(int) +
               sizeof(T) * (2 * (size_t)kM2LGroup * (p + 1) + 2 * (size_t)K + 8 + (size_t)g * grp_member_stride<T>(p));
      };
      for (int g = gcap; g >= 1 && !gmax; g--) {
        const int gs = g > 2 ? 4 : g;
        if (nI * gs <= items * nt && m2l_lds(g) <= max_lds_bytes(ctx)) gmax = g;
      }
      if ((m2l_form == 0 || m2l_form == 3) && gmax > 0) {
        const size_t need = m2l_lds(gmax);
        if (m2l_shape == 0)
          hipLaunchKernelGGL((k_m2l_grp<T, 256, 2>), dim3(P.n_m2l), dim3(256), need, s, P, Tb, gmax, mp, lo);
        else if (m2l_shape == 1)
          hipLaunchKernelGGL((k_m2l_grp<T, 256, 1>), dim3(P.n_m2l), dim3(256), need, s, P, Tb, gmax, mp, lo);
        else
          hipLaunchKernelGGL((k_m2l_grp<T, 512, 1>), dim3(P.n_m2l), dim3(512), need, s, P, Tb, gmax, mp, lo);
      } else if (m2l_form != 2) {
        if (lds) hipLaunchKernelGGL((k_m2l<T, true>), dim3(P.n_m2l), dim3(256), sh, s, P, Tb, mp, lo);
        else hipLaunchKernelGGL((k_m2l<T, false>), dim3(P.n_m2l), dim3(256), sh, s, P, Tb, mp, lo);
      } else {
        if (lds) hipLaunchKernelGGL((k_m2l_fact<T, true>), dim3(P.n_m2l), dim3(256), sh, s, P, Tb, mp, lo);
        else hipLaunchKernelGGL((k_m2l_fact<T, false>), dim3(P.n_m2l), dim3(256), sh, s, P, Tb, mp, lo);
      }
    }
  }
  mark(6);
  // ---- STEP 5: L2L, top-down
  for (auto& g : O.l2l_groups)
    hipLaunchKernelGGL((k_l2l<T>), dim3(g.second), dim3(256), lds_ps, s, P, Tb, g.first, lo);
  mark(7);
  // ---- STEP 7: L2P (packed local + gradient expansions, wave-uniform reads)
  if (P.n_l2p) {
    const size_t H = (size_t)(p + 1) * (p + 2) / 2;
    T* pack = (T*)ctx.arena.get("l2p_pack", sizeof(T) * 16 * H * (size_t)P.n_l2p);
    const size_t tot = H * (size_t)P.n_l2p;
    hipLaunchKernelGGL((k_l2p_pack<T>), dim3((unsigned)std::min<size_t>(blocks(tot, 256), 65536)), dim3(256), 0, s,
                       P, fac, lo, pack);
    if (P.n_l2pi)
      hipLaunchKernelGGL((k_l2p2<T>), dim3(item_groups(P.n_l2pi)), dim3(kP2PBlock * kItemWaves), 0, s, P, fac, pack, spos, tpos, f.ifpot, f.iffld,
                         f.ifpottarg, f.iffldtarg, pot_s, fld_s, pot_t, fld_t);
  }
  mark(8);
  // ---- STEP 6: M2P (list 3) from packed multipoles; list 3 of each box is
  // cut into slices of about kSlice packed entries, one wave per (64 points,
  // slice), and the slices' sums are added in order (k_slice_reduce)
  if (P.n_m2p && (f.ifcharge || f.ifdipole || qp)) {
    const int grad = (f.iffld || f.iffldtarg) ? 1 : 0;
    const bool want_src = f.ifpot || f.iffld, want_trg = f.ifpottarg || f.iffldtarg;
    const int pd = grad ? p + 1 : p;
    const int Hd = (pd + 1) * (pd + 2) / 2;
    constexpr int kSlice = 1024;
    const int nbx = tree.nboxes;
    thread_local std::vector<int> slot, pk, off, it, red, hv;
    slot.assign(nbx, -1);
    pk.clear(); off.clear(); it.clear(); red.clear(); hv.clear();
    int tot = 0;
    std::vector<int> cuts;
    for (int b = 0; b < nbx; b++) {
      const int l0 = tree.list[2].off[b], l1 = tree.list[2].off[b + 1];
      if (l0 == l1) continue;
      if (!((want_src && tree.nsrc[b] > 0) || (want_trg && tree.ntrg[b] > 0))) continue;
      cuts.assign(1, l0);
      int cost = 0;
      for (int k = l0; k < l1; k++) {
        const int j = tree.list[2].idx[k];
        if (tree.nsrc[j] == 0) continue;
        cost += Hd;
        if (cost >= kSlice && k + 1 < l1) { cuts.push_back(k + 1); cost = 0; }
        if (slot[j] >= 0) continue;
        slot[j] = tot;
        pk.push_back(j);
        off.push_back(tot);
        tot += Hd;
      }
      cuts.push_back(l1);
      const int nsl = (int)cuts.size() - 1;
      auto slices = [&](int kind, int npts) {
        for (int q = 0; q < npts; q += 64) {
          red.push_back(b); red.push_back(kind); red.push_back(q);
          red.push_back((int)(it.size() / 5)); red.push_back(nsl);
          for (int c = 0; c < nsl; c++) {
            it.push_back(b); it.push_back(kind); it.push_back(q);
            it.push_back(cuts[c]); it.push_back(cuts[c + 1]);
          }
        }
      };
      if (want_src) slices(0, tree.nsrc[b]);
      if (want_trg) slices(1, tree.ntrg[b]);
    }
    if (!pk.empty() && !it.empty()) {
      const size_t o_pk = nbx, o_off = o_pk + pk.size(), o_it = o_off + off.size(), o_red = o_it + it.size();
      for (auto* v : {&slot, &pk, &off, &it, &red}) hv.insert(hv.end(), v->begin(), v->end());
      const int* dh = upload_ints("m2p_i", hv, s);
      T* pack = (T*)ctx.arena.get("m2p_pack", sizeof(T) * 16 * (size_t)tot);
      const size_t nit = it.size() / 5;
      double* partial = (double*)ctx.arena.get("m2p_part", sizeof(double) * 8 * 64 * nit);
      const size_t tp = (size_t)pk.size() * Hd;
      hipLaunchKernelGGL((k_m2p_pack<T>), dim3((unsigned)std::min<size_t>(blocks(tp, 256), 65536)), dim3(256), 0, s,
                         P, fac, mp, dh + o_pk, dh + o_off, (int)pk.size(), grad, pack);
      hipLaunchKernelGGL((k_m2p_eval<T>), dim3(item_groups((int)nit)), dim3(kP2PBlock * kItemWaves), 0, s, P, fac,
                         pack, dh, dh + o_it, (int)nit, grad, spos, tpos, partial);
      hipLaunchKernelGGL(k_slice_reduce, dim3((unsigned)(red.size() / 5)), dim3(64), 0, s, dh + o_red, partial,
                         P.isrc, P.nsrc, P.itrg, P.ntrg, f.ifpot, f.iffld, f.ifpottarg, f.iffldtarg, pot_s, fld_s,
                         pot_t, fld_t);
    }
  }
  mark(9);
  // ---- far + near, scattered back to the caller's order (l3dpsort / l3dfsort)
  FMM_HIP(hipStreamWaitEvent(s, ctx.aux_ev[1], 0));
  // ---- Hessians (FP64): near field over the P2P items, far field from the
  // locals (L2P items) and the list-3 multipoles, then scattered like the field
  if (c.ifhess || c.ifhesstarg) {
    if (!(std::is_same<T, double>::value && std::is_same<TN, double>::value))
      throw AllocError("Hessians are computed in FP64 only", 128);
    const int h0 = c.ifhess ? 1 : 0, h1 = c.ifhesstarg ? 1 : 0;
    const double* fac2 = hess_factors(ctx, p);
    double* hs_f = (double*)ctx.arena.get("hess_sf", sizeof(double) * 12 * (size_t)ns);
    double* hs_n = (double*)ctx.arena.get("hess_sn", sizeof(double) * 12 * (size_t)ns);
    double* ht_f = (double*)ctx.arena.get("hess_tf", sizeof(double) * 12 * (size_t)std::max(nt, 1));
    double* ht_n = (double*)ctx.arena.get("hess_tn", sizeof(double) * 12 * (size_t)std::max(nt, 1));
    if (h0) { zero(hs_f, 12 * (size_t)ns); zero(hs_n, 12 * (size_t)ns); }
    if (h1) { zero(ht_f, 12 * (size_t)nt); zero(ht_n, 12 * (size_t)nt); }
    const double* str_h = reinterpret_cast<const double*>(str64);
    auto near_hess = [&](const int* items, int n, int kind, double* out) {
      if (n == 0 || !(f.ifcharge || f.ifdipole)) return;
      DevPlan Q = P;
      Q.p2p_items = items;
      Q.n_p2p = n;
      dim3 g(item_groups(n)), bl(kP2PBlock * kItemWaves);
      if (f.ifcharge && f.ifdipole)
        hipLaunchKernelGGL((k_p2p_hess<true, true>), g, bl, 0, s, Q, spos, str_h, tpos, kind, out);
      else if (f.ifcharge)
        hipLaunchKernelGGL((k_p2p_hess<true, false>), g, bl, 0, s, Q, spos, str_h, tpos, kind, out);
      else
        hipLaunchKernelGGL((k_p2p_hess<false, true>), g, bl, 0, s, Q, spos, str_h, tpos, kind, out);
    };
    if (h0) near_hess(dIB_near + O.p2p0, O.n_p2p0, 0, hs_n);
    if (h1) near_hess(dIB_near + O.p2p1, O.n_p2p1, 1, ht_n);
    if (P.n_l2pi && (f.ifcharge || f.ifdipole))
      hipLaunchKernelGGL(k_l2p_hess, dim3(item_groups(P.n_l2pi)), dim3(kP2PBlock * kItemWaves), 0, s, P, fac2,
                         reinterpret_cast<const double*>(lo), spos, tpos, h0, h1, hs_f, ht_f);
    // list-3 work items: (box, kind, first point) for every box with list 3
    thread_local std::vector<int> hit;
    hit.clear();
    for (int b = 0; b < tree.nboxes; b++) {
      if (tree.list[2].count(b) == 0) continue;
      if (h0)
        for (int q = 0; q < tree.nsrc[b]; q += kP2PBlock) { hit.push_back(b); hit.push_back(0); hit.push_back(q); }
      if (h1)
        for (int q = 0; q < tree.ntrg[b]; q += kP2PBlock) { hit.push_back(b); hit.push_back(1); hit.push_back(q); }
    }
    if (!hit.empty() && (f.ifcharge || f.ifdipole)) {
      const int* dh = upload_ints("hess_m2p", hit, s);
      const int nit = (int)(hit.size() / 3);
      hipLaunchKernelGGL(k_m2p_hess, dim3(item_groups(nit)), dim3(kP2PBlock * kItemWaves), 0, s, P, fac2,
                         reinterpret_cast<const double*>(mp), dh, nit, spos, tpos, hs_f, ht_f);
    }
    if (h0) hipLaunchKernelGGL(k_scatter_sum, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, 6, sperm, hs_f, hs_n,
                               c.hess_dev);
    if (h1) hipLaunchKernelGGL(k_scatter_sum, dim3(blocks(nt, 256)), dim3(256), 0, s, nt, 6, tperm, ht_f, ht_n,
                               c.hesstarg_dev);
  }
  auto scatter = [&](int n, int w, const int* perm, const double* far, const double* near, double* out) {
    hipLaunchKernelGGL(k_scatter_sum, dim3(blocks(n, 256)), dim3(256), 0, s, n, w, perm, far, near, out);
  };
  if (f.ifpot) scatter(ns, 1, sperm, pot_s, pot_sn, c.pot_dev);
  if (f.iffld) scatter(ns, 3, sperm, fld_s, fld_sn, c.fld_dev);
  if (f.ifpottarg) scatter(nt, 1, tperm, pot_t, pot_tn, c.pottarg_dev);
  if (f.iffldtarg) scatter(nt, 3, tperm, fld_t, fld_tn, c.fldtarg_dev);
  mark(10);
  if (c.graph) {
    capture.on = false;
    FMM_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e1 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    const hipError_t e2 = e1 == hipSuccess ? hipGraphLaunch(exec, s) : e1;
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    FMM_HIP(e2);
  }
  FMM_HIP(hipGetLastError());
  FMM_HIP(hipStreamSynchronize(s));
  auto& lt = last_timings();
  lt.assign(13, 0.0);
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  lt[0] = ms(t0, t1);
  // main stream: upload, P2M, M2M, (near-plan gap), P2L, M2L, L2L, L2P, M2P,
  // wait for the near field + scatter; the P2P time is the second stream's
  auto el = timed ? tm.elapsed() : std::vector<double>();
  el.resize(10, 0.0);
  lt[1] = el[0]; lt[2] = el[1]; lt[3] = el[4]; lt[4] = el[2]; lt[5] = el[5];
  lt[6] = el[6]; lt[7] = el[7]; lt[8] = el[8]; lt[9] = timed ? tn.elapsed()[0] : 0.0;
  double tot = 0;
  for (auto v : el) tot += v;
  lt[10] = tot + lt[0];
  lt[11] = ms(t0, t_tree);
  lt[12] = ms(t_lists0, t_lists1) + ms(t_lists2, t_lists3);
  return 0;
}

template int lap_solve<double, double>(Context&, const LapCall&);
template int lap_solve<double, float>(Context&, const LapCall&);
template int lap_solve<float, double>(Context&, const LapCall&);
template int lap_solve<float, float>(Context&, const LapCall&);

// Direct sum on the device (l3dpartdirect): identity order, one pass.
int lap_direct(Context& ctx, const LapCall& c) {
  const int ns = c.ns, nt = c.nt;
  hipStream_t s = ctx.stream;
  const size_t nmax = (size_t)std::max(std::max(ns, nt), 1);
  std::vector<int> id(nmax);
  for (size_t i = 0; i < id.size(); i++) id[i] = (int)i;
  int* dperm = (int*)ctx.arena.get("plan_i", id.size() * sizeof(int));
  FMM_HIP(hipMemcpyAsync(dperm, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, s));
  double* spos = (double*)ctx.arena.get("spos", sizeof(double) * 4 * (size_t)std::max(ns, 1));
  double* str64 = (double*)ctx.arena.get("str64", sizeof(double) * 8 * (size_t)std::max(ns, 1));
  double* tpos = (double*)ctx.arena.get("tpos", sizeof(double) * 4 * (size_t)std::max(nt, 1));
  double* pot = (double*)ctx.arena.get("pot_s", sizeof(double) * 2 * nmax);
  double* fld = (double*)ctx.arena.get("fld_s", sizeof(double) * 6 * nmax);
  if (ns > 0)
    hipLaunchKernelGGL(k_gather_sources, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, dperm, c.src_dev,
                       c.q_dev, c.ds_dev, c.dv_dev, c.ifcharge ? 1 : 0, c.ifdipole ? 1 : 0, spos, str64,
                       (float*)nullptr);
  auto run = [&](int npts, const double* pts, int are_src) {
    if (npts <= 0) return;
    if (ns == 0 || (!c.ifcharge && !c.ifdipole)) {
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(2 * (size_t)npts, 256)), dim3(256), 0, s, 2 * (size_t)npts, pot);
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(6 * (size_t)npts, 256)), dim3(256), 0, s, 6 * (size_t)npts, fld);
      return;
    }
#define DCASE(C, D)                                                                                   \
  if ((c.ifcharge ? 1 : 0) == C && (c.ifdipole ? 1 : 0) == D)                                         \
    hipLaunchKernelGGL((k_direct<C, D>), dim3(blocks(npts, 256)), dim3(256), 0, s, ns, spos, str64, npts, \
                       pts, are_src, pot, fld);
    DCASE(1, 0) DCASE(0, 1) DCASE(1, 1)
#undef DCASE
  };
  auto copy_out = [&](int npts, int wpot, double* po, int wfld, double* fo) {
    // the direct sum runs in the caller's order (identity permutation)
    if (wpot) FMM_HIP(hipMemcpyAsync(po, pot, sizeof(double) * 2 * npts, hipMemcpyDeviceToDevice, s));
    if (wfld) FMM_HIP(hipMemcpyAsync(fo, fld, sizeof(double) * 6 * npts, hipMemcpyDeviceToDevice, s));
  };
  if (ns > 0 && (c.ifpot || c.iffld)) {
    run(ns, nullptr, 1);
    copy_out(ns, c.ifpot, c.pot_dev, c.iffld, c.fld_dev);
  }
  if (nt > 0 && (c.ifpottarg || c.iffldtarg)) {
    hipLaunchKernelGGL(k_gather_targets, dim3(blocks(nt, 256)), dim3(256), 0, s, nt, dperm, c.trg_dev, tpos);
    run(nt, tpos, 0);
    copy_out(nt, c.ifpottarg, c.pottarg_dev, c.iffldtarg, c.fldtarg_dev);
  }
  FMM_HIP(hipGetLastError());
  FMM_HIP(hipStreamSynchronize(s));
  return 0;
}

// l3dtriadirect (src/lfmm3dtria.f:1385) on the device: caller's order, one
// pass over all triangles per point.
int tri_direct(Context& ctx, const LapCall& c) {
  const int n = c.ns, nt = c.nt;
  hipStream_t s = ctx.stream;
  std::vector<int> id((size_t)std::max(n, 1));
  for (size_t i = 0; i < id.size(); i++) id[i] = (int)i;
  int* dperm = (int*)ctx.arena.get("plan_i", id.size() * sizeof(int));
  FMM_HIP(hipMemcpyAsync(dperm, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, s));
  double* geo = (double*)ctx.arena.get("tri_geo", sizeof(double) * kTriGeo * (size_t)std::max(n, 1));
  double* tstr = (double*)ctx.arena.get("tri_str", sizeof(double) * 4 * (size_t)std::max(n, 1));
  const int ifc = c.ifcharge ? 1 : 0, ifd = c.ifdipole ? 1 : 0;
  hipLaunchKernelGGL(k_tri_prep, dim3(blocks(n, 256)), dim3(256), 0, s, n, dperm, c.tri_dev, c.q_dev, c.ds_dev, ifc,
                     ifd, 0, (const double*)nullptr, geo, tstr, (double*)nullptr, (double*)nullptr);
  const size_t nmax = (size_t)std::max(std::max(n, nt), 1);
  double* pot = (double*)ctx.arena.get("pot_s", sizeof(double) * 2 * nmax);
  double* fld = (double*)ctx.arena.get("fld_s", sizeof(double) * 6 * nmax);
  auto run = [&](int npts, const double* pts, int centroids, int wpot, double* po, int wfld, double* fo) {
    if (npts <= 0 || !(wpot || wfld)) return;
    if (!ifc && !ifd) {
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(2 * (size_t)npts, 256)), dim3(256), 0, s,
                         2 * (size_t)npts, pot);
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(6 * (size_t)npts, 256)), dim3(256), 0, s,
                         6 * (size_t)npts, fld);
    } else if (ifc && ifd) {
      hipLaunchKernelGGL((k_tri_direct<true, true>), dim3(blocks(npts, 256)), dim3(256), 0, s, n, geo, tstr, npts,
                         pts, centroids, pot, fld);
    } else if (ifc) {
      hipLaunchKernelGGL((k_tri_direct<true, false>), dim3(blocks(npts, 256)), dim3(256), 0, s, n, geo, tstr, npts,
                         pts, centroids, pot, fld);
    } else {
      hipLaunchKernelGGL((k_tri_direct<false, true>), dim3(blocks(npts, 256)), dim3(256), 0, s, n, geo, tstr, npts,
                         pts, centroids, pot, fld);
    }
    if (wpot) FMM_HIP(hipMemcpyAsync(po, pot, sizeof(double) * 2 * npts, hipMemcpyDeviceToDevice, s));
    if (wfld) FMM_HIP(hipMemcpyAsync(fo, fld, sizeof(double) * 6 * npts, hipMemcpyDeviceToDevice, s));
  };
  run(n, c.src_dev, 1, c.ifpot, c.pot_dev, c.iffld, c.fld_dev);
  run(nt, c.trg_dev, 0, c.ifpottarg, c.pottarg_dev, c.iffldtarg, c.fldtarg_dev);
  FMM_HIP(hipGetLastError());
  FMM_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace fmm3d

namespace fmm3d {
int lap_solve_mode(Context& ctx, const LapCall& c, int mode) {
  mode = c.tri_dev ? 0 : effective_precision(c.iprec, mode);
  last_precision() = mode;
  switch (mode) {
    case 0: return lap_solve<double, double>(ctx, c);
    case 1: return lap_solve<double, float>(ctx, c);
    case 2: return lap_solve<float, double>(ctx, c);
    case 3: return lap_solve<float, float>(ctx, c);
  }
  return 128;
}
}  // namespace fmm3d
