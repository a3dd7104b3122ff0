// Helmholtz particle FMM driver: host octree (the same d3tstrcr tree as the
// Laplace path) -> per-level orders and translation tables -> one upload ->
// the eight steps of hfmm3dparttargmain (src/hfmm3dpart.f:545-1086) and
// hfmm3d_list2 (src/hfmm3drouts.f:153-564) as HIP kernels on one stream ->
// scatter back to the caller's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "fmm3d_internal.hpp"
#include "helm_kernels.hip"
#include "helm_tables.hpp"
#include "htria_kernels.hpp"
#include "tria_rules.hpp"
#include "tree_gpu.hpp"
#include "octree.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace fmm3d {

namespace {

// Host tables and their device copies, cached for repeated calls with the
// same wavenumber, root box, depth and tolerance.
struct HelmCache {
  HelmHostTables H;
  std::vector<int> I;
  std::vector<double> D;
  std::vector<long> Lg;
  size_t o_offrot = 0, o_offdist = 0, o_octrot = 0, o_plev = 0, o_ptrunc = 0;
  size_t o_fac = 0, o_offph = 0, o_octph = 0, o_rot = 0, o_sclev = 0, o_z = 0;
  size_t o_rotoff = 0, o_coax = 0;
  int m2m_base = 0, l2l_base = 0;
  // grouped M2L: rotation tiles and items at pmax, complex coaxial tiles
  size_t o_rothalf = 0, o_rothalf_off = 0, o_gri = 0, o_grc = 0, o_gzi = 0, o_gzo = 0, o_zt = 0, o_ztoff = 0;
  bool grouped_m2l = false;   // every list-2 order fits its table's tiles
};

HelmCache* helm_tables(Context& ctx, cplx zk, double size, int nlev, double eps, int& ier) {
  ier = 0;
  auto* c = static_cast<HelmCache*>(ctx.helm_tables.get());
  const bool hit = c && c->H.zk == zk && c->H.size == size && c->H.nlev == nlev && c->H.eps == eps;
  if (!hit) {
    auto nc = std::make_shared<HelmCache>();
    if (!build_helm_tables(zk, size, nlev, eps, nc->H)) {
      ier = 16;
      return nullptr;
    }
    const HelmHostTables& H = nc->H;
    auto& I = nc->I;
    auto& D = nc->D;
    auto& Lg = nc->Lg;
    auto put_i = [&](const std::vector<int>& v) { size_t o = I.size(); I.insert(I.end(), v.begin(), v.end()); while (I.size() % 16) I.push_back(0); return o; };
    auto put_d = [&](const std::vector<double>& v) { size_t o = D.size(); D.insert(D.end(), v.begin(), v.end()); while (D.size() % 8) D.push_back(0.0); return o; };
    nc->o_offrot = put_i(H.geo.off_rot);
    nc->o_offdist = put_i(H.geo.off_dist);
    nc->o_octrot = put_i(H.geo.oct_rot);
    nc->o_plev = put_i(H.plev);
    nc->o_ptrunc = put_i(H.ptrunc);
    nc->o_fac = put_d(H.geo.fac);
    nc->o_offph = put_d(H.geo.off_phase);
    nc->o_octph = put_d(H.geo.oct_phase);
    nc->o_rot = put_d(H.geo.rot);
    nc->o_sclev = put_d(H.sclev);
    nc->o_z = put_d(H.z);
    nc->o_rotoff = 0;
    Lg = H.geo.rot_off;
    nc->o_coax = Lg.size();
    const int nd = H.geo.ndist;
    auto put_desc = [&](const CoaxDesc& d) { Lg.push_back(d.base); Lg.push_back(d.pin); Lg.push_back(d.pout); };
    for (const auto& d : H.m2l) put_desc(d);
    nc->m2m_base = (nlev + 1) * nd;
    for (const auto& d : H.m2m) put_desc(d);
    nc->l2l_base = nc->m2m_base + nlev + 1;
    for (const auto& d : H.l2l) put_desc(d);
    // grouped M2L: the geometry tables' rotation tiles and work items (order
    // pmax), and each M2L table's coaxial blocks as complex 4x4 tiles per
    // order am: [row block of n - am][column block of l - am], 16 real parts
    // then 16 imaginary parts of [row][column]
    {
      const int P1 = H.pmax + 1;
      nc->o_gri = put_i(H.geo.grp_ritems);
      nc->o_grc = put_i(H.geo.grp_rcount);
      nc->o_gzi = put_i(H.geo.grp_zitems);
      nc->o_gzo = put_i(H.geo.grp_zoff);
      nc->o_rothalf = put_d(H.geo.rot_half);
      const size_t ntab = H.m2l.size();
      std::vector<long> zt_off(ntab * P1, 0);
      std::vector<double> zt;
      bool ok = true;
      const size_t nzc = H.z.size() / 2;   // complex entries
      for (size_t t = 0; t < ntab; t++) {
        // M2L runs from level 2 on; the descriptors of levels 0 and 1 are empty
        if (t < 2 * (size_t)H.geo.ndist) continue;
        const CoaxDesc& d = H.m2l[t];
        if (d.pin != d.pout || d.pout < 0 || d.pout > H.pmax) { ok = false; continue; }
        const int pz = d.pout;
        // the table's last block ends inside the coaxial data
        if ((size_t)(d.base + coax_block(pz, pz, pz) + 1) > nzc) { ok = false; continue; }
        for (int am = 0; am <= pz; am++) {
          const int nb4 = (pz - am) / 4 + 1;
          const size_t b0 = zt.size();
          zt_off[t * P1 + am] = (long)b0;
          zt.resize(b0 + 32 * (size_t)nb4 * nb4, 0.0);
          const long mb = d.base + coax_block(am, d.pin, d.pout);
          for (int n = am; n <= pz; n++)
            for (int l = am; l <= pz; l++) {
              const long ix = mb + (long)(l - am) * (d.pin + 1 - am) + (n - am);
              const int rr = n - am, cc = l - am;
              const size_t tile = b0 + 32 * ((size_t)(rr / 4) * nb4 + cc / 4);
              zt[tile + (rr % 4) * 4 + cc % 4] = H.z[2 * ix];
              zt[tile + 16 + (rr % 4) * 4 + cc % 4] = H.z[2 * ix + 1];
            }
        }
      }
      // every list-2 order must lie within its table's tiles
      const int nd2 = H.geo.ndist;
      for (int l = 2; l <= nlev && ok; l++)
        for (int code = 0; code < 343 && ok; code++) {
          const int pt = H.ptrunc[(size_t)l * 343 + code];
          if (pt <= 0) continue;
          const CoaxDesc& d = H.m2l[(size_t)l * nd2 + H.geo.off_dist[code]];
          ok = d.pin == d.pout && pt <= d.pout;
        }
      nc->grouped_m2l = ok;
      nc->o_zt = put_d(zt);
      nc->o_rothalf_off = Lg.size();
      Lg.insert(Lg.end(), H.geo.rot_half_off.begin(), H.geo.rot_half_off.end());
      nc->o_ztoff = Lg.size();
      Lg.insert(Lg.end(), zt_off.begin(), zt_off.end());
    }
    void* di = ctx.arena.get("htab_i", I.size() * sizeof(int));
    void* dd = ctx.arena.get("htab_d", D.size() * sizeof(double));
    void* dl = ctx.arena.get("htab_l", Lg.size() * sizeof(long));
    FMM_HIP(hipMemcpyAsync(di, I.data(), I.size() * sizeof(int), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipMemcpyAsync(dd, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipMemcpyAsync(dl, Lg.data(), Lg.size() * sizeof(long), hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    ctx.helm_tables = nc;
    c = nc.get();
  }
  return c;
}

template <typename TN, bool CHG, bool DIP, bool FLD>
void hp2p(const DevPlan& P, const int* items, int n, int kind, double kre, double kim, const double* spos,
          const TN* str, const double* str64, const double* tpos, double* pot, double* fo, hipStream_t s) {
  DevPlan Q = P;
  Q.p2p_items = items;
  Q.n_p2p = n;
  hipLaunchKernelGGL((k_hp2p<TN, CHG, DIP, FLD>), dim3(item_groups(n)), dim3(kP2PBlock * kItemWaves), 0, s, Q, kre, kim,
                     spos, str, str64, tpos, kind, pot, fo);
}

template <typename TN>
void launch_hp2p(const DevPlan& P, const int* items, int n, int kind, double kre, double kim,
                 const double* spos, const TN* str, const double* str64, const double* tpos, int ifcharge, int ifdipole,
                 bool fld, double* pot, double* fo, hipStream_t s) {
  if (n == 0) return;
#define HCASE(C, Dp, F) \
  if (ifcharge == C && ifdipole == Dp && fld == F) return hp2p<TN, C, Dp, F>(P, items, n, kind, kre, kim, spos, str, str64, tpos, pot, fo, s);
  HCASE(1, 0, false) HCASE(1, 0, true) HCASE(0, 1, false) HCASE(0, 1, true) HCASE(1, 1, false) HCASE(1, 1, true)
#undef HCASE
}

}  // namespace

// TN: near-field pair arithmetic (double, or float with FP64 offsets and accumulation)
template <typename TN>
int helm_solve(Context& ctx, const LapCall& c) {
  auto t0 = std::chrono::steady_clock::now();
  double eps;
  int nbox;
  if (!iprec_params(c.iprec, c.ns, c.nt, eps, nbox)) return 32;
  const bool tri = c.tri_dev != nullptr;
  const cplx zk(c.zk_re, c.zk_im);
  // triangle FMM: the near-field rule (of iprec unless the caller names one)
  const std::vector<double>* sym = nullptr;
  if (tri) {
    if (!std::is_same<TN, double>::value) return 128;
    nbox = tri_nbox(c.iprec, c.ns, c.nt);
    sym = &tri_sym_rule(c.tri_nqtri > 0 ? c.tri_nqtri : htri_nqtri(c.iprec));
    if (sym->empty() || sym->size() / 3 > (size_t)kSymRuleMax) return 128;
  }
  const int ns = c.ns, nt = c.nt;
  const Flags f{c.ifcharge ? 1 : 0, c.ifdipole ? 1 : 0, c.ifpot ? 1 : 0, c.iffld ? 1 : 0,
                nt > 0 && c.ifpottarg ? 1 : 0, nt > 0 && c.iffldtarg ? 1 : 0};
  thread_local Octree tree;  // storage reused across calls
  const int* dsperm = nullptr;
  const int* dtperm = nullptr;
  // the interaction lists are built after the first kernels are queued
  if (device_tree_enabled()) {
    if (build_octree_device(ctx, ctx.stream, c.src_dev, ns, c.trg_dev, nt, nbox, tree, &dsperm, &dtperm, false) != 0)
      return 4;
  } else if (build_octree(c.src_host, ns, c.trg_host, nt, nbox, tree, false) != 0) {
    return 4;
  }
  auto t_tree = std::chrono::steady_clock::now();
  int ier = 0;
  HelmCache* HC = helm_tables(ctx, zk, tree.size, tree.nlev, eps, ier);
  if (!HC) return ier;
  const HelmHostTables& H = HC->H;
  const int pmax = H.pmax, K = (pmax + 1) * (pmax + 1);
  // LDS budgets of the kernels at this order
  const size_t ldsmax = max_lds_bytes(ctx);
  const int Hh = (pmax + 2) * (pmax + 3) / 2, Fn = pmax + 2;
  const size_t per_src = 2 * (size_t)Hh + 2 * (size_t)Fn + 12;
  if (sizeof(double) * (4 * (size_t)K + per_src) > ldsmax) return 16;
  int ts = (int)((ldsmax / sizeof(double) - 4 * (size_t)K) / per_src);
  ts = std::max(1, std::min(32, ts));
  const size_t lds_form = sizeof(double) * (4 * (size_t)K + (size_t)ts * per_src);
  // per-thread-sum P2M/P2L: half tables up to degree pmax+1 held by 256
  // threads x kHFormItems entries, outputs by 256 x kHFormOut
  // (LDS: 64 KB, or what holds 4 sources at order pmax+1; the kernel picks
  // its tile from what is left after the factor head)
  const bool form2 = Hh <= 256 * kHFormItems && K <= 256 * kHFormOut;
  const size_t lds_form2 = std::max<size_t>((size_t)64 * 1024,
                                            hform2_fac_bytes(pmax) + sizeof(double) * (2 * ((size_t)Hh + Fn) + 8) * 4);
  const int region2 = (int)((lds_form2 - hform2_fac_bytes(pmax)) / sizeof(double));
  if (form2 && lds_form2 > ldsmax) return 16;
  const size_t lds_ps = sizeof(double) * 4 * (size_t)K;
  const bool m2l_lds = sizeof(double) * 6 * (size_t)K <= ldsmax;
  if (lds_ps > ldsmax) return 16;
  thread_local Flat FA, FB, FC;  // plan parts: boxes, near lists, far lists (storage reused)
  FA.I.clear(); FA.D.clear(); FB.I.clear(); FB.D.clear(); FC.I.clear(); FC.D.clear();
  PlanOffsets O;
  build_plan_boxes(tree, f, FA, O, 2);
  auto t1 = std::chrono::steady_clock::now();
  hipStream_t s = ctx.stream;
  Timer tm;
  tm.start(s, 11);
  // plan uploads go through pinned staging buffers so that they do not wait
  // for the kernels already queued
  auto upload_ints = [&](const char* name, const std::vector<int>& v, hipStream_t st) {
    int* h = (int*)ctx.arena.get_host(std::string(name) + "_h", std::max<size_t>(v.size(), 1) * sizeof(int));
    std::memcpy(h, v.data(), v.size() * sizeof(int));
    int* d = (int*)ctx.arena.get(name, std::max<size_t>(v.size(), 1) * sizeof(int));
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
  DevPlan P = make_devplan(tree, pmax, O, dI, dD);
  const int* ti = (const int*)ctx.arena.get("htab_i", HC->I.size() * sizeof(int));
  const double* td = (const double*)ctx.arena.get("htab_d", HC->D.size() * sizeof(double));
  const long* tl = (const long*)ctx.arena.get("htab_l", HC->Lg.size() * sizeof(long));
  DevTables Tb{};
  Tb.p = pmax;
  Tb.off_rot = ti + HC->o_offrot; Tb.off_dist = ti + HC->o_offdist; Tb.oct_rot = ti + HC->o_octrot;
  Tb.fac = td + HC->o_fac; Tb.off_phase = td + HC->o_offph; Tb.oct_phase = td + HC->o_octph;
  Tb.rot = td + HC->o_rot; Tb.rot_off = tl + HC->o_rotoff;
  HelmDev Hd;
  Hd.kre = zk.real(); Hd.kim = zk.imag();
  Hd.nd = H.geo.ndist;
  Hd.plev = ti + HC->o_plev; Hd.sclev = td + HC->o_sclev; Hd.ptrunc = ti + HC->o_ptrunc;
  Hd.coax = tl + HC->o_coax; Hd.z = td + HC->o_z;
  Hd.rot_half = td + HC->o_rothalf; Hd.rot_half_off = tl + HC->o_rothalf_off;
  Hd.grp_ritems = ti + HC->o_gri; Hd.grp_rcount = ti + HC->o_grc;
  Hd.grp_zitems = ti + HC->o_gzi; Hd.grp_zoff = ti + HC->o_gzo;
  Hd.zt = td + HC->o_zt; Hd.zt_off = tl + HC->o_ztoff;
  const size_t nb = tree.nboxes;
  double* spos = (double*)ctx.arena.get("spos", sizeof(double) * 4 * (size_t)ns);
  double* str64 = (double*)ctx.arena.get("str64", sizeof(double) * 8 * (size_t)ns);
  constexpr bool near32 = sizeof(TN) < sizeof(double);
  float* str32 = near32 ? (float*)ctx.arena.get("str32", sizeof(float) * 8 * (size_t)ns) : nullptr;
  const TN* strn = near32 ? (const TN*)(const void*)str32 : (const TN*)(const void*)str64;
  double* tpos = (double*)ctx.arena.get("tpos", sizeof(double) * 4 * (size_t)std::max(nt, 1));
  double* mp = (double*)ctx.arena.get("mp", sizeof(double) * 2 * K * nb, 16);
  double* lo = (double*)ctx.arena.get("lo", sizeof(double) * 2 * K * nb, 16);
  double* pot_s = (double*)ctx.arena.get("pot_s", sizeof(double) * 2 * (size_t)ns);
  double* fld_s = (double*)ctx.arena.get("fld_s", sizeof(double) * 6 * (size_t)ns);
  double* pot_t = (double*)ctx.arena.get("pot_t", sizeof(double) * 2 * (size_t)std::max(nt, 1));
  double* fld_t = (double*)ctx.arena.get("fld_t", sizeof(double) * 6 * (size_t)std::max(nt, 1));
  double* pot_sn = (double*)ctx.arena.get("pot_sn", sizeof(double) * 2 * (size_t)ns);
  double* fld_sn = (double*)ctx.arena.get("fld_sn", sizeof(double) * 6 * (size_t)ns);
  double* pot_tn = (double*)ctx.arena.get("pot_tn", sizeof(double) * 2 * (size_t)std::max(nt, 1));
  double* fld_tn = (double*)ctx.arena.get("fld_tn", sizeof(double) * 6 * (size_t)std::max(nt, 1));
  const int* sperm = dsperm ? dsperm : dI + O.sperm;
  const int* tperm = dsperm ? dtperm : dI + O.tperm;
  P.src_perm = sperm;
  P.trg_perm = tperm;
  // (triangles: the point strengths are not used -- k_tri_prep forms the
  // triangle densities and quadrature points -- and no dipole vectors are read)
  hipLaunchKernelGGL(k_gather_sources, dim3(blocks(ns, 256)), dim3(256), 0, s, ns, sperm, c.src_dev,
                     c.q_dev, c.ds_dev, c.dv_dev, tri ? 0 : f.ifcharge, tri ? 0 : f.ifdipole, spos, str64, str32);
  if (nt > 0)
    hipLaunchKernelGGL(k_gather_targets, dim3(blocks(nt, 256)), dim3(256), 0, s, nt, tperm, c.trg_dev, tpos);
  auto zero = [&](double* a, size_t n) {
    hipLaunchKernelGGL((k_fill_zero<double>), dim3(std::min<size_t>(blocks(n, 256), 4096)), dim3(256), 0, s, n, a);
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
  // triangle FMM: frames and densities in sorted order, and the quadrature
  // points that stand for the triangles in P2M / P2L (their box ranges
  // counted in points)
  const double* fpos = spos;
  const double* fstr = str64;
  double *tgeo = nullptr, *tstr = nullptr;
  int* tri_rng = nullptr;
  if (tri) {
    const std::vector<double> rule = tri_rule(c.iprec, true);
    const int nq = (int)rule.size() / 3;
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
  tm.mark(1);
  const double* fac = Tb.fac;
  auto form = [&](auto irreg, int nbx, double* out) {
    constexpr bool IR = decltype(irreg)::value;
    if (!nbx) return;
    DevPlan PF = P;  // the current plan (P2L needs list 4), source ranges in points for triangles
    if (tri) {
      PF.isrc = tri_rng;
      PF.nsrc = tri_rng + nb;
    }
    if (!form2) {
      hipLaunchKernelGGL((k_hform<IR>), dim3(nbx), dim3(256), lds_form, s, PF, Hd, fac, fpos, fstr,
                         f.ifcharge, f.ifdipole, ts, out);
    } else if (f.ifcharge && f.ifdipole) {
      hipLaunchKernelGGL((k_hform2<IR, true, true>), dim3(nbx), dim3(256), lds_form2, s, PF, Hd, fac, fpos, fstr, region2, out);
    } else if (f.ifcharge) {
      hipLaunchKernelGGL((k_hform2<IR, true, false>), dim3(nbx), dim3(256), lds_form2, s, PF, Hd, fac, fpos, fstr, region2, out);
    } else if (f.ifdipole) {
      hipLaunchKernelGGL((k_hform2<IR, false, true>), dim3(nbx), dim3(256), lds_form2, s, PF, Hd, fac, fpos, fstr, region2, out);
    }
  };
  // STEP 1: P2M (hfmm3dpart.f:681-757)
  form(std::false_type{}, P.n_p2m, mp);
  tm.mark(2);
  // STEP 3: M2M, bottom-up
  for (size_t g = 0; g < O.m2m_groups.size(); g++)
    hipLaunchKernelGGL(k_hm2m, dim3(O.m2m_groups[g].second), dim3(256), lds_ps, s, P, Tb, Hd,
                       O.m2m_groups[g].first, HC->m2m_base + O.m2m_group_level[g], mp);
  tm.mark(3);
  // while the device runs P2M and M2M: list 1 and the P2P plan
  auto t_lists0 = std::chrono::steady_clock::now();
  tree_lists_near(tree);
  build_plan_near(tree, f, FB, O);
  // the P2P work items are addressed explicitly: they must lie in this part
  if (O.p2p0 + 3 * (size_t)O.n_p2p0 > FB.I.size() || O.p2p1 + 3 * (size_t)O.n_p2p1 > FB.I.size())
    throw HipError("plan: P2P work items outside the near-field part of the plan");
  auto t_lists1 = std::chrono::steady_clock::now();
  // STEP 8: P2P (hfmm3dpart.f:1101-1180) on the second stream, once the
  // inputs are gathered; its plan part is uploaded on that stream
  FMM_HIP(hipStreamWaitEvent(s2, ctx.aux_ev[0], 0));
  Timer tn;
  tn.start(s2, 2);
  const int* dIB = upload_ints("plan_ib", FB.I, s2);
  devplan_near(P, O, dIB);
  if (tri && c.far_only) {
    // the far field alone
  } else if (tri && (f.ifcharge || f.ifdipole)) {
    double* hr = (double*)ctx.arena.get_host("htri_rule_h", sizeof(double) * 3 * kSymRuleMax);
    std::memcpy(hr, sym->data(), sizeof(double) * sym->size());
    double* dr = (double*)ctx.arena.get("htri_rule", sizeof(double) * 3 * kSymRuleMax);
    FMM_HIP(hipMemcpyAsync(dr, hr, sizeof(double) * sym->size(), hipMemcpyHostToDevice, s2));
    const HtriParams hp{zk.real(), zk.imag(), std::abs(zk), c.tri_lapc, c.tri_lapd, (int)sym->size() / 3, dr};
    auto tri_near = [&](const int* items, int n, int kind, double* po, double* fo) {
      launch_htri_p2p(P, items, n, f.ifcharge, f.ifdipole, tgeo, tstr, spos, tpos, kind, hp, po, fo, s2);
    };
    tri_near(dIB + O.p2p0, O.n_p2p0, 0, pot_sn, f.iffld ? fld_sn : nullptr);
    tri_near(dIB + O.p2p1, O.n_p2p1, 1, pot_tn, f.iffldtarg ? fld_tn : nullptr);
  } else if (f.ifcharge || f.ifdipole) {
    launch_hp2p<TN>(P, dIB + O.p2p0, O.n_p2p0, 0, Hd.kre, Hd.kim, spos, strn, str64, tpos, f.ifcharge, f.ifdipole,
                    f.iffld != 0, pot_sn, fld_sn, s2);
    launch_hp2p<TN>(P, dIB + O.p2p1, O.n_p2p1, 1, Hd.kre, Hd.kim, spos, strn, str64, tpos, f.ifcharge, f.ifdipole,
                    f.iffldtarg != 0, pot_tn, fld_tn, s2);
  }
  tn.mark(1);
  FMM_HIP(hipEventRecord(ctx.aux_ev[1], s2));  // near field done
  tm.mark(4);
  // while the device runs P2P: lists 2-4 and the far-field plan
  auto t_lists2 = std::chrono::steady_clock::now();
  tree_lists_far(tree);
  build_plan_far(tree, f, FC, O);
  devplan_far(P, O, upload_ints("plan_ic", FC.I, s));
  auto t_lists3 = std::chrono::steady_clock::now();
  // STEP 2: P2L (hfmm3dpart.f:765-830)
  form(std::true_type{}, P.n_p2l, lo);
  tm.mark(5);
  // STEP 4: M2L, grouped by translation class (k_hm2l_grp) when the group's
  // expansions fit LDS and a stage's work items fit one pass, else pair by
  // pair (k_hm2l)
  if (P.n_m2l) {
    int nI = 0;
    for (int n = 0; n <= pmax; n++) nI += n / 4 + 1;
    auto grp_lds = [&](int g) {
      return (6 * kM2LChunk + (size_t)grp_kn_ints(pmax)) * sizeof(int) +
             sizeof(double) * (2 * (size_t)kM2LGroup * (pmax + 1) + 2 * (size_t)K + 8 +
                               (size_t)g * grp_member_stride<double>(pmax));
    };
    int gmax = 0;
    for (int g = kM2LGroup; g >= 1 && !gmax && HC->grouped_m2l; g--) {
      const int gs = g > 2 ? 4 : g;
      if (nI * gs <= 256 && grp_lds(g) <= ldsmax) gmax = g;
    }
    // (one work item per thread: with two, four members' expansions take
    // ~92 KB of LDS at pmax 23 and leave one workgroup per CU)
    if (gmax > 0)
      hipLaunchKernelGGL((k_hm2l_grp<256, 1>), dim3(P.n_m2l), dim3(256), grp_lds(gmax), s, P, Tb, Hd, gmax, mp, lo);
    else if (m2l_lds)
      hipLaunchKernelGGL((k_hm2l<true>), dim3(P.n_m2l), dim3(256), sizeof(double) * 6 * (size_t)K, s, P, Tb, Hd, mp, lo);
    else
      hipLaunchKernelGGL((k_hm2l<false>), dim3(P.n_m2l), dim3(256), lds_ps, s, P, Tb, Hd, mp, lo);
  }
  tm.mark(6);
  // STEP 5: L2L, top-down
  for (size_t g = 0; g < O.l2l_groups.size(); g++)
    hipLaunchKernelGGL(k_hl2l, dim3(O.l2l_groups[g].second), dim3(256), lds_ps, s, P, Tb, Hd,
                       O.l2l_groups[g].first, HC->l2l_base + O.l2l_group_level[g], lo);
  tm.mark(7);
  // STEP 7: L2P and STEP 6: M2P (list 3), from packed expansions
  {
    const int grad = (f.iffld || f.iffldtarg) ? 1 : 0;
    const bool want_src = f.ifpot || f.iffld, want_trg = f.ifpottarg || f.iffldtarg;
    const int nbx = tree.nboxes;
    // host side: boxes to pack, their entry offsets, the box -> offset maps and
    // the work items (box, kind, first point) in chunks of 64 points
    std::vector<int> hv;                  // m2p slots | l2p slots | m2p pack | l2p pack | items
    std::vector<int> slot_m(nbx, -1), slot_l(nbx, -1), pk_m, off_m, pk_l, off_l, it_m, it_l, red_m;
    int tot_m = 0, tot_l = 0, pdmax = 1;
    auto pdof = [&](int lev) { return H.plev[lev] + grad; };
    auto entries = [&](int lev) { const int pd = pdof(lev); return (pd + 1) * (pd + 2) / 2; };
    auto items = [&](std::vector<int>& it, int b) {
      if (want_src)
        for (int q = 0; q < tree.nsrc[b]; q += 64) { it.push_back(b); it.push_back(0); it.push_back(q); }
      if (want_trg)
        for (int q = 0; q < tree.ntrg[b]; q += 64) { it.push_back(b); it.push_back(1); it.push_back(q); }
    };
    for (int b = 0; b < nbx; b++) {
      const bool has_pts = (want_src && tree.nsrc[b] > 0) || (want_trg && tree.ntrg[b] > 0);
      if (tree.nchild[b] == 0 && tree.level[b] >= 2 && has_pts) {
        slot_l[b] = tot_l;
        pk_l.push_back(b);
        off_l.push_back(tot_l);
        tot_l += entries(tree.level[b]);
        pdmax = std::max(pdmax, pdof(tree.level[b]));
        items(it_l, b);
      }
      if (tree.list[2].count(b) > 0 && has_pts) {
        // list 3 in slices of about kSlice half-table entries: one workgroup per
        // (64 points, slice), the slices summed in order by k_slice_reduce
        constexpr int kSlice = 1024;
        std::vector<int> cuts = {tree.list[2].off[b]};
        int cost = 0;
        for (int k = tree.list[2].off[b]; k < tree.list[2].off[b + 1]; k++) {
          const int j = tree.list[2].idx[k];
          if (tree.level[j] < 2 || tree.nsrc[j] == 0) continue;
          cost += entries(tree.level[j]);
          if (cost >= kSlice && k + 1 < tree.list[2].off[b + 1]) { cuts.push_back(k + 1); cost = 0; }
          if (slot_m[j] >= 0) continue;
          slot_m[j] = tot_m;
          pk_m.push_back(j);
          off_m.push_back(tot_m);
          tot_m += entries(tree.level[j]);
          pdmax = std::max(pdmax, pdof(tree.level[j]));
        }
        cuts.push_back(tree.list[2].off[b + 1]);
        const int nsl = (int)cuts.size() - 1;
        auto slices = [&](int kind, int npts) {
          for (int q = 0; q < npts; q += 64) {
            red_m.push_back(b); red_m.push_back(kind); red_m.push_back(q);
            red_m.push_back((int)(it_m.size() / 5)); red_m.push_back(nsl);
            for (int c = 0; c < nsl; c++) {
              it_m.push_back(b); it_m.push_back(kind); it_m.push_back(q);
              it_m.push_back(cuts[c]); it_m.push_back(cuts[c + 1]);
            }
          }
        };
        if (want_src) slices(0, tree.nsrc[b]);
        if (want_trg) slices(1, tree.ntrg[b]);
      }
    }
    size_t o_sm = 0, o_sl = nbx, o_pm = 2 * (size_t)nbx, o_om = o_pm + pk_m.size(), o_pl = o_om + off_m.size(),
           o_ol = o_pl + pk_l.size(), o_im = o_ol + off_l.size(), o_il = o_im + it_m.size(),
           o_rm = o_il + it_l.size();
    hv.reserve(o_rm + red_m.size());
    for (auto* v : {&slot_m, &slot_l, &pk_m, &off_m, &pk_l, &off_l, &it_m, &it_l, &red_m})
      hv.insert(hv.end(), v->begin(), v->end());
    (void)o_sm;
    int* hh = (int*)ctx.arena.get_host("heval_i_h", std::max<size_t>(hv.size(), 1) * sizeof(int));
    std::memcpy(hh, hv.data(), hv.size() * sizeof(int));
    int* dh = (int*)ctx.arena.get("heval_i", std::max<size_t>(hv.size(), 1) * sizeof(int));
    FMM_HIP(hipMemcpyAsync(dh, hh, hv.size() * sizeof(int), hipMemcpyHostToDevice, s));
    const size_t lds3 = sizeof(double) * 2 * 64 * (size_t)(pdmax + 2);
    if (lds3 > ldsmax) return 16;
    if (!it_l.empty()) {
      double* pack = (double*)ctx.arena.get("hpack_l", sizeof(double) * 16 * (size_t)std::max(tot_l, 1));
      hipLaunchKernelGGL((k_hpack<false>), dim3((unsigned)pk_l.size()), dim3(256), 0, s, P, Hd, fac, lo, dh + o_pl,
                         dh + o_ol, grad, pack);
      hipLaunchKernelGGL((k_heval3<false>), dim3((unsigned)(it_l.size() / 3)), dim3(64), lds3, s, P, Hd, fac, pack,
                         dh + o_sl, dh + o_il, grad, spos, tpos, f.ifpot, f.iffld, f.ifpottarg, f.iffldtarg, pot_s,
                         fld_s, pot_t, fld_t, (double*)nullptr);
    }
    tm.mark(8);
    if (!it_m.empty() && !pk_m.empty()) {
      double* pack = (double*)ctx.arena.get("hpack_m", sizeof(double) * 16 * (size_t)std::max(tot_m, 1));
      hipLaunchKernelGGL((k_hpack<true>), dim3((unsigned)pk_m.size()), dim3(256), 0, s, P, Hd, fac, mp, dh + o_pm,
                         dh + o_om, grad, pack);
      const size_t nsub = it_m.size() / 5;
      double* partial = (double*)ctx.arena.get("hpart_m", sizeof(double) * 8 * 64 * nsub);
      hipLaunchKernelGGL((k_heval3<true>), dim3((unsigned)nsub), dim3(64), 0, s, P, Hd, fac, pack, dh + o_sm,
                         dh + o_im, grad, spos, tpos, f.ifpot, f.iffld, f.ifpottarg, f.iffldtarg, pot_s, fld_s, pot_t,
                         fld_t, partial);
      hipLaunchKernelGGL(k_slice_reduce, dim3((unsigned)(red_m.size() / 5)), dim3(64), 0, s, dh + o_rm, partial,
                         P.isrc, P.nsrc, P.itrg, P.ntrg, f.ifpot, f.iffld, f.ifpottarg, f.iffldtarg, pot_s, fld_s,
                         pot_t, fld_t);
    }
  }
  tm.mark(9);
  // far + near, scattered back to the caller's order (h3dpsort / h3dfsort)
  FMM_HIP(hipStreamWaitEvent(s, ctx.aux_ev[1], 0));
  auto scatter = [&](int n, int w, const int* perm, const double* far, const double* near, double* out) {
    hipLaunchKernelGGL(k_scatter_sum, dim3(blocks(n, 256)), dim3(256), 0, s, n, w, perm, far, near, out);
  };
  if (f.ifpot) scatter(ns, 1, sperm, pot_s, pot_sn, c.pot_dev);
  if (f.iffld) scatter(ns, 3, sperm, fld_s, fld_sn, c.fld_dev);
  if (f.ifpottarg) scatter(nt, 1, tperm, pot_t, pot_tn, c.pottarg_dev);
  if (f.iffldtarg) scatter(nt, 3, tperm, fld_t, fld_tn, c.fldtarg_dev);
  tm.mark(10);
  FMM_HIP(hipGetLastError());
  FMM_HIP(hipStreamSynchronize(s));
  auto& lt = last_timings();
  lt.assign(13, 0.0);
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  lt[0] = ms(t0, t1);
  // main stream: upload, P2M, M2M, (near-plan gap), P2L, M2L, L2L, L2P, M2P,
  // wait for the near field + scatter; the P2P time is the second stream's
  auto el = tm.elapsed();
  el.resize(10, 0.0);
  lt[1] = el[0]; lt[2] = el[1]; lt[3] = el[4]; lt[4] = el[2]; lt[5] = el[5];
  lt[6] = el[6]; lt[7] = el[7]; lt[8] = el[8]; lt[9] = tn.elapsed()[0];
  double tot = 0;
  for (auto v : el) tot += v;
  lt[10] = tot + lt[0];
  lt[11] = ms(t0, t_tree);
  lt[12] = ms(t_lists0, t_lists1) + ms(t_lists2, t_lists3);
  return 0;
}

int helm_solve_mode(Context& ctx, const LapCall& c, int mode) {
  mode = c.tri_dev ? 0 : effective_precision(c.iprec, mode);
  last_precision() = mode;
  switch (mode) {
    case 0: return helm_solve<double>(ctx, c);
    case 1: return helm_solve<float>(ctx, c);
  }
  return 128;
}

// h3dpartdirect on the device: identity order, one pass.
int helm_direct(Context& ctx, const LapCall& c) {
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
#define HDCASE(C, D)                                                                                       \
  if ((c.ifcharge ? 1 : 0) == C && (c.ifdipole ? 1 : 0) == D)                                              \
    hipLaunchKernelGGL((k_hdirect<C, D>), dim3(blocks(npts, 256)), dim3(256), 0, s, c.zk_re, c.zk_im, ns, \
                       spos, str64, npts, pts, are_src, pot, fld);
    HDCASE(1, 0) HDCASE(0, 1) HDCASE(1, 1)
#undef HDCASE
  };
  auto copy_out = [&](int npts, int wpot, double* po, int wfld, double* fo) {
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

// h3dtriadirect (src/hfmm3dtria.f:1456) on the device: caller's order, one
// pass over all triangles per point, the near-field rule of degree
// c.tri_nqtri.
int htri_direct(Context& ctx, const LapCall& c) {
  const int n = c.ns, nt = c.nt;
  const std::vector<double>& sr = tri_sym_rule(c.tri_nqtri);
  if (sr.empty() || sr.size() / 3 > (size_t)kSymRuleMax) return 128;
  hipStream_t s = ctx.stream;
  double* hr = (double*)ctx.arena.get_host("htri_rule_h", sizeof(double) * 3 * kSymRuleMax);
  std::memcpy(hr, sr.data(), sizeof(double) * sr.size());
  double* dr = (double*)ctx.arena.get("htri_rule", sizeof(double) * 3 * kSymRuleMax);
  FMM_HIP(hipMemcpyAsync(dr, hr, sizeof(double) * sr.size(), hipMemcpyHostToDevice, s));
  const cplx zk(c.zk_re, c.zk_im);
  const HtriParams hp{zk.real(), zk.imag(), std::abs(zk), c.tri_lapc, c.tri_lapd, (int)sr.size() / 3, dr};
  std::vector<int> id((size_t)std::max(n, 1));
  for (size_t i = 0; i < id.size(); i++) id[i] = (int)i;
  int* dperm = (int*)ctx.arena.get("plan_i", id.size() * sizeof(int));
  FMM_HIP(hipMemcpyAsync(dperm, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, s));
  double* geo = (double*)ctx.arena.get("tri_geo", sizeof(double) * kTriGeo * (size_t)std::max(n, 1));
  double* tstr = (double*)ctx.arena.get("tri_str", sizeof(double) * 4 * (size_t)std::max(n, 1));
  const int ifc = c.ifcharge ? 1 : 0, ifd = c.ifdipole ? 1 : 0;
  if (n > 0)
    hipLaunchKernelGGL(k_tri_prep, dim3(blocks(n, 256)), dim3(256), 0, s, n, dperm, c.tri_dev, c.q_dev, c.ds_dev,
                       ifc, ifd, 0, (const double*)nullptr, geo, tstr, (double*)nullptr, (double*)nullptr);
  const size_t nmax = (size_t)std::max(std::max(n, nt), 1);
  double* pot = (double*)ctx.arena.get("pot_s", sizeof(double) * 2 * nmax);
  double* fld = (double*)ctx.arena.get("fld_s", sizeof(double) * 6 * nmax);
  auto run = [&](int npts, const double* pts, int centroids, int wpot, double* po, int wfld, double* fo) {
    if (npts <= 0 || !(wpot || wfld)) return;
    if (n == 0 || (!ifc && !ifd)) {
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(2 * (size_t)npts, 256)), dim3(256), 0, s,
                         2 * (size_t)npts, pot);
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(blocks(6 * (size_t)npts, 256)), dim3(256), 0, s,
                         6 * (size_t)npts, fld);
    } else {
      launch_htri_direct(n, geo, tstr, npts, pts, centroids, hp, ifc, ifd, pot, fld, s);
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
