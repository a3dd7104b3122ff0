// Request server: a persistent loop that accepts FMM evaluations of any size,
// kernel and precision as they arrive and runs several of them at once on the
// GPU.  Each worker owns a context (its own HIP stream, grow-only device
// arena and cached translation tables, so trees, expansions and tables stay
// resident in HBM between requests); a request is admitted when a worker is
// free and its device-memory estimate fits the remaining budget, so requests
// that arrive while others run start as soon as a slot and memory free up and
// their kernels overlap on the device.  Latency is measured from submission
// to completion.
//
// Small Laplace requests (at most kFuseMaxPoints points) that are queued
// together with the same tolerance, precision and output flags are fused: a
// worker takes up to kFuseMaxBatch of them, builds their trees on the host,
// joins them into one forest (merge_forest) and runs all of them through one
// set of launches, captured in a HIP graph and launched as one.  Every
// request's boxes, lists and per-box work are those of the same call made
// alone, so its results are bitwise the same.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <string>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/fmm3d_mi355x.h"
#include "fmm3d_internal.hpp"
#include "octree.hpp"
#include "plan.hpp"

namespace fmm3d {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  long long id = 0;
  int kernel = 0, prec = 0, iprec = 1;
  double zk[2] = {0, 0};
  int ns = 0, nt = 0, ifcharge = 0, ifdipole = 0, ifpot = 0, iffld = 0, ifpottarg = 0, iffldtarg = 0;
  std::vector<double> src, q, ds, dv, trg;        // inputs (copied at submission)
  std::vector<double> pot, fld, pt, ft;          // outputs
  size_t bytes = 0;                              // device-memory estimate
  int bypassed = 0;                              // times a later request was admitted first
  int ier = -1;
  bool done = false;
  Clock::time_point t_submit, t_start, t_end;
};

// Device bytes a request needs: inputs, sorted copies and outputs per point,
// plus multipole and local expansions for the boxes a tree with leaves of
// about nbox points has, at the order of the request's tolerance.
size_t estimate_bytes(const Request& r) {
  static const int nb[] = {40, 50, 80, 160, 400, 800, 1200, 1400, 0};
  static const int pl[] = {2, 5, 9, 14, 26, 39, 51, 64, 1};
  const int k = std::min(std::max(r.iprec + 2, 0), 8);
  const size_t npts = (size_t)r.ns + r.nt;
  // in, sorted, strengths, outputs, far and near sums
  const size_t per_point = 24 + 32 + 64 + 32 + 16 + 48 + 16 + 48 + 16 + 48;
  size_t leaves = nb[k] ? npts / nb[k] + 1 : 1;
  size_t boxes = leaves * 8 / 7 + 1;
  int p = pl[k];
  if (r.kernel == 1) p = std::max(p, 40);   // Helmholtz orders grow with |k| size
  const size_t expn = 2 * 16 * (size_t)(p + 1) * (p + 1) * boxes;
  return npts * per_point + expn + (64u << 20);
}

struct Worker {
  Context ctx;
  std::thread th;
};

// admissions of later requests before the queue head is served first
constexpr int kMaxBypass = 8;

// fusion of small Laplace requests: per-request size, requests per batch,
// points per batch
constexpr int kFuseMaxPoints = 65536;
constexpr int kFuseMaxBatch = 16;
constexpr int kFuseMaxTotal = 262144;

bool fusable(const Request& r) {
  return r.kernel == 0 && r.iprec <= 5 && (long)r.ns + r.nt <= kFuseMaxPoints;
}
// requests that can share one forest: same tolerance (hence order and leaf
// size), precision mode and output flags
bool same_batch(const Request& a, const Request& b) {
  return a.iprec == b.iprec && a.prec == b.prec && a.ifcharge == b.ifcharge && a.ifdipole == b.ifdipole &&
         a.ifpot == b.ifpot && a.iffld == b.iffld && (a.nt > 0) == (b.nt > 0) && a.ifpottarg == b.ifpottarg &&
         a.iffldtarg == b.iffldtarg;
}

bool fusion_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("FMM3D_SERVER_FUSE");
    return !(v && std::strcmp(v, "0") == 0);
  }();
  return on;
}

}  // namespace

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<long long, std::shared_ptr<Request>> all;
  std::vector<std::unique_ptr<Worker>> workers;
  size_t budget = 0, in_use = 0;
  long long next_id = 0;
  bool stop = false;
  int busy = 0;
  long long fused = 0;  // requests served in fused batches

  void run(Worker* w) {
    if (hipSetDevice(w->ctx.device) != hipSuccess) (void)hipGetLastError();
    for (;;) {
      std::shared_ptr<Request> r;
      std::vector<std::shared_ptr<Request>> batch;
      {
        std::unique_lock<std::mutex> lk(mu);
        // admit the oldest queued request whose estimate fits what is left of
        // the budget (backfill: smaller requests behind a large one start
        // when it does not fit); a request larger than the whole budget runs
        // alone; once the head has been passed over kMaxBypass times, only
        // the head is admitted, so a large request is not starved
        auto pick = [&]() -> std::deque<std::shared_ptr<Request>>::iterator {
          if (queue.empty()) return queue.end();
          const auto head = queue.begin();
          if (in_use + (*head)->bytes <= budget || in_use == 0) return head;
          if ((*head)->bypassed >= kMaxBypass) return queue.end();
          for (auto it = std::next(head); it != queue.end(); ++it)
            if (in_use + (*it)->bytes <= budget) return it;
          return queue.end();
        };
        std::deque<std::shared_ptr<Request>>::iterator sel;
        cv_work.wait(lk, [&] {
          if (stop && queue.empty()) return true;
          sel = pick();
          return sel != queue.end();
        });
        if (stop && queue.empty()) return;
        if (sel != queue.begin()) queue.front()->bypassed++;
        r = *sel;
        queue.erase(sel);
        in_use += r->bytes;
        busy++;
        batch.push_back(r);
        // small Laplace requests queued with the same parameters join it
        if (fusion_enabled() && fusable(*r)) {
          long pts = (long)r->ns + r->nt;
          for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < kFuseMaxBatch;) {
            const Request& q = **it;
            if (fusable(q) && same_batch(*r, q) && pts + q.ns + q.nt <= kFuseMaxTotal &&
                in_use + q.bytes <= budget) {
              pts += q.ns + q.nt;
              in_use += q.bytes;
              batch.push_back(*it);
              it = queue.erase(it);
            } else {
              ++it;
            }
          }
        }
        const auto now = Clock::now();
        for (auto& b : batch) b->t_start = now;
      }
      int ier = 0;
      try {
        ier = batch.size() > 1 ? solve_fused(w->ctx, batch) : solve(w->ctx, *r);
      } catch (const AllocError& e) {
        ier = e.ier;
      } catch (const HipError&) {
        (void)hipGetLastError();
        ier = 64;
      } catch (const std::bad_alloc&) {
        ier = 8;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        const auto now = Clock::now();
        for (auto& b : batch) {
          b->ier = ier;
          b->t_end = now;
          b->done = true;
          in_use -= b->bytes;
        }
        if (batch.size() > 1) fused += (long long)batch.size();
        busy--;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  static int solve(Context& ctx, Request& r) {
    auto up = [&](const char* name, const std::vector<double>& h) -> double* {
      if (h.empty()) return nullptr;
      double* d = (double*)ctx.arena.get(name, h.size() * sizeof(double));
      FMM_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      return d;
    };
    LapCall c;
    c.iprec = r.iprec; c.ns = r.ns; c.nt = r.nt;
    c.ifcharge = r.ifcharge; c.ifdipole = r.ifdipole; c.ifpot = r.ifpot; c.iffld = r.iffld;
    c.ifpottarg = r.nt > 0 && r.ifpottarg; c.iffldtarg = r.nt > 0 && r.iffldtarg;
    c.zk_re = r.zk[0]; c.zk_im = r.zk[1];
    c.src_host = r.src.data();
    c.trg_host = r.nt > 0 ? r.trg.data() : nullptr;
    c.src_dev = up("in_src", r.src);
    c.q_dev = c.ifcharge ? up("in_q", r.q) : nullptr;
    c.ds_dev = c.ifdipole ? up("in_ds", r.ds) : nullptr;
    c.dv_dev = c.ifdipole ? up("in_dv", r.dv) : nullptr;
    c.trg_dev = r.nt > 0 ? up("in_trg", r.trg) : nullptr;
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("out_pot", 2 * sizeof(double) * r.ns);
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("out_fld", 6 * sizeof(double) * r.ns);
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("out_pt", 2 * sizeof(double) * r.nt);
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("out_ft", 6 * sizeof(double) * r.nt);
    int ier = r.kernel == 1 ? helm_solve_mode(ctx, c, r.prec) : lap_solve_mode(ctx, c, r.prec);
    if (ier) return ier;
    auto down = [&](std::vector<double>& h, const double* d, size_t n) {
      if (!d) return;
      h.resize(n);
      FMM_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    };
    down(r.pot, c.pot_dev, 2 * (size_t)r.ns);
    down(r.fld, c.fld_dev, 6 * (size_t)r.ns);
    down(r.pt, c.pottarg_dev, 2 * (size_t)r.nt);
    down(r.ft, c.fldtarg_dev, 6 * (size_t)r.nt);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  }

  // Several small Laplace requests (same_batch) as one forest: host trees,
  // merged; inputs concatenated and uploaded once; one graph-captured solve;
  // outputs split back.  A tree that cannot be built fails the batch (4).
  static int solve_fused(Context& ctx, std::vector<std::shared_ptr<Request>>& batch) {
    const Request& r0 = *batch[0];
    double eps;
    int nbox;
    if (!iprec_params(r0.iprec, r0.ns, r0.nt, eps, nbox)) return 32;
    const int nb = (int)batch.size();
    thread_local std::vector<Octree> trees;
    thread_local Octree forest;
    if ((int)trees.size() < nb) trees.resize(nb);
    std::vector<const Octree*> tp(nb);
    int ns = 0, nt = 0;
    for (int k = 0; k < nb; k++) {
      const Request& q = *batch[k];
      trees[k].reset();
      if (build_octree(q.src.data(), q.ns, q.nt > 0 ? q.trg.data() : nullptr, q.nt, nbox, trees[k]) != 0) return 4;
      tp[k] = &trees[k];
      ns += q.ns;
      nt += q.nt;
    }
    merge_forest(tp, forest);
    // concatenated inputs through one pinned staging buffer
    auto up = [&](const char* name, size_t per, bool use, const std::vector<double> Request::*field,
                  int Request::*count) -> double* {
      if (!use) return nullptr;
      size_t tot = 0;
      for (auto& q : batch) tot += per * (size_t)((*q).*count);
      if (tot == 0) return nullptr;
      double* h = (double*)ctx.arena.get_host(std::string(name) + "_h", tot * sizeof(double));
      size_t o = 0;
      for (auto& q : batch) {
        const std::vector<double>& v = (*q).*field;
        std::memcpy(h + o, v.data(), v.size() * sizeof(double));
        o += per * (size_t)((*q).*count);
      }
      double* d = (double*)ctx.arena.get(name, tot * sizeof(double));
      FMM_HIP(hipMemcpyAsync(d, h, tot * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      return d;
    };
    LapCall c;
    c.iprec = r0.iprec; c.ns = ns; c.nt = nt;
    c.ifcharge = r0.ifcharge; c.ifdipole = r0.ifdipole; c.ifpot = r0.ifpot; c.iffld = r0.iffld;
    c.ifpottarg = nt > 0 && r0.ifpottarg; c.iffldtarg = nt > 0 && r0.iffldtarg;
    c.src_dev = up("fin_src", 3, true, &Request::src, &Request::ns);
    c.q_dev = up("fin_q", 2, c.ifcharge, &Request::q, &Request::ns);
    c.ds_dev = up("fin_ds", 2, c.ifdipole, &Request::ds, &Request::ns);
    c.dv_dev = up("fin_dv", 3, c.ifdipole, &Request::dv, &Request::ns);
    c.trg_dev = up("fin_trg", 3, nt > 0, &Request::trg, &Request::nt);
    if (c.ifpot) c.pot_dev = (double*)ctx.arena.get("fout_pot", 2 * sizeof(double) * ns);
    if (c.iffld) c.fld_dev = (double*)ctx.arena.get("fout_fld", 6 * sizeof(double) * ns);
    if (c.ifpottarg) c.pottarg_dev = (double*)ctx.arena.get("fout_pt", 2 * sizeof(double) * nt);
    if (c.iffldtarg) c.fldtarg_dev = (double*)ctx.arena.get("fout_ft", 6 * sizeof(double) * nt);
    c.tree_in = &forest;
    c.graph = true;
    if (int ier = lap_solve_mode(ctx, c, r0.prec)) return ier;
    auto down = [&](const double* d, size_t per, std::vector<double> Request::*field, int Request::*count) {
      if (!d) return;
      size_t o = 0;
      for (auto& q : batch) {
        const size_t n = per * (size_t)((*q).*count);
        std::vector<double>& v = (*q).*field;
        v.resize(n);
        if (n) FMM_HIP(hipMemcpyAsync(v.data(), d + o, n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
        o += n;
      }
    };
    down(c.pot_dev, 2, &Request::pot, &Request::ns);
    down(c.fld_dev, 6, &Request::fld, &Request::ns);
    down(c.pottarg_dev, 2, &Request::pt, &Request::nt);
    down(c.fldtarg_dev, 6, &Request::ft, &Request::nt);
    FMM_HIP(hipStreamSynchronize(ctx.stream));
    return 0;
  }
};

}  // namespace fmm3d

using fmm3d::Server;

extern "C" {

void* fmm3d_server_create(int nworkers, long long mem_budget_bytes) {
  if (nworkers < 1 || nworkers > 16 || mem_budget_bytes <= 0) return nullptr;
  auto* S = new Server();
  S->budget = (size_t)mem_budget_bytes;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    delete S;
    return nullptr;
  }
  const int host_threads = std::max(1, omp_get_max_threads() / nworkers);
  for (int i = 0; i < nworkers; i++) {
    auto w = std::make_unique<fmm3d::Worker>();
    w->ctx.device = dev;
    w->ctx.near_on_aux = false;   // the workers overlap whole requests instead
    if (hipStreamCreateWithFlags(&w->ctx.stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      w->ctx.stream = nullptr;
    }
    S->workers.push_back(std::move(w));
  }
  for (auto& w : S->workers) {
    fmm3d::Worker* wp = w.get();
    wp->th = std::thread([S, wp, host_threads] {
      omp_set_num_threads(host_threads);
      try {
        S->run(wp);
      } catch (...) {
      }
    });
  }
  return S;
}

long long fmm3d_server_submit(void* server, int kernel, int prec_mode, int iprec, const double* zk,
                              int nsource, const double* source, int ifcharge, const double* charge,
                              int ifdipole, const double* dipstr, const double* dipvec, int ifpot,
                              int iffld, int ntarget, const double* target, int ifpottarg,
                              int iffldtarg) {
  auto* S = static_cast<Server*>(server);
  if (!S || nsource < 1 || ntarget < 0 || (kernel != 0 && kernel != 1)) return -128;
  if (iprec < -2 || iprec > 6) return -32;
  if (kernel == 1 && (!zk || (zk[0] == 0.0 && zk[1] == 0.0) || (prec_mode != 0 && prec_mode != 1))) return -128;
  if (kernel == 0 && (prec_mode < 0 || prec_mode > 3)) return -128;
  if (!source || (ifcharge == 1 && !charge) || (ifdipole == 1 && (!dipstr || !dipvec)) ||
      (ntarget > 0 && !target))
    return -128;
  auto r = std::make_shared<fmm3d::Request>();
  r->kernel = kernel; r->prec = prec_mode; r->iprec = iprec;
  if (zk) { r->zk[0] = zk[0]; r->zk[1] = zk[1]; }
  r->ns = nsource; r->nt = ntarget;
  r->ifcharge = ifcharge == 1; r->ifdipole = ifdipole == 1;
  r->ifpot = ifpot == 1; r->iffld = iffld == 1; r->ifpottarg = ifpottarg == 1; r->iffldtarg = iffldtarg == 1;
  r->src.assign(source, source + 3 * (size_t)nsource);
  if (r->ifcharge) r->q.assign(charge, charge + 2 * (size_t)nsource);
  if (r->ifdipole) {
    r->ds.assign(dipstr, dipstr + 2 * (size_t)nsource);
    r->dv.assign(dipvec, dipvec + 3 * (size_t)nsource);
  }
  if (ntarget > 0) r->trg.assign(target, target + 3 * (size_t)ntarget);
  r->bytes = fmm3d::estimate_bytes(*r);
  r->t_submit = fmm3d::Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    S->all[r->id] = r;
    S->queue.push_back(r);
  }
  S->cv_work.notify_one();
  return r->id;
}

int fmm3d_server_wait(void* server, long long id, int block, double* pot, double* fld, double* pottarg,
                      double* fldtarg, int* ier, double* latency_ms) {
  auto* S = static_cast<Server*>(server);
  if (!S) return -1;
  std::shared_ptr<fmm3d::Request> r;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return -1;
    r = it->second;
    if (!r->done) {
      if (!block) return 0;
      S->cv_done.wait(lk, [&] { return r->done; });
    }
    // another waiter on the same id may have released it while this one slept
    if (S->all.erase(id) == 0) return -1;
  }
  auto cp = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  if (r->ier == 0) {
    if (r->ifpot) cp(pot, r->pot);
    if (r->iffld) cp(fld, r->fld);
    if (r->ifpottarg) cp(pottarg, r->pt);
    if (r->iffldtarg) cp(fldtarg, r->ft);
  }
  if (ier) *ier = r->ier;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_submit).count();
  return 1;
}

void fmm3d_server_stats(void* server, long long* queued, int* running, long long* bytes_in_use,
                        long long* fused_requests) {
  auto* S = static_cast<Server*>(server);
  if (!S) return;
  std::lock_guard<std::mutex> lk(S->mu);
  if (queued) *queued = (long long)S->queue.size();
  if (running) *running = S->busy;
  if (bytes_in_use) *bytes_in_use = (long long)S->in_use;
  if (fused_requests) *fused_requests = S->fused;
}

void fmm3d_server_destroy(void* server) {
  auto* S = static_cast<Server*>(server);
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& w : S->workers)
    if (w->th.joinable()) w->th.join();
  for (auto& w : S->workers) {
    w->ctx.arena.release_all();
    if (w->ctx.stream) (void)hipStreamDestroy(w->ctx.stream);
  }
  delete S;
}

}  // extern "C"
