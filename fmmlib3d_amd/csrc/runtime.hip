// Device runtime: arena, timers, default context.
#include "runtime.hpp"

#include <cstdlib>
#include <cstring>

namespace fmm3d {

Arena::~Arena() { release_all(); }

Context::~Context() {
  for (auto& e : aux_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : stage_ev)
    if (e) (void)hipEventDestroy(e);
  if (aux) (void)hipStreamDestroy(aux);
}

static void aux_events(Context& ctx) {
  for (auto& e : ctx.aux_ev)
    if (!e) FMM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

hipStream_t aux_stream(Context& ctx) {
  if (!ctx.aux) FMM_HIP(hipStreamCreateWithFlags(&ctx.aux, hipStreamNonBlocking));
  aux_events(ctx);
  return ctx.aux;
}

hipStream_t near_stream(Context& ctx) {
  const char* v = std::getenv("FMM3D_NEAR_STREAM");
  if (!ctx.near_on_aux || (v && std::strcmp(v, "0") == 0)) {
    aux_events(ctx);  // the drivers order the two sums with them either way
    return ctx.stream;
  }
  return aux_stream(ctx);
}

void Arena::use_fixed(void* base, size_t cap) {
  release_all();
  fbase_ = (char*)base;
  fcap_ = cap;
  fused_ = 0;
}

void Arena::reset_fixed() {
  bufs_.clear();
  fused_ = 0;
}

void* Arena::get(const std::string& name, size_t bytes, int ier_on_fail) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  if (fbase_) {
    // a slice of the fixed region: reused while large enough, else a new
    // slice after the last one (the old one is not reclaimed until reset)
    auto it = bufs_.find(name);
    if (it != bufs_.end() && it->second.n >= bytes) return it->second.p;
    if (bytes > fcap_ - fused_)
      throw AllocError("fixed device region too small for " + name, kIerScratch);
    Buf& b = bufs_[name];
    b.p = fbase_ + fused_;
    b.n = bytes;
    fused_ += bytes;
    return b.p;
  }
  Buf& b = bufs_[name];
  if (b.n >= bytes) return b.p;
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
  }
  // grow with headroom so that slightly larger calls reuse the buffer
  size_t want = bytes + bytes / 8;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      b.p = nullptr;
      throw AllocError("device allocation failed for " + name, ier_on_fail);
    }
    want = bytes;
  }
  b.n = want;
  return b.p;
}

void* Arena::get_host(const std::string& name, size_t bytes, int ier_on_fail) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  if (fbase_) {
    std::vector<char>& v = heap_[name];
    if (v.size() < bytes) v.resize(bytes);
    return v.data();
  }
  Buf& b = hbufs_[name];
  if (b.n >= bytes) return b.p;
  if (b.p) {
    (void)hipHostFree(b.p);
    b.p = nullptr;
    b.n = 0;
  }
  const size_t want = bytes + bytes / 8;
  if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    throw AllocError("pinned host allocation failed for " + name, ier_on_fail);
  }
  b.n = want;
  return b.p;
}

void Arena::release_all() {
  if (!fbase_)
    for (auto& kv : bufs_)
      if (kv.second.p) (void)hipFree(kv.second.p);
  bufs_.clear();
  heap_.clear();
  for (auto& kv : hbufs_)
    if (kv.second.p) (void)hipHostFree(kv.second.p);
  hbufs_.clear();
}

void Timer::start(hipStream_t stream, int n) {
  s = stream;
  if ((int)ev.size() < n) {
    for (int i = (int)ev.size(); i < n; i++) {
      hipEvent_t e;
      FMM_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
  }
  FMM_HIP(hipEventRecord(ev[0], s));
}

void Timer::mark(int i) { FMM_HIP(hipEventRecord(ev[i], s)); }

std::vector<double> Timer::elapsed() {
  std::vector<double> out;
  for (size_t i = 1; i < ev.size(); i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev[i - 1], ev[i]) != hipSuccess) {
      (void)hipGetLastError();
      ms = 0;
    }
    out.push_back(ms);
  }
  return out;
}

Timer::~Timer() {
  for (auto e : ev) (void)hipEventDestroy(e);
}

Context& default_context() {
  static Context* ctx = [] {
    auto* c = new Context();
    (void)hipGetDevice(&c->device);
    return c;
  }();
  return *ctx;
}

std::vector<double>& last_timings() {
  static thread_local std::vector<double> t(11, 0.0);
  return t;
}

// ---- staged host copies ----

static void par_copy(void* dst, const void* src, size_t n) {
  constexpr size_t kBlock = (size_t)1 << 20;
  const long nb = (long)((n + kBlock - 1) / kBlock);
  if (nb <= 2) {
    std::memcpy(dst, src, n);
    return;
  }
#pragma omp parallel for schedule(static)
  for (long b = 0; b < nb; b++) {
    const size_t o = (size_t)b * kBlock, len = o + kBlock <= n ? kBlock : n - o;
    std::memcpy((char*)dst + o, (const char*)src + o, len);
  }
}

static char* stage_buffer(Context& ctx, int k) {
  for (auto& e : ctx.stage_ev)
    if (!e) FMM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return (char*)ctx.arena.get_host(k ? "stage_1" : "stage_0", kStageBytes);
}

// the staging buffer k is free once its last DMA has finished
static void stage_wait(Context& ctx, int k) {
  if (ctx.stage_pending[k]) {
    FMM_HIP(hipEventSynchronize(ctx.stage_ev[k]));
    ctx.stage_pending[k] = false;
  }
}

void staged_upload(Context& ctx, void* dev, const void* host, size_t bytes) {
  for (size_t off = 0; off < bytes; off += kStageBytes) {
    const size_t len = off + kStageBytes <= bytes ? kStageBytes : bytes - off;
    const int k = ctx.stage_next;
    ctx.stage_next ^= 1;
    char* pin = stage_buffer(ctx, k);
    stage_wait(ctx, k);
    par_copy(pin, (const char*)host + off, len);
    FMM_HIP(hipMemcpyAsync((char*)dev + off, pin, len, hipMemcpyHostToDevice, ctx.stream));
    FMM_HIP(hipEventRecord(ctx.stage_ev[k], ctx.stream));
    ctx.stage_pending[k] = true;
  }
}

void staged_download(Context& ctx, void* host, const void* dev, size_t bytes) {
  const size_t nchunk = (bytes + kStageBytes - 1) / kStageBytes;
  int kk[2] = {0, 0};
  auto issue = [&](size_t c) {
    const size_t off = c * kStageBytes, len = off + kStageBytes <= bytes ? kStageBytes : bytes - off;
    const int k = ctx.stage_next;
    ctx.stage_next ^= 1;
    char* pin = stage_buffer(ctx, k);
    stage_wait(ctx, k);
    FMM_HIP(hipMemcpyAsync(pin, (const char*)dev + off, len, hipMemcpyDeviceToHost, ctx.stream));
    FMM_HIP(hipEventRecord(ctx.stage_ev[k], ctx.stream));
    ctx.stage_pending[k] = true;
    kk[c & 1] = k;
  };
  if (nchunk == 0) return;
  issue(0);
  for (size_t c = 0; c < nchunk; c++) {
    if (c + 1 < nchunk) issue(c + 1);  // the next chunk's DMA runs while this one is copied out
    const int k = kk[c & 1];
    const size_t off = c * kStageBytes, len = off + kStageBytes <= bytes ? kStageBytes : bytes - off;
    stage_wait(ctx, k);
    par_copy((char*)host + off, stage_buffer(ctx, k), len);
  }
}

}  // namespace fmm3d
