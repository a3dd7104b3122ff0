// Device runtime shared by the FMM drivers: a per-device context with its
// stream, a grow-only buffer arena, event timers and HIP error mapping.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmm3d {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct AllocError : std::runtime_error {
  int ier;
  AllocError(const std::string& s, int code) : std::runtime_error(s), ier(code) {}
};

#define FMM_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::fmm3d::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ier of a call whose buffers do not fit a fixed region (see Arena::use_fixed)
constexpr int kIerScratch = 256;

// Grow-only named device buffers; released when the context dies.
//
// Fixed mode (use_fixed): every device buffer is a slice of one region the
// caller owns, handed out in request order, and host staging buffers are
// ordinary heap memory; the arena then never calls hipMalloc, hipFree,
// hipHostMalloc or hipHostFree.  reset_fixed() forgets the slices (start of
// a call); a request beyond the region throws AllocError(kIerScratch).
class Arena {
 public:
  ~Arena();
  void* get(const std::string& name, size_t bytes, int ier_on_fail = 8);
  // pinned host buffer (hipHostMalloc), same growth rule
  void* get_host(const std::string& name, size_t bytes, int ier_on_fail = 8);
  void release_all();
  void use_fixed(void* base, size_t cap);
  void reset_fixed();
  bool fixed() const { return fbase_ != nullptr; }

 private:
  struct Buf { void* p = nullptr; size_t n = 0; };
  std::map<std::string, Buf> bufs_, hbufs_;
  std::map<std::string, std::vector<char>> heap_;  // host buffers in fixed mode
  char* fbase_ = nullptr;
  size_t fcap_ = 0, fused_ = 0;
};

struct Timer {
  std::vector<hipEvent_t> ev;
  hipStream_t s = nullptr;
  void start(hipStream_t stream, int n);
  void mark(int i);
  std::vector<double> elapsed();  // ms between consecutive marks
  ~Timer();
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;  // null = legacy default stream
  Arena arena;
  std::mutex mu;
  // (order p, tolerance eps) of the Laplace translation tables in this
  // context's arena (the host tables are built once per process and key;
  // eps sets the per-offset truncation orders, so it is part of the key)
  std::set<std::pair<int, double>> lap_tabs;
  // cached Helmholtz tables (owned by the Helmholtz driver)
  std::shared_ptr<void> helm_tables;
  // second stream (non-blocking, default priority) for the near field, and
  // the two events that order it against `stream`; made on first use
  hipStream_t aux = nullptr;
  hipEvent_t aux_ev[2] = {nullptr, nullptr};
  // false: the near field stays on `stream` (the request server, whose
  // workers already overlap whole requests)
  bool near_on_aux = true;
  // staged host copies (staged_upload / staged_download): the pinned
  // buffer in turn and, per buffer, the event after its last DMA
  int stage_next = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_pending[2] = {false, false};
  ~Context();
  Context() = default;
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
};

// the context's second stream, created on first use on the current device
hipStream_t aux_stream(Context& ctx);

// Copies between pageable host memory and the device on ctx.stream through
// two pinned staging buffers of kStageBytes: the host-side copy of one chunk
// (OpenMP threads) overlaps the DMA of the other.  staged_upload returns
// once every chunk is queued (later work on ctx.stream is ordered after
// it); staged_download returns with the host array filled.
constexpr size_t kStageBytes = (size_t)16 << 20;
void staged_upload(Context& ctx, void* dev, const void* host, size_t bytes);
void staged_download(Context& ctx, void* host, const void* dev, size_t bytes);

// the stream the near field of this call runs on: the second stream, or
// `ctx.stream` when the context keeps it there or FMM3D_NEAR_STREAM=0
hipStream_t near_stream(Context& ctx);

Context& default_context();

// last per-stage timings of this thread
std::vector<double>& last_timings();

}  // namespace fmm3d
