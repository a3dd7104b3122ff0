// Device runtime shared by the FMM drivers: a per-device context with its
// stream, a grow-only buffer arena, event timers and HIP error mapping.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
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

// Grow-only named device buffers; released when the context dies.
class Arena {
 public:
  ~Arena();
  void* get(const std::string& name, size_t bytes, int ier_on_fail = 8);
  // pinned host buffer (hipHostMalloc), same growth rule
  void* get_host(const std::string& name, size_t bytes, int ier_on_fail = 8);
  void release_all();

 private:
  struct Buf { void* p = nullptr; size_t n = 0; };
  std::map<std::string, Buf> bufs_, hbufs_;
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
  // cached host-built translation tables (by order p)
  int tab_p = -1;
  std::vector<int> tab_host_i;
  std::vector<double> tab_host_d;
  std::vector<long> tab_host_l;
  size_t tab_off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t tab_off_l = 0, tab_off_re = 0;
  size_t tab_off_i[4] = {0, 0, 0, 0};  // grouped M2L items: ritems, rcount, zitems, zoff
  size_t tab_off_g[4] = {0, 0, 0, 0};  // grouped M2L tables: rot_half, z_m2l_t (doubles); their offsets (longs)
  long tab_zt_size = 0;
  std::vector<float> tab_host_f;  // FP32 copies of the grouped-M2L tables
  size_t tab_off_f = 0;
  // cached Helmholtz tables (owned by the Helmholtz driver)
  std::shared_ptr<void> helm_tables;
  // second stream (non-blocking, default priority) for the near field, and
  // the two events that order it against `stream`; made on first use
  hipStream_t aux = nullptr;
  hipEvent_t aux_ev[2] = {nullptr, nullptr};
  // false: the near field stays on `stream` (the request server, whose
  // workers already overlap whole requests)
  bool near_on_aux = true;
  ~Context();
};

// the context's second stream, created on first use on the current device
hipStream_t aux_stream(Context& ctx);

// the stream the near field of this call runs on: the second stream, or
// `ctx.stream` when the context keeps it there or FMM3D_NEAR_STREAM=0
hipStream_t near_stream(Context& ctx);

Context& default_context();

// last per-stage timings of this thread
std::vector<double>& last_timings();

}  // namespace fmm3d
