// Shared host/device infrastructure of the gfx950 LSMPS library: error
// propagation (C++ exceptions inside, status codes at the C-ABI), RAII device
// arrays, the context (device, stream, precision, scratch) and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <array>
#include <atomic>
#include <map>
#include <memory>
#include <vector>
#include <stdexcept>
#include <string>
#include <utility>

#include "lsmps3d_amd.h"

namespace lsmps {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void raise(int code, const std::string& message) { throw Error(code, message); }

inline void require(bool ok, const char* message, int code = LSMPS_ERR_INVALID_ARGUMENT) {
  if (!ok) raise(code, message);
}

#define LSMPS_HIP(call)                                                                   \
  do {                                                                                    \
    hipError_t lsmps_e_ = (call);                                                         \
    if (lsmps_e_ != hipSuccess)                                                           \
      ::lsmps::raise(lsmps_e_ == hipErrorOutOfMemory ? LSMPS_ERR_OUT_OF_MEMORY            \
                                                     : LSMPS_ERR_DEVICE,                  \
                     std::string(#call) + ": " + hipGetErrorString(lsmps_e_));            \
  } while (0)

#define LSMPS_LAUNCH_CHECK() LSMPS_HIP(hipGetLastError())

// Regions of host code that must not allocate or free device memory (work
// enqueued there may still read every buffer it was given, and a stream being
// captured into a graph cannot allocate): while an AllocationFreeze lives on
// this thread, any DeviceArray reallocation raises LSMPS_ERR_INTERNAL instead.
inline int& allocation_freeze_depth() {
  thread_local int depth = 0;
  return depth;
}
struct AllocationFreeze {
  AllocationFreeze() { ++allocation_freeze_depth(); }
  ~AllocationFreeze() { --allocation_freeze_depth(); }
  AllocationFreeze(const AllocationFreeze&) = delete;
  AllocationFreeze& operator=(const AllocationFreeze&) = delete;
};

// Move-only owner of one device allocation.  Reallocates only when the
// requested element count grows (reserve) or changes (resize).
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr_(o.ptr_), n_(o.n_) {
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      std::swap(ptr_, o.ptr_);
      std::swap(n_, o.n_);
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    if (allocation_freeze_depth() > 0)
      raise(LSMPS_ERR_INTERNAL, "device allocation inside an allocation-free region");
    release();
    if (n > 0) LSMPS_HIP(hipMalloc(reinterpret_cast<void**>(&ptr_), n * sizeof(T)));
    n_ = n;
  }
  void reserve(size_t n) {
    if (n > n_) resize(n);
  }
  void release() noexcept {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  T* data() const { return ptr_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
};

constexpr int kBlock = 256;

inline unsigned blocks_for(size_t n, int block = kBlock) {
  const size_t b = (n + static_cast<size_t>(block) - 1) / static_cast<size_t>(block);
  return static_cast<unsigned>(b > 0 ? b : 1);
}

// Workgroups are dispatched round-robin over the GPU's 8 XCDs, each with its
// own L2.  Kernels whose threads gather the data of spatial neighbours
// (particles in lattice or cell order) number their work by this index instead
// of blockIdx.x: XCD x (= blockIdx.x % 8) gets the contiguous range of about
// gridDim.x / 8 blocks starting at x * (g / 8) + min(x, g % 8), so the
// neighbour data a block gathers was mostly brought into that XCD's L2 by the
// blocks before it.  A permutation of [0, gridDim.x) for any grid size.
constexpr unsigned kXcdCount = 8;
__device__ __forceinline__ size_t xcd_slab_block() {
  const unsigned g = gridDim.x, b = blockIdx.x;
  const unsigned x = b % kXcdCount, q = g / kXcdCount, r = g % kXcdCount;
  return static_cast<size_t>(x) * q + (x < r ? x : r) + b / kXcdCount;
}

// Kernels that a captured solver cycle replays take an optional `skip` word:
// when it points at a non-zero value the whole launch returns at once (the
// value is uniform, so no barrier is split).  Eager launches pass null.
__device__ __forceinline__ bool launch_skipped(const int32_t* skip) {
  return skip != nullptr && *skip != 0;
}

// Fill kernels replace hipMemsetAsync so that every clear is an ordinary
// kernel node if the stream is captured into a graph.
void fill_i32(hipStream_t s, int32_t* p, size_t n, int32_t v);
void fill_f64(hipStream_t s, double* p, size_t n, double v);
void fill_f32(hipStream_t s, float* p, size_t n, float v);
void copy_f64(hipStream_t s, double* dst, const double* src, size_t n);

// Device-wide exclusive scan of n int32 values into out[0..n]; out[n] = total.
// `in` and `out` may alias.  Uses ctx scratch.
struct Context;
void exclusive_scan_i32(Context& ctx, const int32_t* in, int32_t* out, size_t n);

// Lanes-per-particle choices of the neighbour-list kernels (tune.hpp).
enum TunedKernel : int {
  kTuneScanCount = 0,  // neighbour count pass (one table)
  kTuneScanWrite,      // neighbour fill pass
  kTuneProvision,
  kTuneAssembly,
  kTuneGradient,
  kTuneShifting,
  kTuneSmoothing,
  kTunedKernels
};

struct TuneTable {
  std::map<std::vector<int32_t>, int> best;  // key -> lanes
  std::array<int, kTunedKernels> forced{};   // > 0: every launch uses this L
  std::array<int, kTunedKernels> last{};     // L of the latest launch
  int64_t tuned_keys = 0;                    // keys timed so far
  hipEvent_t ev[2] = {nullptr, nullptr};
  DeviceArray<int32_t> hist;                 // row-length histogram of the latest key
};

struct Context {
  int device = 0;
  int precision = LSMPS_PRECISION_FP64;
  uint32_t flags = 0;
  hipStream_t stream = nullptr;  // null = legacy default stream of `device`
  // scratch owned by the context
  DeviceArray<int32_t> scan_partials;
  DeviceArray<int32_t> tmp_counts;
  DeviceArray<int32_t> tmp_cursor;
  DeviceArray<int32_t> tmp_labels;   // unsorted cell id per particle
  DeviceArray<int32_t> tmp_primary;  // primary surface labels before the near-surface pass
  DeviceArray<char> tmp_points;      // sorted source points for the neighbor scan
  DeviceArray<double> reduce_scratch;  // block maxima of lsmps_max_velocity
  DeviceArray<double> view_ops;        // operator rows read from caller moment views
  int32_t* host_i32 = nullptr;   // pinned host scratch
  double* host_f64 = nullptr;
  // PPE SpMV timing (LSMPS_FLAG_KERNEL_TIMING)
  // two sets of 2 (s + 1) SpMV events: eager cycles use set 0, the two
  // captured cycle graphs one set each
  static constexpr int kTimingSet = 2 * 17;
  hipEvent_t timing_events[2 * kTimingSet] = {};
  bool timing_ready = false;
  int64_t spmv_launches = 0;
  double spmv_ms = 0.0;
  double spmv_bytes = 0.0;

  bool timing() const { return (flags & LSMPS_FLAG_KERNEL_TIMING) != 0; }
  bool graph() const { return (flags & LSMPS_FLAG_GRAPH) != 0; }
  bool autotune() const { return (flags & LSMPS_FLAG_AUTOTUNE) != 0; }
  bool schwarz_half() const { return (flags & LSMPS_FLAG_SCHWARZ_FP32) == 0; }
  TuneTable tune;

  void activate() const { LSMPS_HIP(hipSetDevice(device)); }
  void finish() const {
    if (!(flags & LSMPS_FLAG_ASYNC)) LSMPS_HIP(hipStreamSynchronize(stream));
  }
  void sync() const { LSMPS_HIP(hipStreamSynchronize(stream)); }
};

inline double config_radius(const lsmps_config& c) { return c.support_radius; }

}  // namespace lsmps

namespace lsmps {
class PpeSolver;
}
struct lsmps_context {
  lsmps::Context impl;
  // Workspace of the free lsmps_ppe_solve entry (created on first use; the
  // complete type is visible where contexts are destroyed).
  std::unique_ptr<lsmps::PpeSolver> solver;
  // Owners: the caller's handle plus every moment matrix and stepper created
  // on this context.  lsmps_context_destroy drops the caller's reference; the
  // context is freed when the last object using it is destroyed, in whatever
  // order the caller destroys them.
  std::atomic<int> refs{1};
};
