// Neighbour-count histogram and timing of the autotuner (tune.hpp).
#include <cmath>

#include "tune.hpp"

namespace lsmps {
namespace {

constexpr int kHistBuckets = 8;

// Row-length bucket: [0,16) [16,32) [32,64) [64,96) [96,128) [128,160)
// [160,256) [256, inf).
__device__ __forceinline__ int bucket_of(int32_t len) {
  if (len < 16) return 0;
  if (len < 32) return 1;
  if (len < 64) return 2;
  if (len < 96) return 3;
  if (len < 128) return 4;
  if (len < 160) return 5;
  if (len < 256) return 6;
  return 7;
}

__global__ void __launch_bounds__(kBlock) row_histogram_kernel(size_t n, const int32_t* offsets,
                                                              int32_t* hist) {
  __shared__ int32_t local[kHistBuckets];
  if (threadIdx.x < kHistBuckets) local[threadIdx.x] = 0;
  __syncthreads();
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    atomicAdd(&local[bucket_of(offsets[i + 1] - offsets[i])], 1);
  __syncthreads();
  if (threadIdx.x < kHistBuckets) atomicAdd(&hist[threadIdx.x], local[threadIdx.x]);
}

__global__ void clear_histogram_kernel(int32_t* hist) {
  if (threadIdx.x < kHistBuckets) hist[threadIdx.x] = 0;
}

}  // namespace

void tune_prepare(Context& ctx) {
  TuneTable& t = ctx.tune;
  t.hist.reserve(kHistBuckets);
  for (auto& e : t.ev)
    if (!e) LSMPS_HIP(hipEventCreate(&e));
}

std::vector<int32_t> tune_key(Context& ctx, TunedKernel kernel, size_t n, const int32_t* offsets) {
  TuneTable& t = ctx.tune;
  std::vector<int32_t> key{static_cast<int32_t>(kernel),
                           n > 0 ? static_cast<int32_t>(std::log2(static_cast<double>(n))) : -1};
  if (!offsets || n == 0) return key;
  require(t.hist.size() >= static_cast<size_t>(kHistBuckets), "tuner not prepared",
          LSMPS_ERR_INTERNAL);
  clear_histogram_kernel<<<1, 64, 0, ctx.stream>>>(t.hist.data());
  const unsigned grid = static_cast<unsigned>(std::min<size_t>(blocks_for(n), 1024));
  row_histogram_kernel<<<grid, kBlock, 0, ctx.stream>>>(n, offsets, t.hist.data());
  LSMPS_LAUNCH_CHECK();
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, t.hist.data(), kHistBuckets * sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  for (int b = 0; b < kHistBuckets; ++b)  // fractions in sixteenths
    key.push_back(static_cast<int32_t>(std::lround(16.0 * ctx.host_i32[b] / static_cast<double>(n))));
  return key;
}

void tune_mark(Context& ctx, int which) {
  LSMPS_HIP(hipEventRecord(ctx.tune.ev[which], ctx.stream));
}

float tune_elapsed(Context& ctx) {
  LSMPS_HIP(hipEventSynchronize(ctx.tune.ev[1]));
  float ms = 0.0f;
  LSMPS_HIP(hipEventElapsedTime(&ms, ctx.tune.ev[0], ctx.tune.ev[1]));
  return ms;
}

}  // namespace lsmps
