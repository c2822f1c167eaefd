// Restricted additive Schwarz blocks around the PPE's wrong-sign rows (schwarz.hip).
#pragma once

#include "common.hpp"

namespace lsmps {

constexpr int kMaxBlock = 256;  // rows per dense block (= workgroup size of its kernels)

struct SchwarzPrec {
  size_t n = 0;
  size_t blocks = 0;
  bool fp32 = false;
  DeviceArray<int32_t> eligible, seed_flag, seed_pos, seeds;
  DeviceArray<int32_t> members;  // blocks x kMaxBlock, ascending, padded with INT32_MAX
  DeviceArray<int32_t> sizes;
  DeviceArray<double> scratch;   // LU + column solves
  DeviceArray<double> inv64;     // blocks x kMaxBlock^2, column-major
  DeviceArray<float> inv32;
  DeviceArray<int32_t> cover;          // per-row counter (setup only)
  DeviceArray<int32_t> cover_start;    // n + 1: row i's slots are cover_entries[start[i], start[i+1])
  DeviceArray<int32_t> cover_entries;  // slots b * kMaxBlock + position, ascending per row
  DeviceArray<double> ybuf;            // blocks x kMaxBlock: last block products
  size_t bytes() const {
    return eligible.bytes() + seed_flag.bytes() + seed_pos.bytes() + seeds.bytes() +
           members.bytes() + sizes.bytes() + scratch.bytes() + inv64.bytes() + inv32.bytes() +
           cover.bytes() + cover_start.bytes() + cover_entries.bytes() + ybuf.bytes();
  }
};

void schwarz_setup(Context& ctx, SchwarzPrec& P, const lsmps_csr& A, bool fp32_inverse);
// ybuf[b * kMaxBlock + p] = (inv_b w_b)_p for every block b; rows covered by
// blocks then take the mean of their slots (cover_start / cover_entries).
template <typename T>
void schwarz_accumulate(Context& ctx, const SchwarzPrec& P, const T* w);
void schwarz_stats(Context& ctx, const SchwarzPrec& P, int64_t* blocks, int64_t* members);

}  // namespace lsmps
