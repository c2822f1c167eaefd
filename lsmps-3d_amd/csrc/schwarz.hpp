// Restricted additive Schwarz blocks around the PPE's wrong-sign rows (schwarz.hip).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace lsmps {

constexpr int kMaxBlock = 256;  // rows per dense block (= workgroup size of its kernels)
constexpr int kApplyParts = 4;                       // workgroups per block in the application
constexpr int kApplyCols = kMaxBlock / kApplyParts;  // columns per workgroup

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
  // mixed precision: FP16 copies of inv32 read by the application, column c
  // of block b stored as inv32 * 2^e and col_scale[b * kMaxBlock + c] = 2^-e
  // (the column's largest magnitude lands in [2^14, 2^15)); entry (r, c) at
  // half_index(b, r, c): the four columns of a quad side by side in each row
  bool half = false;
  DeviceArray<uint16_t> inv16;
  DeviceArray<float> col_scale;
  DeviceArray<int32_t> cover;          // per-row counter (setup only)
  DeviceArray<int32_t> cover_start;    // n + 1: row i's slots are cover_entries[start[i], start[i+1])
  DeviceArray<int32_t> cover_entries;  // slots b * kMaxBlock + position, ascending per row
  DeviceArray<int32_t> entry_of;       // blocks x kMaxBlock: slot -> its index in cover_entries
  DeviceArray<int32_t> covered;        // the covered rows, ascending
  size_t covered_count = 0;
  DeviceArray<double> pdiag;           // 1 on covered rows, 1 / A_ii elsewhere
  DeviceArray<double> ybuf;            // kApplyParts x ybuf_stride: partial p of entry e at p * stride + e
  size_t ybuf_stride = 0;              // blocks x kMaxBlock (at least the number of cover entries)
  size_t bytes() const {
    return eligible.bytes() + seed_flag.bytes() + seed_pos.bytes() + seeds.bytes() +
           members.bytes() + sizes.bytes() + scratch.bytes() + inv64.bytes() + inv32.bytes() +
           inv16.bytes() + col_scale.bytes() +
           cover.bytes() + cover_start.bytes() + cover_entries.bytes() + entry_of.bytes() +
           covered.bytes() +
           pdiag.bytes() + ybuf.bytes();
  }
};

// FP16 inverses: each quad of columns is a kMaxBlock x 4 row-major tile, so a
// row's four entries of the quad are one 8-byte load.
__host__ __device__ __forceinline__ size_t half_index(size_t b, int r, int c) {
  return (b * kMaxBlock + static_cast<size_t>(c & ~3)) * kMaxBlock + static_cast<size_t>(r) * 4 +
         static_cast<size_t>(c & 3);
}

// Value of covered row i: the mean of its blocks' products (each the sum of
// its kApplyParts partials), in ascending block order.  The partials are
// stored in cover-entry order (partial p of entry e at p * stride + e), so
// row i reads kApplyParts contiguous pieces [start[i], start[i + 1]) right
// after its range; a gather through the block slots instead cost a chain of
// dependent loads per block (11 us per IDR step at 1M rows).
__device__ __forceinline__ double schwarz_row_mean(const int32_t* start, const double* ybuf,
                                                   size_t stride, int32_t i) {
  const int32_t lo = start[i], hi = start[i + 1];
  double z = 0.0;
  for (int32_t e = lo; e < hi; ++e) {
#pragma unroll
    for (int p = 0; p < kApplyParts; ++p) z += ybuf[static_cast<size_t>(p) * stride + e];
  }
  return z / static_cast<double>(hi - lo);
}

// y_b = inv_b w_b, split over kApplyParts workgroups per block b so that a
// few hundred blocks still fill the GPU: workgroup (b, p) = group multiplies
// the column range [p * 64, p * 64 + 64) and thread t (row t) writes its
// partial sum to ybuf[p * stride + entry_of[b * kMaxBlock + t]].  The column
// slices are read coalesced across the rows.  With FP16 inverses
// (V = _Float16) column j holds inv * 2^e_j and the input slice takes the
// exact factor scale[j] = 2^-e_j instead (scale null: none).  Called by every
// thread of a kMaxBlock-thread workgroup.
template <typename T, typename V>
__device__ __forceinline__ void schwarz_block_product(unsigned group, const int32_t* members,
                                                      const int32_t* sizes, const V* inverse,
                                                      const float* scale, const T* w,
                                                      const int32_t* entry_of, double* ybuf,
                                                      size_t stride) {
  __shared__ double wb[kApplyCols];
  const int t = threadIdx.x;
  const size_t b = group / kApplyParts;
  const int part = static_cast<int>(group % kApplyParts);
  const int n = sizes[b];
  const int j0 = part * kApplyCols;
  const int j1 = min(n, j0 + kApplyCols);
  constexpr int ld = kMaxBlock;
  const int32_t* mem = members + b * kMaxBlock;
  const V* inv = inverse + b * ld * ld;
  if (t < kApplyCols) {
    double x = j0 + t < j1 ? static_cast<double>(w[mem[j0 + t]]) : 0.0;
    if (scale && j0 + t < j1) x *= static_cast<double>(scale[b * ld + j0 + t]);
    wb[t] = x;
  }
  __syncthreads();
  if (t >= n) return;
  double y[4] = {0.0, 0.0, 0.0, 0.0};
  int j = j0;
  if constexpr (std::is_same<V, _Float16>::value) {
    // quad tiles (half_index): one 8-byte load per row and quad; j0 is a multiple of 4
    typedef _Float16 half4 __attribute__((ext_vector_type(4)));
    const V* hb = inverse + b * ld * ld;
    for (; j + 4 <= j1; j += 4) {
      const half4 q = *reinterpret_cast<const half4*>(hb + static_cast<size_t>(j) * ld + 4 * t);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        y[u] += static_cast<double>(static_cast<float>(q[u])) * wb[j + u - j0];
    }
    for (; j < j1; ++j)
      y[0] += static_cast<double>(static_cast<float>(hb[half_index(0, t, j)])) * wb[j - j0];
  } else {
    for (; j + 4 <= j1; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        y[u] += static_cast<double>(static_cast<float>(inv[static_cast<size_t>(j + u) * ld + t])) *
                wb[j + u - j0];
    }
    for (; j < j1; ++j)
      y[0] += static_cast<double>(static_cast<float>(inv[static_cast<size_t>(j) * ld + t])) *
              wb[j - j0];
  }
  ybuf[static_cast<size_t>(part) * stride + entry_of[b * kMaxBlock + t]] =
      (y[0] + y[1]) + (y[2] + y[3]);
}

// Builds the blocks of matrix A from its per-row statistics: diagonal A_ii,
// Jacobi diagonal 1 / A_ii and the sum of the off-diagonal magnitudes.
// fp32_inverse: mixed precision (FP32 inverses; FP16 application copies when
// half_application).
void schwarz_setup(Context& ctx, SchwarzPrec& P, const lsmps_csr& A, const double* diag,
                   const double* dinv, const double* offabs, bool fp32_inverse,
                   bool half_application);
// The blocks of A's preconditioner for A^T: the same member sets, whose dense
// blocks of A^T are the transposes of A's, so their inverses are the
// transposed inverses of src (restricted to the same covered rows; the
// diagonal part 1 / A_ii is shared).
void schwarz_transpose_of(Context& ctx, SchwarzPrec& P, const SchwarzPrec& src,
                          bool half_application);
// In-place completion of a preconditioned vector: on entry v = pdiag * w (w on
// covered rows, Jacobi values elsewhere); on exit each covered row holds the mean
// of its blocks' values (inv_b w_b), summed in block order.
template <typename T>
void schwarz_apply(Context& ctx, const SchwarzPrec& P, T* v, const int32_t* skip = nullptr);
// The completion alone (covered rows of v take their blocks' mean from P.ybuf),
// after the products were formed by schwarz_products or products_and_restrict.
template <typename T>
void schwarz_complete(Context& ctx, const SchwarzPrec& P, T* v, const int32_t* skip = nullptr);
// The block products alone (into P.ybuf); a consumer that reads v row by row
// can then take covered rows from schwarz_row_mean instead of the completion.
template <typename T>
void schwarz_products(Context& ctx, const SchwarzPrec& P, const T* v,
                      const int32_t* skip = nullptr);
void schwarz_stats(Context& ctx, const SchwarzPrec& P, int64_t* blocks, int64_t* members);

}  // namespace lsmps
