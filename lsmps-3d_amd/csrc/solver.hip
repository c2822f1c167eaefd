// PPE solve on gfx950: replaces AmgxPpeSolver::solve (src/ppe/ppe_matrix.cu:586-679).
//
// The assembled PPE matrix is non-symmetric and, near wall corners of the free
// surface, indefinite (NearSurface rows with a positive diagonal), so the
// default method is preconditioned IDR(s) with biorthogonalisation
// (Sonneveld & van Gijzen, SIAM J. Sci. Comput. 31, 2008; van Gijzen &
// Sonneveld, ACM TOMS 38, 2011), written from the published algorithm:
// short recurrences, s+1 SpMVs per cycle, s-dimensional shadow space.  IDR(1)
// is mathematically BiCGStab, which LSMPS_SOLVER_BICGSTAB selects.  The
// preconditioner is Jacobi, or Jacobi plus restricted additive Schwarz blocks
// around the wrong-sign rows (schwarz.hip).  Jacobi-PCG (Hestenes-Stiefel) is
// kept for symmetric systems.
//
// GPU structure:
//  * SpMV: sliced ELLPACK with 64-row chunks, one wavefront per chunk and one
//    row per lane, 8 loads in flight per lane, XCD-aware chunk slabs, fused
//    epilogue dots (shadow-space products P^T (A u), or t.r / t.t) into
//    per-block partial records of 16 doubles.
//  * all s x s algebra (M, f, c, alphas, beta, omega) lives in a device struct
//    updated by one-block "scalar" kernels that reduce the partials in a fixed
//    order: no host round trip per iteration and bit-reproducible reductions.
//    The host reads one double per cycle to test convergence.
//  * FP64 mode: IDR(s) on the FP64 matrix, FP64 vectors; at convergence of the
//    recursive residual the true FP64 residual b - A x is recomputed and must
//    meet the tolerance.
//  * mixed mode: the SpMVs of IDR(s) read an FP32 copy of the matrix values
//    (4 instead of 8 bytes per nonzero); the Krylov basis, the iterate, the
//    residual, every dot product and every accumulation stay FP64, and the
//    recursive residual is replaced by the true FP64 residual b - A x each time
//    it falls by `inner_tolerance` (reliable updates), so the solve converges to
//    the FP64 system at FP64 accuracy.
#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

#include <exception>

#include "common.hpp"
#include "solver.hpp"
#include "stages.hpp"

namespace lsmps {

void PpeSolver::release_graphs() noexcept {
  for (auto& e : cycle_exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  cycle_signature.clear();
}

PpeSolver::~PpeSolver() {
  release_graphs();
  for (auto& e : cycle_done)
    if (e) (void)hipEventDestroy(e);
  if (capture_stream) {
    (void)hipStreamSynchronize(capture_stream);
    (void)hipStreamDestroy(capture_stream);
  }
  if (host_control) (void)hipHostFree(host_control);
  if (host) (void)hipHostFree(host);
  if (seg_host) (void)hipHostFree(seg_host);
}

namespace {

constexpr int kLanes = 16;
constexpr uint32_t kShadowSeed = 0x5eedu;

__device__ __forceinline__ double hashed_sign(uint32_t row, int q, uint32_t seed) {
  uint32_t h = row * 0x9E3779B1u + static_cast<uint32_t>(q) * 0x85EBCA77u + seed;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return (h & 1u) ? 1.0 : -1.0;
}

// Shadow vector q: P_0 = initial residual (stored), P_q (q >= 1) random signs.
template <typename T>
__device__ __forceinline__ double shadow(const T* p0, size_t row, int q) {
  return q == 0 ? static_cast<double>(p0[row]) : hashed_sign(static_cast<uint32_t>(row), q, kShadowSeed);
}
// Sums of kMaxShadow (16) values over the 64 lanes of a wavefront by recursive
// halving: at offset 32, 16, 8, 4 each lane keeps half of its remaining sums
// (the half its lane bit selects) and adds the partner's copy of that half,
// then the last two offsets add whole values.  Lane l ends with the sum of
// value l >> 2 (17 double shuffles instead of 16 x 6 for one butterfly per
// value); the order of the additions is fixed, so the sums are reproducible.
__device__ __forceinline__ double wave_sum16(const double (&v)[kMaxShadow], int lane) {
  static_assert(kMaxShadow == 16, "four halving steps");
  double a8[8], a4[4], a2[2];
  const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0,
             b2 = (lane & 4) != 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double keep = b5 ? v[8 + i] : v[i], give = b5 ? v[i] : v[8 + i];
    a8[i] = keep + __shfl_xor(give, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b4 ? a8[4 + i] : a8[i], give = b4 ? a8[i] : a8[4 + i];
    a4[i] = keep + __shfl_xor(give, 16, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b3 ? a4[2 + i] : a4[i], give = b3 ? a4[i] : a4[2 + i];
    a2[i] = keep + __shfl_xor(give, 8, 64);
  }
  double z;
  {
    const double keep = b2 ? a2[1] : a2[0], give = b2 ? a2[0] : a2[1];
    z = keep + __shfl_xor(give, 4, 64);
  }
  z += __shfl_xor(z, 2, 64);
  z += __shfl_xor(z, 1, 64);
  return z;
}

template <int NQ, int BLOCK = kBlock>
__device__ __forceinline__ void block_reduce_store(const double (&acc)[NQ], int nq,
                                                   double* partials) {
  __shared__ double red[BLOCK / 64][kPartialStride];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if constexpr (NQ == kMaxShadow) {  // the 16 shadow dots: recursive halving
    const double v = wave_sum16(acc, lane);
    if ((lane & 3) == 0) red[wave][lane >> 2] = v;
  } else {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (q < nq) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][q] = v;
      }
    }
  }
  __syncthreads();
  // All kPartialStride slots are written (unused ones as 0) so the all-column
  // reduction below only ever reads initialised memory.
  if (static_cast<int>(threadIdx.x) < kPartialStride) {
    double s = 0.0;
    if (static_cast<int>(threadIdx.x) < nq) {
#pragma unroll
      for (int w = 0; w < BLOCK / 64; ++w) s += red[w][threadIdx.x];
    }
    partials[static_cast<size_t>(blockIdx.x) * kPartialStride + threadIdx.x] = s;
  }
}

// Fixed-order sums over b < nb of partials[b][q] for every q < kPartialStride
// (nb * kPartialStride doubles, record b at b * kPartialStride), by the
// kReduceThreads threads of one block.  Thread t owns column t % kPartialStride
// and the records t / kPartialStride + 64 j: its kReduceLoads loads of a round
// (all of nb <= 2048 in one) are issued before the first add and summed in a
// fixed pairwise tree, so every sweep over j is coalesced and one memory
// latency long; the four same-column lanes of a wavefront then meet in a fixed
// butterfly and the 16 wavefronts in LDS, in wavefront order.  Results land in
// out[0 .. kPartialStride).
constexpr int kReduceThreads = 1024;
constexpr int kReduceRows = kReduceThreads / kPartialStride;  // records per sweep
constexpr int kReduceLoads = 32;                               // sweeps in flight
static_assert(kReduceThreads % 64 == 0 && 64 % kPartialStride == 0, "column-stable lanes");
__device__ void reduce_partials(const double* partials, int nb, double* out) {
  __shared__ double red[kReduceThreads / 64][kPartialStride];
  const int t = threadIdx.x;
  const int q = t % kPartialStride;
  const int r0 = t / kPartialStride;
  double s = 0.0;
  for (int base = 0; base < nb; base += kReduceRows * kReduceLoads) {
    double v[kReduceLoads];
#pragma unroll
    for (int j = 0; j < kReduceLoads; ++j) {
      const int b = base + r0 + j * kReduceRows;
      v[j] = b < nb ? partials[static_cast<size_t>(b) * kPartialStride + q] : 0.0;
    }
#pragma unroll
    for (int w = kReduceLoads / 2; w > 0; w >>= 1)
#pragma unroll
      for (int j = 0; j < w; ++j) v[j] += v[j + w];
    s += v[0];
  }
#pragma unroll
  for (int off = kPartialStride; off < 64; off <<= 1) s += __shfl_xor(s, off, 64);
  if ((t & 63) < kPartialStride) red[t >> 6][t & 63] = s;
  __syncthreads();
  if (t < kPartialStride) {
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < kReduceThreads / 64; ++w) tot += red[w][t];
    out[t] = tot;
  }
  __syncthreads();
}

template <typename T, typename V>
struct SpmvArgs {
  size_t n;
  const int32_t* rows;
  const int32_t* cols;
  const V* vals;
  const T* x;
  T* y;
  const T* p0;  // mode 1
  int s;        // mode 1
  const T* r;   // mode 2
  double* partials;
};

// MODE 0: y = A x.  1: + shadow dots P_q . y.  2: + y.r, y.y.  3: + y.x (CG p.Ap).
template <typename T, typename V, int MODE>
__global__ void __launch_bounds__(kBlock) spmv_kernel(SpmvArgs<T, V> a) {
  constexpr int NQ = MODE == 1 ? kMaxShadow : (MODE == 2 ? 2 : 1);
  const int lane = threadIdx.x & (kLanes - 1);
  const size_t group = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kLanes;
  const size_t ngroups = static_cast<size_t>(gridDim.x) * blockDim.x / kLanes;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  for (size_t row = group; row < a.n; row += ngroups) {
    const int32_t b = a.rows[row];
    const int32_t e = a.rows[row + 1];
    T sum = T(0);
    for (int32_t c = b + lane; c < e; c += kLanes) sum += static_cast<T>(a.vals[c]) * a.x[a.cols[c]];
#pragma unroll
    for (int off = kLanes / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, kLanes);
    if (lane == 0) {
      a.y[row] = sum;
      const double ys = static_cast<double>(sum);
      if (MODE == 1) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (q < a.s) acc[q] += shadow(a.p0, row, q) * ys;
      } else if (MODE == 2) {
        acc[0] += ys * static_cast<double>(a.r[row]);
        acc[1 % NQ] += ys * ys;
      } else if (MODE == 3) {
        acc[0] += ys * static_cast<double>(a.x[row]);
      }
    }
  }
  if (MODE != 0) block_reduce_store<NQ>(acc, MODE == 1 ? a.s : (MODE == 2 ? 2 : 1), a.partials);
}

// ---- sliced ELLPACK (SELL-64) ---------------------------------------------

// Bitonic sort of one row's (column, position) keys held by a wavefront, E keys
// per lane (key i = lane * E + e); P = 64 E keys, padded with all-ones keys.
template <int E>
__device__ __forceinline__ void wave_bitonic_sort(uint64_t (&key)[E], int lane) {
  constexpr int P = 64 * E;
#pragma unroll
  for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < E) {  // partner in the same lane
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int f = e ^ j;
          if (f > e) {
            const bool up = ((lane * E + e) & k) == 0;
            const uint64_t a = key[e], b = key[f];
            if ((a > b) == up) {
              key[e] = b;
              key[f] = a;
            }
          }
        }
      } else {  // partner lane = lane ^ (j / E), same slot
        const int m = j / E;
        const bool lower = (lane & m) == 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint64_t other = __shfl_xor(key[e], m, 64);
          const bool up = ((lane * E + e) & k) == 0;
          const bool keep_min = lower == up;
          key[e] = keep_min ? (other < key[e] ? other : key[e]) : (other > key[e] ? other : key[e]);
        }
      }
    }
  }
}

template <int E>
__device__ __forceinline__ void sort_row(int32_t b, int32_t len, const int32_t* cols,
                                         const double* vals, int32_t* out_cols, double* out_vals,
                                         int lane) {
  uint64_t key[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = lane * E + e;
    key[e] = i < len ? (static_cast<uint64_t>(static_cast<uint32_t>(cols[b + i])) << 32) |
                           static_cast<uint32_t>(i)
                     : ~0ull;
  }
  wave_bitonic_sort<E>(key, lane);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = lane * E + e;
    if (i < len) {
      out_cols[b + i] = static_cast<int32_t>(key[e] >> 32);
      out_vals[b + i] = vals[b + static_cast<int32_t>(key[e] & 0xFFFFFFFFu)];
    }
  }
}

// Copy of the CSR with every row's entries in ascending column order (one
// wavefront per row).  In the SELL layout, lane r of a chunk then reads the
// k-th smallest column of its row, and neighbouring particles' k-th columns
// are neighbours too, so the gathers of x coalesce; in CSR (neighbor-list)
// order they scatter over ~50 cache lines per load.  Rows longer than 512
// entries are copied unsorted (still correct).
__global__ void __launch_bounds__(kBlock) sort_rows_kernel(size_t n, const int32_t* rows,
                                                           const int32_t* cols,
                                                           const double* vals,
                                                           int32_t* out_cols, double* out_vals) {
  const int lane = threadIdx.x & 63;
  const size_t wave = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const size_t waves = (static_cast<size_t>(gridDim.x) * blockDim.x) >> 6;
  for (size_t row = wave; row < n; row += waves) {
    const int32_t b = rows[row];
    const int32_t len = rows[row + 1] - b;
    if (len <= 64) {
      sort_row<1>(b, len, cols, vals, out_cols, out_vals, lane);
    } else if (len <= 128) {
      sort_row<2>(b, len, cols, vals, out_cols, out_vals, lane);
    } else if (len <= 256) {
      sort_row<4>(b, len, cols, vals, out_cols, out_vals, lane);
    } else if (len <= 512) {
      sort_row<8>(b, len, cols, vals, out_cols, out_vals, lane);
    } else {
      for (int32_t i = lane; i < len; i += 64) {
        out_cols[b + i] = cols[b + i];
        out_vals[b + i] = vals[b + i];
      }
    }
  }
}

// Index of offset d in a chunk's ascending DIA list, or -1.
__device__ __forceinline__ int dia_find(const int32_t* list, int nd, int32_t d) {
  int lo = 0, hi = nd - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int32_t v = list[mid];
    if (v == d) return mid;
    if (v < d) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

constexpr int kDiaTable = 2048;  // LDS hash slots (power of two)
// Remainder entries of one row that a lane reads with one vector load: 4 values
// (16 bytes in FP32) and 4 16-bit columns (8 bytes).  Slot k of a chunk's
// remainder is at (k / kQuad) * kQuad * kSellChunk + lane * kQuad + k % kQuad.
// A DIA quad holds one diagonal's values of a lane's 4 consecutive rows.
constexpr int kQuad = 4;
__device__ __forceinline__ int64_t quad_slot(int32_t k, int lane) {
  return static_cast<int64_t>(k / kQuad) * (kQuad * kSellChunk) + lane * kQuad + k % kQuad;
}
typedef float quad_f32 __attribute__((ext_vector_type(4)));
typedef double pair_f64 __attribute__((ext_vector_type(2)));
typedef unsigned short quad_u16 __attribute__((ext_vector_type(4)));
typedef int quad_i32 __attribute__((ext_vector_type(4)));
// A lane's quad in one streaming load (two for FP64 values); p is aligned to
// the quad's size by the layout (chunk parts start at multiples of kQuad rows).
__device__ __forceinline__ void load_quad(const float* p, float v[kQuad]) {
  const quad_f32 q = __builtin_nontemporal_load(reinterpret_cast<const quad_f32*>(p));
  v[0] = q.x;
  v[1] = q.y;
  v[2] = q.z;
  v[3] = q.w;
}
__device__ __forceinline__ void load_quad(const double* p, double v[kQuad]) {
  const pair_f64 a = __builtin_nontemporal_load(reinterpret_cast<const pair_f64*>(p));
  const pair_f64 b = __builtin_nontemporal_load(reinterpret_cast<const pair_f64*>(p) + 1);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void load_quad(const uint16_t* p, uint16_t v[kQuad]) {
  const quad_u16 q = __builtin_nontemporal_load(reinterpret_cast<const quad_u16*>(p));
  v[0] = q.x;
  v[1] = q.y;
  v[2] = q.z;
  v[3] = q.w;
}
__device__ __forceinline__ void load_quad(const int32_t* p, int32_t v[kQuad]) {
  const quad_i32 q = __builtin_nontemporal_load(reinterpret_cast<const quad_i32*>(p));
  v[0] = q.x;
  v[1] = q.y;
  v[2] = q.z;
  v[3] = q.w;
}
constexpr int32_t kDiaEmpty = INT32_MIN;
constexpr int kBucketBits = kSellWindowBits - 1;  // 2048-column buckets

// The build kernels walk a chunk's entries (row-sorted, contiguous in the CSR)
// 64 at a time, lane i taking entry base + i, so every load is coalesced.  The
// entry's row within the chunk comes from the chunk's 65 row starts in LDS: the
// last row whose start is <= e (empty rows share a start with the next row).
__device__ __forceinline__ int chunk_row(const int32_t* rs, int32_t e) {
  int lo = 0, hi = kSellChunk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rs[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Loads the chunk's row starts (rows past n start at the CSR's end).
__device__ __forceinline__ void load_row_starts(size_t n, size_t chunk, const int32_t* rows,
                                                int32_t* rs, int lane) {
  const size_t r = chunk * kSellChunk + lane;
  rs[lane] = rows[r < n ? r : n];
  if (lane == 0) {
    const size_t end = chunk * kSellChunk + kSellChunk;
    rs[kSellChunk] = rows[end < n ? end : n];
  }
}

// Rank of a flagged entry among the flagged entries of its row that lie in
// earlier lanes of the same 64-entry group (the row's entries are contiguous
// lanes), and whether this lane is the row's last lane in the group.
struct SegmentRank {
  int before;
  bool tail;
};
__device__ __forceinline__ SegmentRank segment_rank(int lane, int row, bool flag) {
  const int prev = __shfl(row, lane > 0 ? lane - 1 : 0, 64);
  const int next = __shfl(row, lane < 63 ? lane + 1 : 63, 64);
  const bool head = lane == 0 || prev != row;
  const bool tail = lane == 63 || next != row;
  const unsigned long long heads = __ballot(head);
  const unsigned long long flags = __ballot(flag);
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int start = 63 - __clzll(heads & (below | (1ull << lane)));
  const unsigned long long seg = below & ~((1ull << start) - 1ull);
  return {static_cast<int>(__popcll(flags & seg)), tail};
}

// One 256-thread block per group of four chunks: counts the offsets
// column - row of the group's entries in an LDS hash table (wave w hashes the
// entries of chunk 4g + w), keeps those present in at least kDiaMinRows of the
// group's rows as its diagonals (ascending) and plans their storage.  A table
// overflow or more than kMaxDia qualifying offsets leaves the group without
// diagonals, so the selection never depends on the order of the atomics.
//
// Plan (thread 0, nd <= kMaxDia): the sorted offsets are cut into runs of
// consecutive offsets of at most kMaxRun; a run of length m is class m when the
// window of every lane (x[row + d] for the group's first row up to the last
// lane's row + d + 2 * ((m + 4) / 2) - 1) lies inside [0, n), else each of its
// offsets is class 0.  Plan order and storage slots: classes kMaxRun .. 1 (one
// plan entry per run, m slots), then class 0 (one entry and slot per offset);
// within a class, ascending offsets.
__global__ void __launch_bounds__(kGroupRows) dia_group_kernel(
    size_t n, const int32_t* rows, const int32_t* cols, int32_t* dia_count,
    int32_t* dia_offsets, int32_t* dia_slot, int32_t* dia_plan, int32_t* dia_runs,
    int32_t* dia_slots) {
  __shared__ int32_t keys[kDiaTable];
  __shared__ int32_t counts[kDiaTable];
  __shared__ int32_t list[kMaxDia];
  __shared__ int32_t nlist;
  __shared__ int32_t failed;
  __shared__ int32_t rs[kDiaGroup][kSellChunk + 1];
  const size_t g = blockIdx.x;
  const int wave = static_cast<int>(threadIdx.x) >> 6;
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const size_t chunk = g * kDiaGroup + wave;
  const int32_t row0 = static_cast<int32_t>(chunk * kSellChunk);
  load_row_starts(n, chunk, rows, rs[wave], lane);
  for (int t = threadIdx.x; t < kDiaTable; t += kGroupRows) {
    keys[t] = kDiaEmpty;
    counts[t] = 0;
  }
  if (threadIdx.x == 0) {
    nlist = 0;
    failed = 0;
  }
  __syncthreads();
  {
    const int32_t* r = rs[wave];
    const int32_t e0 = r[0], e1 = r[kSellChunk];
    // the next group's column is loaded while this group's entry is hashed
    int32_t col_next = e0 + lane < e1 ? cols[e0 + lane] : 0;
    for (int32_t base = e0; base < e1; base += kSellChunk) {
      const int32_t c = base + lane;
      if (c >= e1) break;
      const int32_t col = col_next;
      if (c + kSellChunk < e1) col_next = cols[c + kSellChunk];
      const int32_t d = col - (row0 + chunk_row(r, c));
      uint32_t h = (static_cast<uint32_t>(d) * 2654435761u) & (kDiaTable - 1);
      bool placed = false;
      for (int probe = 0; probe < kDiaTable; ++probe) {
        const int32_t prev = atomicCAS(&keys[h], kDiaEmpty, d);
        if (prev == kDiaEmpty || prev == d) {
          atomicAdd(&counts[h], 1);
          placed = true;
          break;
        }
        h = (h + 1) & (kDiaTable - 1);
      }
      if (!placed) failed = 1;
    }
  }
  __syncthreads();
  if (!failed) {
    for (int t = threadIdx.x; t < kDiaTable; t += kGroupRows) {
      if (keys[t] != kDiaEmpty && counts[t] >= kDiaMinRows) {
        const int32_t slot = atomicAdd(&nlist, 1);
        if (slot < kMaxDia) list[slot] = keys[t]; else failed = 1;
      }
    }
  }
  __syncthreads();
  const int nd = failed ? 0 : nlist;
  // sort the list ascending (signed order via the sign-flipped high word)
  if (wave == 0) {
    uint64_t key[kMaxDia / kSellChunk];
#pragma unroll
    for (int q = 0; q < kMaxDia / kSellChunk; ++q) {
      const int i = lane * (kMaxDia / kSellChunk) + q;
      key[q] = i < nd ? static_cast<uint64_t>(static_cast<uint32_t>(list[i]) ^ 0x80000000u) << 32
                      : ~0ull;
    }
    wave_bitonic_sort<kMaxDia / kSellChunk>(key, lane);
#pragma unroll
    for (int q = 0; q < kMaxDia / kSellChunk; ++q) {
      const int i = lane * (kMaxDia / kSellChunk) + q;
      if (i < nd) {
        const int32_t d = static_cast<int32_t>(static_cast<uint32_t>(key[q] >> 32) ^ 0x80000000u);
        list[i] = d;
        dia_offsets[g * kMaxDia + i] = d;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // runs (first sorted index, length, class) in ascending order, in the
    // hash table's memory, which is free now
    int32_t* run_first = keys;
    int32_t* run_len = keys + kMaxDia;
    int32_t* run_cls = counts;
    const int64_t grow0 = static_cast<int64_t>(g) * kGroupRows;
    const int64_t last_lane_row = grow0 + kGroupRows - 4;
    int32_t per_class[kRunClasses];
#pragma unroll
    for (int c = 0; c < kRunClasses; ++c) per_class[c] = 0;
    int nr = 0;
    for (int i = 0; i < nd;) {
      int m = 1;
      while (i + m < nd && m < kMaxRun && list[i + m] == list[i] + m) ++m;
      const int64_t lo = grow0 + list[i];
      const int64_t hi = last_lane_row + list[i] + 2 * ((m + 4) / 2) - 1;
      const int cls = lo >= 0 && hi <= static_cast<int64_t>(n) - 1 ? m : 0;
      run_first[nr] = i;
      run_len[nr] = m;
      run_cls[nr] = cls;
      per_class[cls] += cls > 0 ? 1 : m;
      ++nr;
      i += m;
    }
    // plan and slot positions of each class: kMaxRun .. 1, then 0
    int32_t plan_at[kRunClasses], slot_at[kRunClasses];
    int32_t p = 0, sl = 0;
    for (int c = kMaxRun; c >= 0; --c) {
      plan_at[c] = p;
      slot_at[c] = sl;
      p += per_class[c];
      sl += c > 0 ? per_class[c] * c : per_class[c];
    }
    int32_t* plan = dia_plan + g * kMaxDia;
    int32_t* slot_of = dia_slot + g * kMaxDia;
    for (int r = 0; r < nr; ++r) {
      const int i = run_first[r], m = run_len[r], c = run_cls[r];
      if (c > 0) {
        plan[plan_at[c]++] = list[i];
        for (int q = 0; q < m; ++q) slot_of[i + q] = slot_at[c]++;
      } else {
        for (int q = 0; q < m; ++q) {
          plan[plan_at[0]++] = list[i + q];
          slot_of[i + q] = slot_at[0]++;
        }
      }
    }
    for (int c = 0; c < kRunClasses; ++c) dia_runs[g * kRunClasses + c] = per_class[c];
    dia_count[g] = nd;
    dia_slots[g] = nd * kGroupRows;
  }
}

// One 64-thread block per chunk: the entries that are not on one of its
// group's diagonals form the remainder.  Sizes it (the chunk's widest remainder
// row) and lays its column windows over the 2048-wide column buckets it
// touches (a 128-bit mask relative to the chunk's first row).
__global__ void __launch_bounds__(kSellChunk) rem_select_kernel(
    size_t n, const int32_t* rows, const int32_t* cols, const int32_t* dia_count,
    const int32_t* dia_offsets, int32_t* rem_width, int32_t* rem_scaled,
    unsigned long long* rem_total, int32_t* nwin, int32_t* windows, int32_t* col_rows) {
  __shared__ int32_t list[kMaxDia];
  __shared__ int32_t rs[kSellChunk + 1];
  __shared__ int32_t rem_count[kSellChunk + 1];
  const size_t chunk = blockIdx.x;
  const size_t g = chunk / kDiaGroup;
  const int lane = threadIdx.x;
  const int32_t row0 = static_cast<int32_t>(chunk * kSellChunk);
  const int nd = dia_count[g];
  for (int t = lane; t < nd; t += kSellChunk) list[t] = dia_offsets[g * kMaxDia + t];
  load_row_starts(n, chunk, rows, rs, lane);
  rem_count[lane] = 0;
  __syncthreads();
  const int32_t e0 = rs[0], e1 = rs[kSellChunk];
  const int64_t first_bucket = static_cast<int64_t>(row0 >> kBucketBits) - 64;
  uint64_t lo = 0, hi = 0;
  int outside = 0;
  int32_t col_next = e0 + lane < e1 ? cols[e0 + lane] : 0;
  for (int32_t base = e0; base < e1; base += kSellChunk) {
    const int32_t c = base + lane;
    const bool in = c < e1;
    const int lr = in ? chunk_row(rs, c) : kSellChunk;
    const int32_t col = in ? col_next : 0;
    if (c + kSellChunk < e1) col_next = cols[c + kSellChunk];
    const bool rem = in && dia_find(list, nd, col - (row0 + lr)) < 0;
    if (rem) {
      const int64_t rel = (static_cast<int64_t>(col) >> kBucketBits) - first_bucket;
      if (rel < 0 || rel >= 128) {
        outside = 1;
      } else if (rel < 64) {
        lo |= 1ull << rel;
      } else {
        hi |= 1ull << (rel - 64);
      }
    }
    const SegmentRank sr = segment_rank(lane, lr, rem);
    if (in && sr.tail) rem_count[lr] += sr.before + (rem ? 1 : 0);
    __syncthreads();
  }
  const int32_t rem_row = rem_count[lane];
  int32_t width = rem_row;
  unsigned long long total = static_cast<unsigned long long>(rem_row);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    width = max(width, __shfl_xor(width, off, 64));
    total += __shfl_xor(total, off, 64);
    lo |= __shfl_xor(lo, off, 64);
    hi |= __shfl_xor(hi, off, 64);
    outside |= __shfl_xor(outside, off, 64);
  }
  // Windows of two buckets (4096 columns), laid greedily from the lowest
  // uncovered bucket; columns outside the 128-bucket span, or more than
  // kSellWindows windows, give the chunk 32-bit columns.
  int32_t* wb = windows + chunk * kSellWindows;
  int count = 0;
  bool wide = outside != 0;
  for (int bit = 0; bit < 128 && !wide; ++bit) {
    const bool set = bit < 64 ? ((lo >> bit) & 1ull) : ((hi >> (bit - 64)) & 1ull);
    if (!set) continue;
    if (count == kSellWindows) {
      wide = true;
      break;
    }
    if (lane == 0) wb[count] = static_cast<int32_t>((first_bucket + bit) << kBucketBits);
    ++count;
    ++bit;  // the window also covers the next bucket
  }
  if (!wide && count == 0) {  // a chunk without remainder: one window at 0 for the padding
    if (lane == 0) wb[0] = 0;
    count = 1;
  }
  if (wide) {  // 32-bit columns; padding points at the chunk's first row
    if (lane == 0) wb[0] = row0;
    count = 1;
  }
  if (lane == 0) {
    for (int j = count; j < kSellWindows; ++j) wb[j] = wb[count - 1];
    nwin[chunk] = wide ? -1 : count;
    const int32_t width4 = (width + kQuad - 1) / kQuad * kQuad;
    col_rows[chunk] = width4 * (wide ? 2 : 1);
    rem_width[chunk] = width4;
    rem_scaled[chunk] = width4 * kSellChunk;
    atomicAdd(rem_total, total);
  }
}

// One 64-thread block per chunk, one lane per row: each lane walks its row's
// entries in CSR order and writes a value on one of the
// group's diagonals into its slot (the DIA array is zeroed beforehand) and
// every other entry into its row's next remainder slot, with the column
// encoded against the chunk's windows (or as a 32-bit index); then pads its
// row's remainder to the chunk's width (base_0, value 0).  The rows of a
// lattice-ordered chunk mostly sit on the same diagonal at the same step, so
// the lanes' DIA stores land in consecutive words of one slot.
constexpr int kFillStep = 16;
template <typename V>
__global__ void __launch_bounds__(kSellChunk) sell_fill_kernel(
    size_t n, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const int32_t* __restrict__ dia_count,
    const int32_t* __restrict__ dia_offsets, const int32_t* __restrict__ dia_slot,
    const int32_t* __restrict__ dia_start, V* __restrict__ dia_vals,
    const int32_t* __restrict__ chunk_start, const int32_t* __restrict__ width,
    const int32_t* __restrict__ col_start, const int32_t* __restrict__ nwin,
    const int32_t* __restrict__ windows, uint16_t* __restrict__ out_cols,
    V* __restrict__ out_vals) {
  __shared__ int32_t s_off[kMaxDia];
  __shared__ int32_t s_slot[kMaxDia];
  __shared__ int32_t s_win[kSellWindows];
  const size_t chunk = blockIdx.x;
  const size_t g = chunk / kDiaGroup;
  const int lane = threadIdx.x;
  const int nd = dia_count[g];
  for (int t = lane; t < nd; t += kSellChunk) {
    s_off[t] = dia_offsets[g * kMaxDia + t];
    s_slot[t] = dia_slot[g * kMaxDia + t];
  }
  if (lane < kSellWindows) s_win[lane] = windows[chunk * kSellWindows + lane];
  __syncthreads();
  // row r of the chunk is row (chunk % 4) * 64 + r of its group
  V* dia = dia_vals + static_cast<int64_t>(dia_start[g]) + (chunk % kDiaGroup) * kSellChunk + lane;
  const int64_t vbase = chunk_start[chunk];
  const int64_t cbase = static_cast<int64_t>(col_start[chunk]) * kSellChunk;
  const int32_t w = width[chunk];
  const int32_t nw = nwin[chunk];
  auto put = [&](int32_t k, int32_t c, V v) {
    const int64_t slot = quad_slot(k, lane);
    out_vals[vbase + slot] = v;
    if (nw < 0) {
      reinterpret_cast<int32_t*>(out_cols + cbase)[slot] = c;
    } else {
      int j = 0;
      while (j + 1 < nw && s_win[j + 1] <= c) ++j;
      out_cols[cbase + slot] = static_cast<uint16_t>((j << kSellWindowBits) | (c - s_win[j]));
    }
  };
  const size_t row = chunk * kSellChunk + lane;
  const int32_t b = row < n ? rows[row] : 0, e = row < n ? rows[row + 1] : 0;
  const int32_t r32 = static_cast<int32_t>(row);
  int32_t rem = 0;
  auto take = [&](int32_t c, double v) {
    const int j = dia_find(s_off, nd, c - r32);
    if (j >= 0) {
      dia[static_cast<int64_t>(s_slot[j]) * kGroupRows] = static_cast<V>(v);
    } else {
      put(rem++, c, static_cast<V>(v));
    }
  };
  // 16 entries (64 bytes of columns, 128 of values) per step, read together,
  // so that each cache line of the lane's row is fetched about once
  int32_t q = b;
  for (; q + kFillStep <= e; q += kFillStep) {
    int32_t c[kFillStep];
    double v[kFillStep];
#pragma unroll
    for (int u = 0; u < kFillStep; ++u) {
      c[u] = cols[q + u];
      v[u] = vals[q + u];
    }
#pragma unroll
    for (int u = 0; u < kFillStep; ++u) take(c[u], v[u]);
  }
  for (; q < e; ++q) take(cols[q], vals[q]);
  for (int32_t k = rem; k < w; ++k) put(k, s_win[0], V(0));
}

template <typename T, typename V>
struct SellArgs {
  size_t n;
  size_t chunks;
  size_t groups;
  const int32_t* dia_runs;
  const int32_t* dia_plan;
  const int32_t* dia_start;
  const V* dia_vals;
  const int32_t* chunk_start;
  const int32_t* width;
  const int32_t* col_start;
  const int32_t* nwin;
  const int32_t* windows;
  const uint16_t* cols;
  const V* vals;
  const T* x;
  T* y;
  const T* p0;       // mode 1
  int s;             // mode 1
  const double* r;   // mode 2: the FP64 residual
  double* partials;
  const int32_t* skip = nullptr;  // captured cycles: the launch returns when *skip != 0
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Give XCD k a contiguous slab of the rows (and so of the gathered x entries,
// which are spatial neighbors) by renumbering block b as
// (b % 8) * (gridDim / 8) + b / 8; the grid is a multiple of 8.
constexpr int kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// Gathers of x through a raw buffer resource: 32-bit byte offsets (no 64-bit
// address arithmetic per entry) and hardware range checking, which returns 0
// for any offset outside [0, n) -- negative ones wrap to huge unsigned values.
// The padding slots of the diagonals (value 0) may therefore point anywhere.
// Word 3 is the gfx9-family raw-buffer format (32-bit data, no swizzle).
constexpr int kBufferWord3 = 0x00020000;
template <typename T>
struct XGather;
template <>
struct XGather<double> {
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ XGather(const double* p, size_t n)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0,
                                               static_cast<int>(n * sizeof(double)), kBufferWord3)) {}
  __device__ double operator()(uint32_t byte_offset) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, byte_offset, 0, 0));
  }
  // x[i] and x[i + 1] in one 16-byte load (8-byte aligned; the plan only pairs
  // offsets whose two columns lie inside [0, n) for every row of the chunk).
  __device__ void pair(uint32_t byte_offset, double& lo, double& hi) const {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 q = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte_offset, 0, 0));
    lo = __builtin_bit_cast(double, static_cast<unsigned long long>(q.x) |
                                        (static_cast<unsigned long long>(q.y) << 32));
    hi = __builtin_bit_cast(double, static_cast<unsigned long long>(q.z) |
                                        (static_cast<unsigned long long>(q.w) << 32));
  }
};
template <>
struct XGather<float> {
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ XGather(const float* p, size_t n)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), 0,
                                               static_cast<int>(n * sizeof(float)), kBufferWord3)) {}
  __device__ float operator()(uint32_t byte_offset) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, byte_offset, 0, 0));
  }
  __device__ void pair(uint32_t byte_offset, float& lo, float& hi) const {
    const unsigned long long q =
        __builtin_bit_cast(unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(rsrc, byte_offset, 0, 0));
    lo = __builtin_bit_cast(float, static_cast<unsigned int>(q));
    hi = __builtin_bit_cast(float, static_cast<unsigned int>(q >> 32));
  }
};
// Largest system the 32-bit buffer offsets address.
constexpr size_t kMaxGatherRows = (size_t{1} << 31) / sizeof(double) - 1;

// A group's plan held across the lanes of a wavefront (lane j holds entries
// j, j + 64, j + 128, j + 192), read back one entry at a time with readlane:
// the offsets then live in scalar registers and a run's window loads issue
// without first waiting for a vector load of its offset.
struct GroupPlan {
  int32_t v0, v1, v2, v3;
  __device__ int32_t operator[](int p) const {  // p is wave-uniform
    const int lane = p & 63;
    if (p < 128) return __builtin_amdgcn_readlane(p < 64 ? v0 : v1, lane);
    return __builtin_amdgcn_readlane(p < 192 ? v2 : v3, lane);
  }
};
static_assert(kMaxDia == 256, "four plan registers per lane");

// Lane i receives `src` of lane i + 1 (DPP wave_shl:1 on both 32-bit halves);
// lane 63, whose source would be lane 64, keeps `old`.
__device__ __forceinline__ double from_next_lane(double old, double src) {
  const uint64_t o = __builtin_bit_cast(uint64_t, old), v = __builtin_bit_cast(uint64_t, src);
  constexpr int kWaveShl1 = 0x130;
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(o)), static_cast<int>(static_cast<uint32_t>(v)),
      kWaveShl1, 0xF, 0xF, false));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(o >> 32)), static_cast<int>(static_cast<uint32_t>(v >> 32)),
      kWaveShl1, 0xF, 0xF, false));
  return __builtin_bit_cast(double, static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32));
}

// Runs of class M (M consecutive offsets each) of one group, plan entries
// [pbase, pbase + count).  The lane owning rows i .. i + 3 needs the window
// x[i + d .. i + d + M + 2] and the run's M value quads (4 rows each), then adds
// its 4 M products.  Wave w of the block takes every fourth run, starting at
// `first`.  FP64 values: each lane reads its window as (M + 4) / 2 16-byte
// loads.  FP32 values: each lane loads only its own quad x[i + d .. i + d + 3]
// and takes the next quads of the window from the lanes above it (DPP), so a
// run costs two full-wave x loads instead of up to six; lane 63 loads the rest
// of its window itself, in loads of that one lane (111 against 115 us per
// launch at 1M rows).
template <int M, typename T, typename V>
__device__ __forceinline__ void dia_run_class(int count, int first, const GroupPlan& plan,
                                              int pbase, const V* vp, uint32_t rowb,
                                              const XGather<T>& xg, double (&acc)[4]) {
  constexpr bool kShift = sizeof(V) == 4;
  constexpr int NL = (M + 4) / 2;                  // 16-byte loads of a whole window
  constexpr int NQ = (M + 6) / 4;                  // quads holding the M + 3 window elements
  constexpr int NW = kShift ? 4 * NQ : 2 * NL;     // window registers
  const bool top_lane = (threadIdx.x & 63) == 63;
  for (int i = first; i < count; i += kDiaGroup) {
    const uint32_t off = rowb + static_cast<uint32_t>(plan[pbase + i]) * static_cast<uint32_t>(sizeof(T));
    T xw[NW];
    if constexpr (kShift) {
      xg.pair(off, xw[0], xw[1]);
      xg.pair(off + 16u, xw[2], xw[3]);
#pragma unroll
      for (int e = 4; e < NW; ++e) xw[e] = 0.0;
      if (top_lane) {
        // the last quad may reach past the window: those columns are unused (0 past n)
#pragma unroll
        for (int j = 2; j < NW / 2; ++j) xg.pair(off + 16u * j, xw[2 * j], xw[2 * j + 1]);
      }
#pragma unroll
      for (int e = 4; e < NW; ++e) xw[e] = from_next_lane(xw[e], xw[e - 4]);
    } else {
#pragma unroll
      for (int j = 0; j < NL; ++j) xg.pair(off + 16u * j, xw[2 * j], xw[2 * j + 1]);
    }
    V v[M][kQuad];
#pragma unroll
    for (int q = 0; q < M; ++q) load_quad(vp + static_cast<int64_t>(i * M + q) * kGroupRows, v[q]);
#pragma unroll
    for (int q = 0; q < M; ++q)
#pragma unroll
      for (int t = 0; t < kQuad; ++t)
        acc[t] += static_cast<double>(v[q][t]) * static_cast<double>(xw[t + q]);
  }
}

// One 256-thread workgroup per 256-row group at a time; each workgroup owns a
// contiguous range of groups (an even split in the XCD-aware order above).
// DIA part: lane l of every wave holds rows 4 l .. 4 l + 3 of the group and the
// four waves share out the group's runs (run j of the plan goes to wave j mod 4);
// the four partial sums of a row meet in LDS.  Remainder: wave w takes chunk
// 4 g + w with one row per lane, which is also the row whose sum, y and dots
// that lane finishes.  MODE 1: shadow dots P_q . y.  MODE 2: y.r and y.y.
//
// The group index is wave-uniform, so the run counts, the plan offsets and the
// chunk descriptors come through scalar loads and the loops branch on scalars.
constexpr int kSellBlock = kGroupRows;
template <typename T, typename V, int MODE>
__global__ void __launch_bounds__(kSellBlock) sell_spmv_kernel(SellArgs<T, V> a) {
  static_assert(sizeof(T) == sizeof(double), "x windows are read as 16-byte pairs of doubles");
  if (launch_skipped(a.skip)) return;
  __shared__ double part[kDiaGroup][kGroupRows];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const size_t b = xcd_block();
  const size_t first = b * a.groups / gridDim.x;
  const size_t last = (b + 1) * a.groups / gridDim.x;
  const int nq = MODE == 1 ? a.s : (MODE == 2 ? 2 : 0);
  const XGather<T> xg(a.x, a.n);
  // Lane q keeps the wavefront's running dot product q (one register instead of
  // one per dot in every lane, which keeps occupancy up).
  double lane_acc = 0.0;
  for (size_t g = first; g < last; ++g) {
    // this wave's remainder chunk and its descriptors, loaded ahead of the DIA part
    const size_t chunk = g * kDiaGroup + wave;
    const size_t row = chunk * kSellChunk + lane;
    const bool has_chunk = chunk < a.chunks;
    const int32_t w = has_chunk ? __builtin_amdgcn_readfirstlane(a.width[chunk]) : 0;
    const int32_t nw = has_chunk ? __builtin_amdgcn_readfirstlane(a.nwin[chunk]) : 0;
    const int32_t rstart = has_chunk ? __builtin_amdgcn_readfirstlane(a.chunk_start[chunk]) : 0;
    const int32_t cstart = has_chunk ? __builtin_amdgcn_readfirstlane(a.col_start[chunk]) : 0;
    const int32_t wbase =
        has_chunk && nw >= 0 ? a.windows[chunk * kSellWindows + (lane & (kSellWindows - 1))] : 0;
    {
      // the group's plan, run counts and DIA start in lanes, read back as scalars
      const int32_t* pg = a.dia_plan + g * kMaxDia + lane;
      const GroupPlan plan{pg[0], pg[64], pg[128], pg[192]};
      const int32_t cnt_lane = lane < kRunClasses ? a.dia_runs[g * kRunClasses + lane] : 0;
      const int32_t dstart = __builtin_amdgcn_readfirstlane(a.dia_start[g]);
      const uint32_t rowb = static_cast<uint32_t>(g * kGroupRows + kQuad * lane) *
                            static_cast<uint32_t>(sizeof(T));
      const V* vp = a.dia_vals + dstart + kQuad * lane;
      double acc[kQuad] = {0.0, 0.0, 0.0, 0.0};
      int pbase = 0;   // plan entries of the classes already done
      int before = 0;  // runs of the classes already done: sets each class's first run per wave
      auto first_of = [&](int runs_before) {
        return (wave - runs_before % kDiaGroup + kDiaGroup) % kDiaGroup;
      };
#define LSMPS_DIA_CLASS(M)                                                        \
  {                                                                               \
    const int c = __builtin_amdgcn_readlane(cnt_lane, M);                         \
    if (c > 0) {                                                                  \
      dia_run_class<M>(c, first_of(before), plan, pbase, vp, rowb, xg, acc);      \
      pbase += c;                                                                 \
      vp += static_cast<int64_t>(c) * (M * kGroupRows);                           \
      before += c;                                                                \
    }                                                                             \
  }
      static_assert(kMaxRun == 8, "one case per run class");
      LSMPS_DIA_CLASS(8)
      LSMPS_DIA_CLASS(7)
      LSMPS_DIA_CLASS(6)
      LSMPS_DIA_CLASS(5)
      LSMPS_DIA_CLASS(4)
      LSMPS_DIA_CLASS(3)
      LSMPS_DIA_CLASS(2)
      LSMPS_DIA_CLASS(1)
#undef LSMPS_DIA_CLASS
      // class 0: offsets whose windows would leave [0, n) for some lane, one
      // value at a time.  Each column is formed and range-tested on its own:
      // a byte offset that wrapped below zero plus a constant increment may be
      // split by the compiler into the instruction's immediate, and the buffer
      // range check then sees the wrapped base and returns 0 for a column
      // inside [0, n).
      const int64_t lane_row = static_cast<int64_t>(g) * kGroupRows + kQuad * lane;
      const int c0 = __builtin_amdgcn_readlane(cnt_lane, 0);
      for (int i = first_of(before); i < c0; i += kDiaGroup) {
        const int64_t base = lane_row + plan[pbase + i];
        V v[kQuad];
        load_quad(vp + static_cast<int64_t>(i) * kGroupRows, v);
#pragma unroll
        for (int t = 0; t < kQuad; ++t) {
          const int64_t col = base + t;
          const double xe =
              col >= 0 && col < static_cast<int64_t>(a.n)
                  ? static_cast<double>(xg(static_cast<uint32_t>(col) * static_cast<uint32_t>(sizeof(T))))
                  : 0.0;
          acc[t] += static_cast<double>(v[t]) * xe;
        }
      }
#pragma unroll
      for (int t = 0; t < kQuad; ++t) part[wave][kQuad * lane + t] = acc[t];
    }
    double s[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] = 0.0;
    if (has_chunk) {
      constexpr int U = 8;  // entries (two quads) per step
      const V* vp = a.vals + rstart + lane * kQuad;
      const int64_t cbase = static_cast<int64_t>(cstart) * kSellChunk;
      int32_t k = 0;
      if (nw >= 0) {
        // lane l holds window base l % 16; an entry's base is fetched across lanes
        const uint16_t* cp = a.cols + cbase + lane * kQuad;
        constexpr int kMask = (1 << kSellWindowBits) - 1;
        for (; k + U <= w; k += U) {
          uint16_t c[U];
          V v[U];
          load_quad(cp + k * kSellChunk, c);
          load_quad(cp + (k + kQuad) * kSellChunk, c + kQuad);
          load_quad(vp + k * kSellChunk, v);
          load_quad(vp + (k + kQuad) * kSellChunk, v + kQuad);
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int32_t col = __shfl(wbase, c[u] >> kSellWindowBits, 64) + (c[u] & kMask);
            const T xv = xg(static_cast<uint32_t>(col) * static_cast<uint32_t>(sizeof(T)));
            s[u] += static_cast<double>(v[u]) * static_cast<double>(xv);
          }
        }
        for (; k < w; k += kQuad) {
          uint16_t c[kQuad];
          V v[kQuad];
          load_quad(cp + k * kSellChunk, c);
          load_quad(vp + k * kSellChunk, v);
#pragma unroll
          for (int u = 0; u < kQuad; ++u) {
            const int32_t col = __shfl(wbase, c[u] >> kSellWindowBits, 64) + (c[u] & kMask);
            const T xv = xg(static_cast<uint32_t>(col) * static_cast<uint32_t>(sizeof(T)));
            s[u] += static_cast<double>(v[u]) * static_cast<double>(xv);
          }
        }
      } else {
        const int32_t* cp = reinterpret_cast<const int32_t*>(a.cols + cbase) + lane * kQuad;
        for (; k < w; k += kQuad) {
          int32_t c[kQuad];
          V v[kQuad];
          load_quad(cp + k * kSellChunk, c);
          load_quad(vp + k * kSellChunk, v);
#pragma unroll
          for (int u = 0; u < kQuad; ++u) {
            const T xv = xg(static_cast<uint32_t>(c[u]) * static_cast<uint32_t>(sizeof(T)));
            s[u] += static_cast<double>(v[u]) * static_cast<double>(xv);
          }
        }
      }
    }
    __syncthreads();
    const int r = wave * kSellChunk + lane;  // the row within the group
    const double dia = (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
    __syncthreads();  // part is rewritten by the next group
    const double sum = dia + (((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])));
    double ys = 0.0;
    if (row < a.n) {
      // the dots use the stored (rounded) value, which the later updates read
      const T yv = static_cast<T>(sum);
      a.y[row] = yv;
      ys = static_cast<double>(yv);
    }
    if (MODE == 1) {  // the shadow dots: lane l accumulates dot l >> 2
      double d[kMaxShadow];
#pragma unroll
      for (int q = 0; q < kMaxShadow; ++q)
        d[q] = row < a.n && q < nq ? shadow(a.p0, row, q) * ys : 0.0;
      lane_acc += wave_sum16(d, lane);
    } else {
      for (int q = 0; q < nq; ++q) {  // uniform trip count
        double d = 0.0;
        if (row < a.n) d = (q == 0 ? a.r[row] : ys) * ys;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
        if (lane == q) lane_acc += d;
      }
    }
  }
  if (MODE != 0) {
    __shared__ double red[kSellBlock / 64][kPartialStride];
    if (MODE == 1) {  // dot q sits in lanes 4q .. 4q + 3
      if ((lane & 3) == 0) red[wave][lane >> 2] = (lane >> 2) < nq ? lane_acc : 0.0;
    } else if (lane < kPartialStride) {
      red[wave][lane] = lane < nq ? lane_acc : 0.0;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < kPartialStride) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kSellBlock / 64; ++w) t += red[w][threadIdx.x];
      a.partials[static_cast<size_t>(blockIdx.x) * kPartialStride + threadIdx.x] = t;
    }
  }
}

// r = b - A x in FP64; optional FP32 copy of r; partial r.r.  Batched systems
// (row_seg non-null): only the rows of segments k with take[k] != 0.
__global__ void __launch_bounds__(kBlock) residual_kernel(size_t n, const int32_t* rows,
                                                          const int32_t* cols, const double* vals,
                                                          const double* x, const double* b,
                                                          double* r, float* r32,
                                                          double* partials,
                                                          const int32_t* skip = nullptr,
                                                          const int32_t* row_seg = nullptr,
                                                          const int32_t* take = nullptr) {
  if (launch_skipped(skip)) return;
  const int lane = threadIdx.x & (kLanes - 1);
  const size_t group = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kLanes;
  const size_t ngroups = static_cast<size_t>(gridDim.x) * blockDim.x / kLanes;
  double acc[1] = {0.0};
  for (size_t row = group; row < n; row += ngroups) {
    if (row_seg && !take[row_seg[row]]) continue;  // uniform per lane group
    double sum = 0.0;
    for (int32_t c = rows[row] + lane; c < rows[row + 1]; c += kLanes) sum += vals[c] * x[cols[c]];
#pragma unroll
    for (int off = kLanes / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, kLanes);
    if (lane == 0) {
      const double rv = b[row] - sum;
      if (r) r[row] = rv;
      if (r32) r32[row] = static_cast<float>(rv);
      acc[0] += rv * rv;
    }
  }
  block_reduce_store<1>(acc, 1, partials);
}

// Per row: the diagonal (sum of the entries in column i), its inverse for
// Jacobi (1 where it is 0), and the sum of the off-diagonal magnitudes.  Sixteen
// lanes share a row and stride over its entries, so the CSR streams are read
// in coalesced runs; the lane sums meet in a fixed butterfly.
constexpr int kRowLanes = 16;
__global__ void __launch_bounds__(kBlock) row_stats_kernel(size_t n, const int32_t* rows,
                                                           const int32_t* cols,
                                                           const double* vals, double* diag,
                                                           double* dinv, double* offabs) {
  const size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
  const int l = threadIdx.x % kRowLanes;
  double d = 0.0, o = 0.0;
  if (i < n) {
    for (int32_t c = rows[i] + l; c < rows[i + 1]; c += kRowLanes) {
      const double v = vals[c];
      if (cols[c] == static_cast<int32_t>(i)) d += v; else o += fabs(v);
    }
  }
#pragma unroll
  for (int off = kRowLanes / 2; off > 0; off >>= 1) {
    d += __shfl_xor(d, off, kRowLanes);
    o += __shfl_xor(o, off, kRowLanes);
  }
  if (i < n && l == 0) {
    diag[i] = d;
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    offabs[i] = o;
  }
}

template <typename T>
__global__ void fill_kernel(size_t n, T* p, T v, const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename S, typename D>
__global__ void copy_kernel(size_t n, const S* src, D* dst, const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<D>(src[i]);
}

__global__ void dot_self_kernel(size_t n, const double* v, double* partials) {
  double acc[1] = {0.0};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    acc[0] += v[i] * v[i];
  block_reduce_store<1>(acc, 1, partials);
}

__global__ void clamp_kernel(size_t n, double* x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && x[i] < 0.0) x[i] = 0.0;
}

// ---- IDR(s) -------------------------------------------------------------

// IDR(s) state: the residual r and iterate x are FP64; the Krylov basis
// (G, U), the work vectors v, t and the shadow vector P_0 have type T (FP32 in
// mixed mode, so the basis updates move half the bytes).
template <typename T>
struct IdrVec {
  size_t n;
  size_t ld;
  int s;
  double* r;
  double* x;
  T* v;
  T* t;
  T* G;
  T* U;
  const T* p0;
  SolverScalars* sc;
  double* partials;
  // Diagonal part of the preconditioner: 1 / A_ii, except 1 on the rows the
  // Schwarz blocks cover, whose values the block application then replaces.
  const double* pdiag;
  const int32_t* row_seg;  // batched systems: each row's segment (scalars sc[row_seg[i]])
  // Schwarz completion read by idr_a2g_kernel (null without blocks)
  const int32_t* cover_start;
  const double* ybuf;
  size_t ybuf_stride;
  // coarse-space term added by idr_a2g_kernel (null without a coarse space)
  const int32_t* agg_of_row;
  const double* yc;
  const int32_t* skip = nullptr;  // captured cycles: the launch returns when *skip != 0
};

// Scalars that govern row i: the system's, or its segment's in a batched solve.
template <bool kSeg, typename T>
__device__ __forceinline__ const SolverScalars* row_sc(const IdrVec<T>& a, size_t i) {
  if (kSeg) return a.sc + a.row_seg[i];
  return a.sc;
}

// Linear combinations of basis columns for row i, accumulated in q order.  The
// column loads are issued four at a time (the trip counts are run-time values,
// so a plain loop would wait for each load before the next), which keeps four
// (eight for the paired form) HBM reads in flight per lane.
template <typename T>
__device__ __forceinline__ double comb_sub(double acc, const double* c, const T* vec, size_t ld,
                                           size_t i, int q0, int q1) {
  int q = q0;
  for (; q + 4 <= q1; q += 4) {
    const double e0 = static_cast<double>(vec[static_cast<size_t>(q) * ld + i]);
    const double e1 = static_cast<double>(vec[static_cast<size_t>(q + 1) * ld + i]);
    const double e2 = static_cast<double>(vec[static_cast<size_t>(q + 2) * ld + i]);
    const double e3 = static_cast<double>(vec[static_cast<size_t>(q + 3) * ld + i]);
    acc -= c[q] * e0;
    acc -= c[q + 1] * e1;
    acc -= c[q + 2] * e2;
    acc -= c[q + 3] * e3;
  }
  for (; q < q1; ++q) acc -= c[q] * static_cast<double>(vec[static_cast<size_t>(q) * ld + i]);
  return acc;
}
template <typename T>
__device__ __forceinline__ double comb_add(double acc, const double* c, const T* vec, size_t ld,
                                           size_t i, int q0, int q1) {
  int q = q0;
  for (; q + 4 <= q1; q += 4) {
    const double e0 = static_cast<double>(vec[static_cast<size_t>(q) * ld + i]);
    const double e1 = static_cast<double>(vec[static_cast<size_t>(q + 1) * ld + i]);
    const double e2 = static_cast<double>(vec[static_cast<size_t>(q + 2) * ld + i]);
    const double e3 = static_cast<double>(vec[static_cast<size_t>(q + 3) * ld + i]);
    acc += c[q] * e0;
    acc += c[q + 1] * e1;
    acc += c[q + 2] * e2;
    acc += c[q + 3] * e3;
  }
  for (; q < q1; ++q) acc += c[q] * static_cast<double>(vec[static_cast<size_t>(q) * ld + i]);
  return acc;
}
// Both updates of a biorthogonalisation over q in [0, k): g -= c_q G_q and
// u -= c_q U_q, pairwise as before, loads issued two columns of each at a time.
template <typename T>
__device__ __forceinline__ void comb_sub2(double& g, double& u, const double* c, const T* G,
                                          const T* U, size_t ld, size_t i, int k) {
  int q = 0;
  for (; q + 2 <= k; q += 2) {
    const size_t o0 = static_cast<size_t>(q) * ld + i, o1 = o0 + ld;
    const double g0 = static_cast<double>(G[o0]), g1 = static_cast<double>(G[o1]);
    const double u0 = static_cast<double>(U[o0]), u1 = static_cast<double>(U[o1]);
    g -= c[q] * g0;
    u -= c[q] * u0;
    g -= c[q + 1] * g1;
    u -= c[q + 1] * u1;
  }
  for (; q < k; ++q) {
    const size_t o = static_cast<size_t>(q) * ld + i;
    g -= c[q] * static_cast<double>(G[o]);
    u -= c[q] * static_cast<double>(U[o]);
  }
}

// Schwarz path, first half of idr_a: v = pdiag (r - sum_{q>=k} c_q G_q): Jacobi
// on uncovered rows, the raw value (the block application's input) on covered ones.
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_a1_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double v = comb_sub(a.r[i], sc->c, a.G, a.ld, i, k, a.s);
    a.v[i] = static_cast<T>(a.pdiag[i] * v);
  }
}

// Schwarz path, second half (v already preconditioned): U_k = sum_{q>=k} c_q U_q + om v.
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_a2_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  const double om1 = a.sc->om;  // one system: loaded once
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double om = kSeg ? sc->om : om1;
    const double u = comb_add(0.0, sc->c, a.U, a.ld, i, k, a.s);
    a.U[k * a.ld + i] = static_cast<T>(u + om * static_cast<double>(a.v[i]));
  }
}

// idr_a2_kernel with the Schwarz completion fused: covered rows take their
// blocks' mean (schwarz_row_mean) in place of v, which is not written back.
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_a2g_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  const double om1 = a.sc->om;  // one system: loaded once
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double om = kSeg ? sc->om : om1;
    const double u = comb_add(0.0, sc->c, a.U, a.ld, i, k, a.s);
    double v = static_cast<double>(a.v[i]);
    if (a.cover_start[i + 1] > a.cover_start[i])
      v = static_cast<double>(static_cast<T>(
          schwarz_row_mean(a.cover_start, a.ybuf, a.ybuf_stride, static_cast<int32_t>(i))));
    if (a.yc) {
      const int32_t g = a.agg_of_row[i];
      if (g >= 0) v += a.yc[g];
    }
    a.U[k * a.ld + i] = static_cast<T>(u + om * v);
  }
}

// One block per system (segment); the grid is 1 for a single system.  `only`
// (optional, per segment) restricts the reset to the flagged segments.
__global__ void idr_init_scalars(SolverScalars* sc, int s, const int32_t* only = nullptr,
                                 const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  if (only && !only[blockIdx.x]) return;
  sc += blockIdx.x;
  const int t = threadIdx.x;
  for (int e = t; e < kMaxShadow * kMaxShadow; e += blockDim.x) {
    const int i = e / kMaxShadow, k = e % kMaxShadow;
    sc->M[e] = (i == k && i < s) ? 1.0 : 0.0;
  }
  if (t == 0) {
    sc->om = 1.0;
    sc->breakdown = 0;
    sc->beta = 0.0;
    if (!only) sc->done = 0;
  }
}

// A frozen (converged) segment keeps every coefficient at zero, so the vector
// kernels leave its rows unchanged.
__device__ void freeze(SolverScalars* sc) {
  for (int q = 0; q < kMaxShadow; ++q) {
    sc->c[q] = 0.0;
    sc->alpha[q] = 0.0;
  }
  sc->beta = 0.0;
  sc->om = 0.0;
}

// Lane `l` (wave-uniform) of a double, through two v_readlane: the scalar
// recurrences broadcast one lane at a time, so this replaces a ds_bpermute
// round trip per step with register moves (same value).
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Lower-triangular solve M[k:, k:] c[k:] = f[k:] (single thread).
// The s x s recurrences run on the first wavefront of the block, lane i owning
// row i: forward substitutions go column by column (lane l finishes its row,
// broadcasts the result, and the rows below subtract its term), so every row
// takes the same operations in the same order as a row-by-row substitution.
// c(k:s) = M(k:s, k:s)^-1 f(k:s), lower triangular.
__device__ void solve_c(SolverScalars* sc, int s, int k) {
  const int i = threadIdx.x;  // lane, < 64
  double t = i >= k && i < s ? sc->f[i] : 0.0;
  int singular = 0;
  for (int l = k; l < s; ++l) {
    double cl = 0.0;
    if (i == l) {
      const double d = sc->M[l * kMaxShadow + l];
      cl = d != 0.0 ? t / d : 0.0;
      sc->c[l] = cl;
      singular = d == 0.0;
    }
    cl = lane_value(cl, l);
    if (i > l && i < s) t -= sc->M[i * kMaxShadow + l] * cl;
  }
  if (__any(singular) && i == 0) sc->breakdown = 1;
}

// The single-thread recurrences below work on an LDS copy of the system's
// scalars (all threads copy in and out), not on global memory.
static_assert(sizeof(SolverScalars) % sizeof(double) == 0, "copied as doubles");
constexpr int kScalarWords = static_cast<int>(sizeof(SolverScalars) / sizeof(double));
static_assert(kScalarWords <= kReduceThreads, "one scalar word per thread");
// The scalar kernels' copy-in in two halves: each thread's word is loaded
// before the partial loads are issued (both round trips overlap) and stored to
// LDS after the reduction.
__device__ __forceinline__ double scalars_word(const SolverScalars* from) {
  return threadIdx.x < kScalarWords ? reinterpret_cast<const double*>(from)[threadIdx.x] : 0.0;
}
__device__ __forceinline__ void scalars_store(double w, SolverScalars* to) {
  if (threadIdx.x < kScalarWords) reinterpret_cast<double*>(to)[threadIdx.x] = w;
  __syncthreads();
}
__device__ void scalars_copy(const SolverScalars* from, SolverScalars* to) {
  const double* src = reinterpret_cast<const double*>(from);
  double* dst = reinterpret_cast<double*>(to);
  for (int e = threadIdx.x; e < static_cast<int>(sizeof(SolverScalars) / sizeof(double));
       e += blockDim.x)
    dst[e] = src[e];
  __syncthreads();
}

// Cycle start: f = P^T r (from partials), c = M^-1 f.
__global__ void __launch_bounds__(kReduceThreads)
    idr_scalar_f(SolverScalars* sc_g, const double* partials, int nb, int s,
                 const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  sc_g += blockIdx.x;
  partials += static_cast<size_t>(blockIdx.x) * nb * kPartialStride;
  __shared__ double d[kPartialStride];
  __shared__ SolverScalars lds;
  const double word = scalars_word(sc_g);
  reduce_partials(partials, nb, d);
  scalars_store(word, &lds);
  SolverScalars* sc = &lds;
  if (threadIdx.x < 64) {  // the first wavefront
    if (sc->done) {
      if (threadIdx.x == 0) freeze(sc);
    } else {
      if (static_cast<int>(threadIdx.x) < s) sc->f[threadIdx.x] = d[threadIdx.x];
      solve_c(sc, s, 0);
    }
  }
  __syncthreads();
  scalars_copy(&lds, sc_g);
}

// After G_k = A U_k: biorthogonalisation coefficients, new column k of M,
// beta, f update and c for the next k.
__device__ void scalar_after_body(SolverScalars* sc, const double* d, int s, int k);

__global__ void __launch_bounds__(kReduceThreads)
    idr_scalar_after(SolverScalars* sc_g, const double* partials, int nb, int s, int k,
                     const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  sc_g += blockIdx.x;
  partials += static_cast<size_t>(blockIdx.x) * nb * kPartialStride;
  __shared__ double d[kPartialStride];
  __shared__ SolverScalars lds;
  const double word = scalars_word(sc_g);
  reduce_partials(partials, nb, d);
  scalars_store(word, &lds);
  if (threadIdx.x < 64) {  // the first wavefront
    if (lds.done) {
      if (threadIdx.x == 0) freeze(&lds);
    } else {
      scalar_after_body(&lds, d, s, k);
    }
  }
  __syncthreads();
  scalars_copy(&lds, sc_g);
}

__device__ void scalar_after_body(SolverScalars* sc, const double* d, int s, int k) {
  const int i = threadIdx.x;  // lane, < 64
  // alpha(0:k) = M(0:k, 0:k)^-1 d(0:k)
  double t = i < k ? d[i] : 0.0;
  double my_alpha = 0.0;
  for (int l = 0; l < k; ++l) {
    double al = 0.0;
    if (i == l) {
      const double mii = sc->M[l * kMaxShadow + l];
      al = mii != 0.0 ? t / mii : 0.0;
      sc->alpha[l] = al;
      my_alpha = al;
    }
    al = lane_value(al, l);
    if (i > l && i < k) t -= al * sc->M[i * kMaxShadow + l];
  }
  // new column k of M: d(k:s) minus the projections on the earlier columns
  double mk = i >= k && i < s ? d[i] : 0.0;
  for (int l = 0; l < k; ++l) {
    const double al = lane_value(my_alpha, l);
    if (i >= k && i < s) mk -= al * sc->M[i * kMaxShadow + l];
  }
  if (i >= k && i < s) sc->M[i * kMaxShadow + k] = mk;
  const double mkk = lane_value(mk, k);
  double beta = 0.0;
  if (mkk == 0.0 || !isfinite(mkk)) {
    if (i == 0) sc->breakdown = 1;
  } else {
    beta = sc->f[k] / mkk;
  }
  if (i == 0) sc->beta = beta;
  if (i > k && i < s) sc->f[i] -= beta * mk;
  if (k + 1 < s) solve_c(sc, s, k + 1);
}

// omega = (t.r)/(t.t) with the 0.7 "maintaining the convergence" safeguard.
__global__ void __launch_bounds__(kReduceThreads)
    idr_scalar_omega(SolverScalars* sc, const double* part_t, const double* part_r, int nb_t,
                     int nb_r, int rr_index = 0, const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  sc += blockIdx.x;
  part_t += static_cast<size_t>(blockIdx.x) * nb_t * kPartialStride;
  part_r += static_cast<size_t>(blockIdx.x) * nb_r * kPartialStride;
  __shared__ double dt[kPartialStride];
  __shared__ double dr[kPartialStride];
  reduce_partials(part_t, nb_t, dt);
  reduce_partials(part_r, nb_r, dr);
  if (threadIdx.x != 0) return;
  if (sc->done) {
    freeze(sc);
    return;
  }
  const double tr = dt[0], tt = dt[1], rr = dr[rr_index];
  if (!(tt > 0.0) || !(rr > 0.0)) {
    sc->om = 0.0;
    if (!(rr > 0.0)) return;  // already converged exactly
    sc->breakdown = 1;
    return;
  }
  double om = tr / tt;
  const double rho = fabs(tr) / (sqrt(tt) * sqrt(rr));
  if (rho < 0.7) om *= 0.7 / rho;
  sc->om = om;
}

__global__ void __launch_bounds__(kReduceThreads)
    scalar_norm(SolverScalars* sc, const double* partials, int nb,
                const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  sc += blockIdx.x;
  partials += static_cast<size_t>(blockIdx.x) * nb * kPartialStride;
  __shared__ double d[kPartialStride];
  reduce_partials(partials, nb, d);
  if (threadIdx.x == 0) sc->rr = d[0];
}

// v = D^-1 (r - sum_{q>=k} c_q G_q);  U_k = sum_{q>=k} c_q U_q + om v
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_a_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double om = sc->om;
    double v = comb_sub(a.r[i], sc->c, a.G, a.ld, i, k, a.s);
    const double u = comb_add(0.0, sc->c, a.U, a.ld, i, k, a.s);
    v *= a.pdiag[i];
    a.U[k * a.ld + i] = static_cast<T>(u + om * v);
  }
}

// G_k, U_k biorthogonalised; r -= beta G_k; x += beta U_k; partial r.r;
// optionally v = D^-1 r for the dimension-reduction step.
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_b_kernel(IdrVec<T> a, int k, int write_v) {
  if (launch_skipped(a.skip)) return;
  double acc[1] = {0.0};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double beta = sc->beta;
    double g = static_cast<double>(a.G[k * a.ld + i]);
    double u = static_cast<double>(a.U[k * a.ld + i]);
    comb_sub2(g, u, sc->alpha, a.G, a.U, a.ld, i, k);
    a.G[k * a.ld + i] = static_cast<T>(g);
    a.U[k * a.ld + i] = static_cast<T>(u);
    const double r = a.r[i] - beta * g;
    a.r[i] = r;
    a.x[i] += beta * u;
    acc[0] += r * r;
    // dimension-reduction input: pdiag r (the Schwarz blocks then fill their rows)
    if (write_v) a.v[i] = static_cast<T>(a.pdiag[i] * r);
  }
  block_reduce_store<1>(acc, 1, a.partials);
}

// idr_b_kernel(k) fused with the first half of step k + 1 (both read only the
// scalars idr_scalar_after(k) wrote): after the biorthogonalisation and the r, x
// updates, v = pdiag (r - sum_{q>k} c_q G_q); without Schwarz blocks also
// U_{k+1} = sum_{q>k} c_q U_q + om v (idr_a_kernel).
template <typename T, bool kBlocks, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_ba_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  double acc[1] = {0.0};
  const double beta1 = a.sc->beta, om1 = a.sc->om;  // one system: loaded once
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const SolverScalars* sc = row_sc<kSeg>(a, i);
    const double beta = kSeg ? sc->beta : beta1, om = kSeg ? sc->om : om1;
    double g = static_cast<double>(a.G[k * a.ld + i]);
    double u = static_cast<double>(a.U[k * a.ld + i]);
    comb_sub2(g, u, sc->alpha, a.G, a.U, a.ld, i, k);
    a.G[k * a.ld + i] = static_cast<T>(g);
    a.U[k * a.ld + i] = static_cast<T>(u);
    const double r = a.r[i] - beta * g;
    a.r[i] = r;
    a.x[i] += beta * u;
    acc[0] += r * r;
    double v = comb_sub(r, sc->c, a.G, a.ld, i, k + 1, a.s);
    v *= a.pdiag[i];
    if (kBlocks) {
      a.v[i] = static_cast<T>(v);
    } else {
      const double un = comb_add(0.0, sc->c, a.U, a.ld, i, k + 1, a.s);
      a.U[(k + 1) * a.ld + i] = static_cast<T>(un + om * v);
    }
  }
  block_reduce_store<1>(acc, 1, a.partials);
}

// ---- one system, shadow dimension fixed at compile time ---------------------
//
// The generic kernels above loop over a run-time number of basis columns and
// read the coefficients from the scalars struct in global memory, which the
// compiler must reload after every store (it may alias), so each row waits
// for several memory round trips.  For the default shadow dimensions the
// launch below picks a kernel with S fixed and dispatches on k inside: the
// coefficients are read once into registers and every column a row needs is
// loaded before the first use.  The arithmetic is the generic kernels', term
// for term in the same order, so results are bit-identical.

// Calls f(std::integral_constant<int, K>{}) for K = k, 0 <= k < S.
template <int S, typename F>
__device__ __forceinline__ void with_step(int k, F&& f) {
  static_assert(S >= 1 && S <= kMaxShadow, "shadow dimension");
  switch (k) {
#define LSMPS_STEP(K) \
  case K:             \
    if constexpr (K < S) f(std::integral_constant<int, K>{}); \
    break;
    LSMPS_STEP(0) LSMPS_STEP(1) LSMPS_STEP(2) LSMPS_STEP(3) LSMPS_STEP(4) LSMPS_STEP(5)
    LSMPS_STEP(6) LSMPS_STEP(7) LSMPS_STEP(8) LSMPS_STEP(9) LSMPS_STEP(10) LSMPS_STEP(11)
    LSMPS_STEP(12) LSMPS_STEP(13) LSMPS_STEP(14) LSMPS_STEP(15)
#undef LSMPS_STEP
    default:
      break;
  }
}

// idr_a2g_kernel for one system with s = S.
template <typename T, int S>
__global__ void __launch_bounds__(kBlock) idr_a2g_fixed_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  const double om = a.sc->om;
  double cs[kMaxShadow];  // cs[j] = c[k + j]
#pragma unroll
  for (int j = 0; j < kMaxShadow; ++j) cs[j] = j < S - k ? a.sc->c[k + j] : 0.0;
  with_step<S>(k, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr int N = S - K;
    const T* U = a.U + static_cast<size_t>(K) * a.ld;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
      double e[N];
#pragma unroll
      for (int j = 0; j < N; ++j) e[j] = static_cast<double>(U[static_cast<size_t>(j) * a.ld + i]);
      double v = static_cast<double>(a.v[i]);
      const int32_t lo = a.cover_start[i], hi = a.cover_start[i + 1];
      const int32_t g = a.yc ? a.agg_of_row[i] : -1;
      double u = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) u += cs[j] * e[j];
      if (hi > lo)
        v = static_cast<double>(static_cast<T>(
            schwarz_row_mean(a.cover_start, a.ybuf, a.ybuf_stride, static_cast<int32_t>(i))));
      if (g >= 0) v += a.yc[g];
      a.U[static_cast<size_t>(K) * a.ld + i] = static_cast<T>(u + om * v);
    }
  });
}

// idr_ba_kernel for one system with s = S.
template <typename T, bool kBlocks, int S>
__global__ void __launch_bounds__(kBlock) idr_ba_fixed_kernel(IdrVec<T> a, int k) {
  if (launch_skipped(a.skip)) return;
  double acc[1] = {0.0};
  const double beta = a.sc->beta, om = a.sc->om;
  double al[kMaxShadow], cs[kMaxShadow];  // al[q] = alpha[q] (q < k), cs[j] = c[k + 1 + j]
#pragma unroll
  for (int j = 0; j < kMaxShadow; ++j) {
    al[j] = j < k ? a.sc->alpha[j] : 0.0;
    cs[j] = j < S - 1 - k ? a.sc->c[k + 1 + j] : 0.0;
  }
  with_step<S>(k, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr int N = S - 1 - K;  // columns after k
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
      const size_t ok = static_cast<size_t>(K) * a.ld + i;
      double g = static_cast<double>(a.G[ok]);
      double u = static_cast<double>(a.U[ok]);
      double eg[K > 0 ? K : 1], eu[K > 0 ? K : 1], en[N > 0 ? N : 1], un_e[N > 0 && !kBlocks ? N : 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        eg[q] = static_cast<double>(a.G[static_cast<size_t>(q) * a.ld + i]);
        eu[q] = static_cast<double>(a.U[static_cast<size_t>(q) * a.ld + i]);
      }
#pragma unroll
      for (int j = 0; j < N; ++j) en[j] = static_cast<double>(a.G[static_cast<size_t>(K + 1 + j) * a.ld + i]);
      if constexpr (!kBlocks) {
#pragma unroll
        for (int j = 0; j < N; ++j) un_e[j] = static_cast<double>(a.U[static_cast<size_t>(K + 1 + j) * a.ld + i]);
      }
      const double ri = a.r[i];
      const double xi = a.x[i];
      const double pd = a.pdiag[i];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        g -= al[q] * eg[q];
        u -= al[q] * eu[q];
      }
      a.G[ok] = static_cast<T>(g);
      a.U[ok] = static_cast<T>(u);
      const double r = ri - beta * g;
      a.r[i] = r;
      a.x[i] = xi + beta * u;
      acc[0] += r * r;
      double v = r;
#pragma unroll
      for (int j = 0; j < N; ++j) v -= cs[j] * en[j];
      v *= pd;
      if constexpr (kBlocks) {
        a.v[i] = static_cast<T>(v);
      } else {
        double un = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) un += cs[j] * un_e[j];
        a.U[static_cast<size_t>(K + 1) * a.ld + i] = static_cast<T>(un + om * v);
      }
    }
  });
  block_reduce_store<1>(acc, 1, a.partials);
}

// x += om v; r -= om t; partial r.r
template <typename T, bool kSeg = false>
__global__ void __launch_bounds__(kBlock) idr_c_kernel(IdrVec<T> a) {
  if (launch_skipped(a.skip)) return;
  double acc[1] = {0.0};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double om = row_sc<kSeg>(a, i)->om;
    a.x[i] += om * static_cast<double>(a.v[i]);
    const double r = a.r[i] - om * static_cast<double>(a.t[i]);
    a.r[i] = r;
    acc[0] += r * r;
  }
  block_reduce_store<1>(acc, 1, a.partials);
}

// f partials: P_q . r
template <typename T>
__global__ void __launch_bounds__(kBlock) shadow_dots_kernel(IdrVec<T> a) {
  if (launch_skipped(a.skip)) return;
  double acc[kMaxShadow];
#pragma unroll
  for (int q = 0; q < kMaxShadow; ++q) acc[q] = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double r = a.r[i];
#pragma unroll
    for (int q = 0; q < kMaxShadow; ++q)
      if (q < a.s) acc[q] += shadow(a.p0, i, q) * r;
  }
  block_reduce_store<kMaxShadow>(acc, a.s, a.partials);
}

// ---- batched systems: per-segment reductions and bookkeeping -------------

constexpr int kSegBlocks = 16;  // partial records per segment

// Per-segment dot products, kSegBlocks blocks per segment, fixed order.
// MODE 1: P_q . y (shadow dots of an SpMV result); 2: t.r, t.t, r.r (y = t);
// 3: r.r; 4: P_q . r.
template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) seg_reduce_kernel(const int32_t* start, int s,
                                                            const T* p0, const T* y,
                                                            const double* r,
                                                            double* partials) {
  constexpr int NQ = (MODE == 1 || MODE == 4) ? kMaxShadow : (MODE == 2 ? 3 : 1);
  const int seg = blockIdx.x / kSegBlocks;
  const int part = blockIdx.x % kSegBlocks;
  const int64_t b = start[seg], e = start[seg + 1];
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  for (int64_t i = b + static_cast<int64_t>(part) * blockDim.x + threadIdx.x; i < e;
       i += static_cast<int64_t>(kSegBlocks) * blockDim.x) {
    if (MODE == 1 || MODE == 4) {
      const double w = MODE == 1 ? static_cast<double>(y[i]) : r[i];
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q < s) acc[q] += shadow(p0, static_cast<size_t>(i), q) * w;
    } else if (MODE == 2) {
      const double t = static_cast<double>(y[i]);
      acc[0] += t * r[i];
      acc[1 % NQ] += t * t;
      acc[2 % NQ] += r[i] * r[i];
    } else {
      acc[0] += r[i] * r[i];
    }
  }
  block_reduce_store<NQ>(acc, (MODE == 1 || MODE == 4) ? s : NQ, partials);
}

// Per segment: rr and the breakdown flag, packed for one device-to-host copy.
__global__ void seg_status_kernel(const SolverScalars* sc, int count, double* out) {
  const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  out[2 * k] = sc[k].rr;
  out[2 * k + 1] = static_cast<double>(sc[k].breakdown);
}

// Marks segments converged (flag 1) or, for flag 2, due for a restart.
__global__ void seg_flags_kernel(SolverScalars* sc, int count, const int32_t* flags) {
  const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  if (flags[k] == 1) sc[k].done = 1;
}

// Restart of the flagged segments' Krylov state: G = U = 0 and P_0 = r on their rows.
template <typename T>
__global__ void seg_restart_kernel(size_t n, const int32_t* row_seg, const int32_t* flags,
                                   size_t ld, int s, const double* r, T* p0, T* G, T* U) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || flags[row_seg[i]] != 2) return;
  p0[i] = static_cast<T>(r[i]);
  for (int q = 0; q < s; ++q) {
    G[q * ld + i] = T(0);
    U[q * ld + i] = T(0);
  }
}

// ---- PCG ----------------------------------------------------------------
__global__ void __launch_bounds__(kReduceThreads)
    cg_scalar_rz(SolverScalars* sc, const double* part, int nb) {
  __shared__ double d[kPartialStride];
  reduce_partials(part, nb, d);
  const double rz = d[0], rr = d[1];
  if (threadIdx.x == 0) {
    sc->cg_beta = sc->rz != 0.0 ? rz / sc->rz : 0.0;
    sc->rz = rz;
    sc->rr = rr;
  }
}
__global__ void __launch_bounds__(kReduceThreads)
    cg_scalar_alpha(SolverScalars* sc, const double* part, int nb) {
  __shared__ double d[kPartialStride];
  reduce_partials(part, nb, d);
  const double pq = d[0];
  if (threadIdx.x == 0) {
    if (pq == 0.0 || !isfinite(pq)) {
      sc->breakdown = 1;
      sc->cg_alpha = 0.0;
    } else {
      sc->cg_alpha = sc->rz / pq;
    }
  }
}
// x += alpha p; r -= alpha q; partials r.z, r.r
__global__ void __launch_bounds__(kBlock) cg_update_kernel(size_t n, double* x, double* r,
                                                           const double* p, const double* q,
                                                           const double* dinv,
                                                           const SolverScalars* sc,
                                                           double* partials) {
  const double alpha = sc->cg_alpha;
  double acc[2] = {0.0, 0.0};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc[0] += rv * rv * dinv[i];
    acc[1] += rv * rv;
  }
  block_reduce_store<2>(acc, 2, partials);
}
// p = D^-1 r + beta p  (first call with beta_zero: p = D^-1 r)
__global__ void cg_direction_kernel(size_t n, double* p, const double* r, const double* dinv,
                                    const SolverScalars* sc, int first) {
  const double beta = first ? 0.0 : sc->cg_beta;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = dinv[i] * r[i] + (first ? 0.0 : beta * p[i]);
}
__global__ void rz_kernel(size_t n, const double* r, const double* dinv, double* partials) {
  double acc[2] = {0.0, 0.0};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    acc[0] += r[i] * r[i] * dinv[i];
    acc[1] += r[i] * r[i];
  }
  block_reduce_store<2>(acc, 2, partials);
}

// ---- host orchestration -------------------------------------------------

unsigned vec_blocks(size_t n) {
  const size_t b = (n + kBlock - 1) / kBlock;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, kMaxVectorBlocks)));
}
unsigned spmv_blocks(size_t n) {
  const size_t rows_per_block = kBlock / kLanes;
  const size_t b = (n + rows_per_block - 1) / rows_per_block;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, kMaxReduceBlocks)));
}

// A multiple of the XCD count (see xcd_block), at most kMaxReduceBlocks.
// SpMV grid: one workgroup per group up to kMaxReduceBlocks (each writes one
// partial record), a multiple of the 8 XCDs.
unsigned sell_blocks(size_t groups) {
  size_t b = std::min<size_t>(std::max<size_t>(groups, 1), kMaxReduceBlocks);
  b = (b + kXcds - 1) / kXcds * kXcds;
  return static_cast<unsigned>(std::min<size_t>(b, kMaxReduceBlocks));
}

// Converts the CSR (FP64 values) into the DIA + SELL-64 layout with values of type V.
template <typename V>
void build_sell(Context& ctx, SellMatrix& M, const lsmps_csr& A_in, V* out_vals_dummy_type) {
  (void)out_vals_dummy_type;
  const size_t n = A_in.rows;
  require(n <= kMaxGatherRows, "PPE solve: the system has more rows than the SpMV's 32-bit "
                               "gather offsets address (2^28)");
  // row-sorted copy of the matrix (same rows and values, columns ascending)
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, A_in.row_offsets + n, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  const size_t used = static_cast<size_t>(std::max(ctx.host_i32[0], 0));
  M.sorted_cols.reserve(std::max<size_t>(used, 1));
  M.sorted_vals.reserve(std::max<size_t>(used, 1));
  {
    const size_t waves = std::min<size_t>(n, static_cast<size_t>(kMaxVectorBlocks) * 16);
    const unsigned blocks = static_cast<unsigned>(std::max<size_t>(1, (waves * 64 + kBlock - 1) / kBlock));
    sort_rows_kernel<<<blocks, kBlock, 0, ctx.stream>>>(n, A_in.row_offsets, A_in.col_indices,
                                                        A_in.values, M.sorted_cols.data(),
                                                        M.sorted_vals.data());
    LSMPS_LAUNCH_CHECK();
  }
  lsmps_csr A = A_in;
  A.col_indices = M.sorted_cols.data();
  A.values = M.sorted_vals.data();
  M.n = n;
  M.chunks = (n + kSellChunk - 1) / kSellChunk;
  M.groups = (M.chunks + kDiaGroup - 1) / kDiaGroup;
  const size_t chunks = M.chunks;
  const size_t groups = M.groups;
  M.dia_count.reserve(groups);
  M.dia_offsets.reserve(groups * kMaxDia);
  M.dia_slot.reserve(groups * kMaxDia);
  M.dia_plan.reserve(groups * kMaxDia);
  M.dia_runs.reserve(groups * kRunClasses);
  M.dia_slots.reserve(groups);
  M.dia_start.reserve(groups + 1);
  M.chunk_width.reserve(chunks);
  M.chunk_scaled.reserve(chunks);
  M.chunk_start.reserve(chunks + 1);
  M.col_rows.reserve(chunks);
  M.col_start.reserve(chunks + 1);
  M.nwin.reserve(chunks);
  M.windows.reserve(chunks * kSellWindows);
  M.rem_counter.reserve(1);
  fill_kernel<unsigned long long><<<1, kBlock, 0, ctx.stream>>>(1, M.rem_counter.data(), 0ull);
  // each group's diagonals and their plan, then each chunk's remainder widths
  // and column windows
  dia_group_kernel<<<static_cast<unsigned>(groups), kGroupRows, 0, ctx.stream>>>(
      n, A.row_offsets, A.col_indices, M.dia_count.data(), M.dia_offsets.data(),
      M.dia_slot.data(), M.dia_plan.data(), M.dia_runs.data(), M.dia_slots.data());
  LSMPS_LAUNCH_CHECK();
  rem_select_kernel<<<static_cast<unsigned>(chunks), kSellChunk, 0, ctx.stream>>>(
      n, A.row_offsets, A.col_indices, M.dia_count.data(), M.dia_offsets.data(),
      M.chunk_width.data(), M.chunk_scaled.data(), M.rem_counter.data(), M.nwin.data(),
      M.windows.data(), M.col_rows.data());
  LSMPS_LAUNCH_CHECK();
  exclusive_scan_i32(ctx, M.dia_slots.data(), M.dia_start.data(), groups);
  exclusive_scan_i32(ctx, M.chunk_scaled.data(), M.chunk_start.data(), chunks);
  exclusive_scan_i32(ctx, M.col_rows.data(), M.col_start.data(), chunks);
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, M.chunk_start.data() + chunks, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32 + 1, A.row_offsets + n, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32 + 2, M.col_start.data() + chunks, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32 + 3, M.dia_start.data() + groups, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  LSMPS_HIP(hipMemcpyAsync(ctx.host_f64, M.rem_counter.data(), sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  require(ctx.host_i32[0] >= 0 && ctx.host_i32[2] >= 0 && ctx.host_i32[3] >= 0,
          "PPE matrix too large for the 32-bit SELL layout");
  M.entries = static_cast<size_t>(ctx.host_i32[0]);
  M.nnz = static_cast<size_t>(ctx.host_i32[1]);
  M.col_units = static_cast<size_t>(ctx.host_i32[2]) * kSellChunk;
  M.dia_entries = static_cast<size_t>(ctx.host_i32[3]);
  unsigned long long rem = 0;
  std::memcpy(&rem, ctx.host_f64, sizeof(rem));
  M.rem_nnz = static_cast<size_t>(rem);
  M.cols.reserve(std::max<size_t>(M.col_units, 1));
  V* vals;
  V* dia;
  if (sizeof(V) == sizeof(double)) {
    M.vals64.reserve(std::max<size_t>(M.entries, 1));
    M.dia64.reserve(std::max<size_t>(M.dia_entries, 1));
    vals = reinterpret_cast<V*>(M.vals64.data());
    dia = reinterpret_cast<V*>(M.dia64.data());
  } else {
    M.vals32.reserve(std::max<size_t>(M.entries, 1));
    M.dia32.reserve(std::max<size_t>(M.dia_entries, 1));
    vals = reinterpret_cast<V*>(M.vals32.data());
    dia = reinterpret_cast<V*>(M.dia32.data());
  }
  fill_kernel<V><<<blocks_for(std::max<size_t>(M.dia_entries, 1)), kBlock, 0, ctx.stream>>>(
      M.dia_entries, dia, V(0));
  sell_fill_kernel<V><<<static_cast<unsigned>(chunks), kSellChunk, 0, ctx.stream>>>(
      n, A.row_offsets, A.col_indices, A.values, M.dia_count.data(),
      M.dia_offsets.data(), M.dia_slot.data(), M.dia_start.data(), dia, M.chunk_start.data(),
      M.chunk_width.data(), M.col_start.data(), M.nwin.data(), M.windows.data(), M.cols.data(),
      vals);
  LSMPS_LAUNCH_CHECK();
}

double read_rr(Context& ctx, PpeSolver& S) {
  LSMPS_HIP(hipMemcpyAsync(S.host, &S.scalars.data()->rr, sizeof(double), hipMemcpyDeviceToHost,
                           ctx.stream));
  ctx.sync();
  return S.host[0];
}

// FP64 true residual r = b - A x; returns ||r||^2.
double true_residual(Context& ctx, PpeSolver& S, const lsmps_csr& A, const double* x,
                     const double* b, double* r, float* r32) {
  const size_t n = A.rows;
  const unsigned nb = spmv_blocks(n);
  residual_kernel<<<nb, kBlock, 0, ctx.stream>>>(n, A.row_offsets, A.col_indices, A.values, x, b,
                                                 r, r32, S.part_b.data());
  LSMPS_LAUNCH_CHECK();
  scalar_norm<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(), static_cast<int>(nb));
  LSMPS_LAUNCH_CHECK();
  return read_rr(ctx, S);
}

// ---- captured IDR cycles ----------------------------------------------------

__global__ void idr_control_init(IdrControl* c, double rr0, double target2, double factor2,
                                 int max_iterations, int s) {
  c->halt = rr0 > target2 && max_iterations > 0 ? 0 : 1;
  c->skip_replace = 1;
  c->skip_reinit = 1;
  c->iterations = 0;
  c->replacements = 0;
  c->max_iterations = max_iterations;
  c->s = s;
  c->pad = 0;
  c->rr = rr0;
  c->rr_replaced = rr0;
  c->target2 = target2;
  c->factor2 = factor2;
}

// First launch of a cycle: once halted, the gated launches are skipped too.
__global__ void idr_control_begin(IdrControl* c) {
  if (c->halt) {
    c->skip_replace = 1;
    c->skip_reinit = 1;
  }
}

// After the cycle's residual norm: the tests the host loop makes on the value
// it reads back (non-finite stop, reliable update, restart after a breakdown).
__global__ void idr_control_cycle(IdrControl* c, const SolverScalars* sc) {
  if (c->halt) return;
  c->iterations += c->s + 1;
  const double rr = sc->rr;
  c->rr = rr;
  if (!isfinite(rr)) {
    c->halt = 1;
    c->skip_replace = 1;
    c->skip_reinit = 1;
    return;
  }
  c->skip_replace = rr <= c->target2 || rr <= c->factor2 * c->rr_replaced ? 0 : 1;
  c->skip_reinit = sc->breakdown ? 0 : 1;
}

// After the gated true-residual launches: take their norm, then the loop test.
__global__ void idr_control_after_replace(IdrControl* c, const SolverScalars* sc) {
  if (c->halt) return;
  if (!c->skip_replace) {
    c->rr = sc->rr;
    c->rr_replaced = c->rr;
    c->replacements += 1;
  }
  if (!(c->rr > c->target2 && c->iterations < c->max_iterations)) c->halt = 1;
}

void append_sell_signature(std::vector<uintptr_t>& sig, const SellMatrix& M) {
  const void* p[] = {M.dia_runs.data(), M.dia_plan.data(),
                     M.dia_start.data(), M.chunk_start.data(), M.chunk_width.data(),
                     M.col_start.data(), M.nwin.data(), M.windows.data(), M.cols.data()};
  for (const void* q : p) sig.push_back(reinterpret_cast<uintptr_t>(q));
  sig.push_back(M.chunks);
}

void append_schwarz_signature(std::vector<uintptr_t>& sig, const SchwarzPrec& P) {
  const void* p[] = {P.members.data(), P.sizes.data(), P.inv32.data(), P.inv64.data(),
                     P.inv16.data(), P.col_scale.data(), P.ybuf.data(), P.covered.data(),
                     P.cover_start.data(), P.entry_of.data(), P.pdiag.data()};
  for (const void* q : p) sig.push_back(reinterpret_cast<uintptr_t>(q));
  sig.push_back(P.blocks);
  sig.push_back(P.covered_count);
  sig.push_back(P.fp32 ? 1 : 0);
  sig.push_back(P.half ? 1 : 0);
}

void append_coarse_signature(std::vector<uintptr_t>& sig, const CoarseSpace& C) {
  const void* p[] = {C.agg_start.data(), C.agg_rows.data(), C.agg_of_row.data(), C.wfac.data(),
                     C.rc.data(), C.yc.data(), C.ainv.data(), C.status.data()};
  for (const void* q : p) sig.push_back(reinterpret_cast<uintptr_t>(q));
  sig.push_back(static_cast<uintptr_t>(C.count));
  sig.push_back(C.n);
}

// Algorithmic bytes of one SpMV y = A x over n rows: the stored nonzeros'
// values, the column indices of the remainder entries (16-bit, or 32-bit in the
// chunks that need it, in the layout's proportion; DIA columns are implicit),
// the per-chunk descriptors, diagonal offsets and window bases, x read once and
// y written once.  Padding is not counted.
double spmv_algorithmic_bytes(const SellMatrix& M, size_t n, size_t value_bytes,
                              size_t vector_bytes) {
  const double col_bytes =
      2.0 * static_cast<double>(M.col_units) / static_cast<double>(std::max<size_t>(M.entries, 1));
  return static_cast<double>(M.nnz) * value_bytes + static_cast<double>(M.rem_nnz) * col_bytes +
         M.chunks * (5.0 + kSellWindows) * sizeof(int32_t) + M.groups * (1.0 + kRunClasses) * sizeof(int32_t) +
         static_cast<double>(M.dia_entries) / kGroupRows * sizeof(int32_t) +
         2.0 * n * vector_bytes;
}

// Whether the SpMVs are timed (LSMPS_FLAG_TIMING); creates the events once.
// With LSMPS_FLAG_KERNEL_TIMING the SpMVs of one IDR cycle in this many are timed.
constexpr int kSpmvTimingStride = 8;

bool spmv_timing_ready(Context& ctx) {
  const bool timing = ctx.timing();
  if (timing && !ctx.timing_ready) {
    for (auto& e : ctx.timing_events) LSMPS_HIP(hipEventCreate(&e));
    ctx.timing_ready = true;
  }
  return timing;
}

// Makes S.cycle_exec[0..1] the captured cycle for `sig`, capturing it (twice:
// SpMV event sets 0 and 1) when the signature changed.  enqueue(halt,
// skip_replace, skip_reinit, set) launches one cycle on ctx.stream.  Returns
// false when the graph path cannot be used (then nothing was left captured).
template <typename F>
bool idr_cycles_captured(Context& ctx, PpeSolver& S, F enqueue, std::vector<uintptr_t> sig) {
  S.control.reserve(1);
  if (!S.host_control)
    LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&S.host_control), 2 * sizeof(IdrControl)));
  for (auto& e : S.cycle_done)
    if (!e) LSMPS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!S.capture_stream) LSMPS_HIP(hipStreamCreateWithFlags(&S.capture_stream, hipStreamNonBlocking));
  sig.push_back(reinterpret_cast<uintptr_t>(S.control.data()));
  sig.push_back(reinterpret_cast<uintptr_t>(S.scalars.data()));
  sig.push_back(reinterpret_cast<uintptr_t>(S.part_a.data()));
  sig.push_back(reinterpret_cast<uintptr_t>(S.part_b.data()));
  if (S.cycle_exec[0] && S.cycle_exec[1] && sig == S.cycle_signature) return true;
  S.release_graphs();
  IdrControl* ctl = S.control.data();
  const int32_t* halt = &ctl->halt;
  const int32_t* skip_replace = &ctl->skip_replace;
  const int32_t* skip_reinit = &ctl->skip_reinit;
  // the capture stream stands in for the library stream while the cycle is
  // recorded (the legacy default stream cannot be captured); nothing may
  // allocate or synchronise in between
  const hipStream_t library_stream = ctx.stream;
  for (int set = 0; set < 2; ++set) {
    hipGraph_t graph = nullptr;
    ctx.stream = S.capture_stream;
    try {
      const AllocationFreeze frozen;
      LSMPS_HIP(hipStreamBeginCapture(S.capture_stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(halt, skip_replace, skip_reinit, set);
      } catch (...) {
        (void)hipStreamEndCapture(S.capture_stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      LSMPS_HIP(hipStreamEndCapture(S.capture_stream, &graph));
      ctx.stream = library_stream;
      const hipError_t e = hipGraphInstantiate(&S.cycle_exec[set], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      LSMPS_HIP(e);
    } catch (const Error& e) {
      // the runtime could not record or instantiate the cycle: the solve runs
      // its launches directly (graph_cycles = 0 in the report says so)
      ctx.stream = library_stream;
      S.release_graphs();
      S.graph_error = e.what();
      (void)hipGetLastError();
      return false;
    }
  }
  S.cycle_signature = std::move(sig);
  S.graph_error.clear();
  return true;
}

// IDR(s) with an FP64 iterate and residual, a Krylov basis of type T and matrix
// values of type V.  x and r hold the initial iterate and its residual.  Each time the recursive residual has
// dropped by `replace_factor` since the last replacement (and when it meets the
// target) it is replaced by the true FP64 residual b - A64 x, keeping the
// Krylov state (reliable updates); the solve stops when the true residual meets
// the target.
template <typename T, typename V>
int idr_solve(Context& ctx, PpeSolver& S, size_t n, const SellMatrix& M, const V* vals,
              const double* dinv, double* x, double* r, T* work, size_t ld, int s, double rr0,
              double target, int max_iterations, const lsmps_csr& A64, const double* b64,
              double replace_factor, int* iterations_out, int* replacements_out,
              double* rr_out) {
  T* v = work;
  T* t = work + ld;
  T* p0 = work + 2 * ld;
  T* G = work + 3 * ld;
  T* U = work + (3 + s) * ld;
  const bool schwarz = S.schwarz.blocks > 0 && S.schwarz.n == n;
  const bool coarse = schwarz && S.coarse.ready && S.coarse.n == n && S.coarse.tanks == 0;
  const V* dia = sizeof(V) == sizeof(double) ? reinterpret_cast<const V*>(M.dia64.data())
                                             : reinterpret_cast<const V*>(M.dia32.data());
  IdrVec<T> iv{n, ld, s, r, x, v, t, G, U, p0, S.scalars.data(), S.part_b.data(),
               schwarz ? S.schwarz.pdiag.data() : dinv, nullptr,
               schwarz ? S.schwarz.cover_start.data() : nullptr,
               schwarz ? S.schwarz.ybuf.data() : nullptr,
               schwarz ? S.schwarz.ybuf_stride : 0,
               coarse ? S.coarse.agg_of_row.data() : nullptr,
               coarse ? S.coarse.yc.data() : nullptr};
  const unsigned vb = vec_blocks(n);
  const unsigned sb = sell_blocks(M.groups);
  fill_kernel<T><<<blocks_for(2 * s * ld), kBlock, 0, ctx.stream>>>(2 * s * ld, G, T(0));
  copy_kernel<double, T><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, r, p0);
  idr_init_scalars<<<1, kBlock, 0, ctx.stream>>>(S.scalars.data(), s);
  LSMPS_LAUNCH_CHECK();
  double rr = rr0;
  int it = 0, replacements = 0;
  const double target2 = target * target;
  const double factor2 = replace_factor * replace_factor;
  // Optional per-launch timing of the SpMVs: events bracket each launch of
  // every kSpmvTimingStride-th cycle (each event pair stalls the stream for a
  // few us, so timing every cycle would slow the solve it measures).
  const bool timing = spmv_timing_ready(ctx);
  const double spmv_bytes = spmv_algorithmic_bytes(M, n, sizeof(V), sizeof(T));
  // One IDR(s) cycle: s steps, each with one SpMV G_k = A U_k (and, with the
  // Schwarz blocks, their products and the coarse term before it), then the
  // omega step's SpMV and the residual norm.  `skip` (null when launched
  // directly) is the captured graphs' halt word; `set` picks the SpMV events,
  // `timed` whether this cycle records them.
  auto cycle_body = [&](const int32_t* skip, int set, bool timed) {
    // SpMV timing: HIP events around each launch (direct launches only)
    auto mark = [&](int slot) {
      if (timed)
        LSMPS_HIP(hipEventRecord(ctx.timing_events[set * Context::kTimingSet + slot], ctx.stream));
    };
    iv.skip = skip;
    iv.partials = S.part_a.data();
    shadow_dots_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv);
    idr_scalar_f<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_a.data(),
                                                       static_cast<int>(vb), s, skip);
    // step 0's first half; later steps get it fused into the previous step's b
    if (schwarz) {
      idr_a1_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv, 0);
    } else {
      idr_a_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv, 0);
    }
    for (int k = 0; k < s; ++k) {
      if (schwarz) {
        if (coarse) {  // block products and restriction in one launch, then yc
          products_and_restrict<T>(ctx, S.schwarz, S.coarse, v, v, true, skip);
          coarse_product(ctx, S.coarse, skip);
        } else {
          schwarz_products<T>(ctx, S.schwarz, v, skip);
        }
        if (s == 14) {
          idr_a2g_fixed_kernel<T, 14><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        } else if (s == 16) {
          idr_a2g_fixed_kernel<T, 16><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        } else {
          idr_a2g_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        }
      }
      SellArgs<T, V> sa{n, M.chunks, M.groups, M.dia_runs.data(), M.dia_plan.data(),
                        M.dia_start.data(), dia, M.chunk_start.data(), M.chunk_width.data(),
                        M.col_start.data(), M.nwin.data(), M.windows.data(), M.cols.data(),
                        vals, U + k * ld, G + k * ld, p0, s, nullptr, S.part_a.data(), skip};
      mark(2 * k);
      sell_spmv_kernel<T, V, 1><<<sb, kSellBlock, 0, ctx.stream>>>(sa);
      mark(2 * k + 1);
      idr_scalar_after<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_a.data(),
                                                             static_cast<int>(sb), s, k, skip);
      iv.partials = S.part_b.data();
      if (schwarz && (s == 14 || s == 16)) {
        // at k = s - 1 there are no later columns: the fixed kernel's a part
        // is then v = pdiag r, what idr_b_kernel(k, write_v) writes
        if (s == 14) {
          idr_ba_fixed_kernel<T, true, 14><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        } else {
          idr_ba_fixed_kernel<T, true, 16><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        }
      } else if (k == s - 1) {
        idr_b_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv, k, 1);
      } else if (schwarz) {
        idr_ba_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
      } else {
        if (s == 14) {
          idr_ba_fixed_kernel<T, false, 14><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        } else if (s == 16) {
          idr_ba_fixed_kernel<T, false, 16><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        } else {
          idr_ba_kernel<T, false><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
        }
      }
    }
    if (schwarz) {
      if (coarse) {  // products of v and restriction of r in one launch
        products_and_restrict<T>(ctx, S.schwarz, S.coarse, v, r, false, skip);
        coarse_product(ctx, S.coarse, skip);
        schwarz_complete_prolong_add<T>(ctx, S.schwarz, S.coarse, v, skip);
      } else {
        schwarz_apply<T>(ctx, S.schwarz, v, skip);
      }
    }
    SellArgs<T, V> st{n, M.chunks, M.groups, M.dia_runs.data(), M.dia_plan.data(),
                      M.dia_start.data(), dia, M.chunk_start.data(), M.chunk_width.data(),
                      M.col_start.data(), M.nwin.data(), M.windows.data(), M.cols.data(), vals,
                      v, t, nullptr, s, r, S.part_a.data(), skip};
    mark(2 * s);
    sell_spmv_kernel<T, V, 2><<<sb, kSellBlock, 0, ctx.stream>>>(st);
    mark(2 * s + 1);
    idr_scalar_omega<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_a.data(),
                                                           S.part_b.data(), static_cast<int>(sb),
                                                           static_cast<int>(vb), 0, skip);
    iv.partials = S.part_b.data();
    idr_c_kernel<T><<<vb, kBlock, 0, ctx.stream>>>(iv);
    scalar_norm<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(),
                                                      static_cast<int>(vb), skip);
    LSMPS_LAUNCH_CHECK();
    iv.skip = nullptr;
  };
  // reliable update: r = b - A64 x in FP64 and its norm
  auto replace_body = [&](const int32_t* skip) {
    const unsigned nb = spmv_blocks(n);
    residual_kernel<<<nb, kBlock, 0, ctx.stream>>>(n, A64.row_offsets, A64.col_indices,
                                                   A64.values, x, b64, r, nullptr,
                                                   S.part_b.data(), skip);
    scalar_norm<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(),
                                                      static_cast<int>(nb), skip);
    LSMPS_LAUNCH_CHECK();
  };
  // restart after a breakdown: fresh basis, shadow space and scalars
  auto reinit_body = [&](const int32_t* skip) {
    fill_kernel<T><<<blocks_for(2 * s * ld), kBlock, 0, ctx.stream>>>(2 * s * ld, G, T(0), skip);
    copy_kernel<double, T><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, r, p0, skip);
    idr_init_scalars<<<1, kBlock, 0, ctx.stream>>>(S.scalars.data(), s, nullptr, skip);
    LSMPS_LAUNCH_CHECK();
  };
  auto add_timing = [&](int set) {
    for (int k = 0; k <= s; ++k) {
      float ms = 0.0f;
      LSMPS_HIP(hipEventElapsedTime(&ms, ctx.timing_events[set * Context::kTimingSet + 2 * k],
                                    ctx.timing_events[set * Context::kTimingSet + 2 * k + 1]));
      ctx.spmv_ms += ms;
      ctx.spmv_bytes += spmv_bytes + (k == s ? n * sizeof(double) : 0.0);  // mode 2 also reads r
      ++ctx.spmv_launches;
    }
  };
  bool use_graph = false;
  // HIP events recorded inside a graph cannot be timed on this runtime, so
  // kernel timing keeps the direct launches
  if (S.graph_requested && !timing && rr > target2 && max_iterations > 0) {
    std::vector<uintptr_t> sig{
        reinterpret_cast<uintptr_t>(x), reinterpret_cast<uintptr_t>(r),
        reinterpret_cast<uintptr_t>(work), ld, n, static_cast<uintptr_t>(s),
        static_cast<uintptr_t>(vb), static_cast<uintptr_t>(sb), reinterpret_cast<uintptr_t>(vals),
        reinterpret_cast<uintptr_t>(dia), reinterpret_cast<uintptr_t>(dinv),
        reinterpret_cast<uintptr_t>(A64.row_offsets), reinterpret_cast<uintptr_t>(A64.col_indices),
        reinterpret_cast<uintptr_t>(A64.values), reinterpret_cast<uintptr_t>(b64),
        schwarz ? 1u : 0u, coarse ? 1u : 0u, sizeof(V), sizeof(T)};
    append_sell_signature(sig, M);
    if (schwarz) append_schwarz_signature(sig, S.schwarz);
    if (coarse) append_coarse_signature(sig, S.coarse);
    use_graph = idr_cycles_captured(
        ctx, S,
        [&](const int32_t* halt, const int32_t* skip_replace, const int32_t* skip_reinit,
            int set) {
          IdrControl* ctl = S.control.data();
          idr_control_begin<<<1, 1, 0, ctx.stream>>>(ctl);
          cycle_body(halt, set, false);
          idr_control_cycle<<<1, 1, 0, ctx.stream>>>(ctl, S.scalars.data());
          replace_body(skip_replace);
          idr_control_after_replace<<<1, 1, 0, ctx.stream>>>(ctl, S.scalars.data());
          reinit_body(skip_reinit);
          LSMPS_LAUNCH_CHECK();
        },
        std::move(sig));
  }
  if (use_graph) {
    // two cycles in flight: while cycle c runs, the host waits for cycle c - 1's
    // control record and enqueues cycle c + 1; the cycle queued after the one
    // that halts returns at once from every launch
    idr_control_init<<<1, 1, 0, ctx.stream>>>(S.control.data(), rr, target2, factor2,
                                              max_iterations, s);
    LSMPS_LAUNCH_CHECK();
    auto launch = [&](int slot) {
      LSMPS_HIP(hipGraphLaunch(S.cycle_exec[slot], ctx.stream));
      LSMPS_HIP(hipMemcpyAsync(S.host_control + slot, S.control.data(), sizeof(IdrControl),
                               hipMemcpyDeviceToHost, ctx.stream));
      LSMPS_HIP(hipEventRecord(S.cycle_done[slot], ctx.stream));
    };
    launch(0);
    launch(1);
    int replays = 2;
    IdrControl last{};
    for (int c = 0;; ++c) {
      const int slot = c & 1;
      LSMPS_HIP(hipEventSynchronize(S.cycle_done[slot]));
      last = S.host_control[slot];
      if (last.halt) break;
      launch(slot);
      ++replays;
    }
    it = last.iterations;
    replacements = last.replacements;
    rr = last.rr;
    S.last_graph_cycles = replays;
  } else {
    // Direct launches under the same device-side loop control as the graph
    // replays (the loop tests run in one-thread kernels, every later launch of
    // a halted solve returns at once), with two cycles in flight: while cycle
    // c runs, the host waits for cycle c - 1's control record and enqueues
    // cycle c + 1.  Cycle c's SpMV events are event set c & 1; a set is read
    // back after its cycle's record and before the set is recorded again.
    S.last_graph_cycles = 0;
    S.control.reserve(1);
    if (!S.host_control)
      LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&S.host_control), 2 * sizeof(IdrControl)));
    for (auto& e : S.cycle_done)
      if (!e) LSMPS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (rr > target2 && max_iterations > 0) {
      IdrControl* ctl = S.control.data();
      idr_control_init<<<1, 1, 0, ctx.stream>>>(ctl, rr, target2, factor2, max_iterations, s);
      LSMPS_LAUNCH_CHECK();
      auto launch = [&](int c) {
        const int slot = c & 1;
        idr_control_begin<<<1, 1, 0, ctx.stream>>>(ctl);
        cycle_body(&ctl->halt, slot, timing && c % kSpmvTimingStride == 0);
        idr_control_cycle<<<1, 1, 0, ctx.stream>>>(ctl, S.scalars.data());
        replace_body(&ctl->skip_replace);
        idr_control_after_replace<<<1, 1, 0, ctx.stream>>>(ctl, S.scalars.data());
        reinit_body(&ctl->skip_reinit);
        LSMPS_LAUNCH_CHECK();
        LSMPS_HIP(hipMemcpyAsync(S.host_control + slot, ctl, sizeof(IdrControl),
                                 hipMemcpyDeviceToHost, ctx.stream));
        LSMPS_HIP(hipEventRecord(S.cycle_done[slot], ctx.stream));
      };
      launch(0);
      launch(1);
      IdrControl last{};
      for (int c = 0;; ++c) {
        const int slot = c & 1;
        LSMPS_HIP(hipEventSynchronize(S.cycle_done[slot]));
        last = S.host_control[slot];
        // cycle c ran: a cycle that starts halted is never waited for (the
        // loop stops at the record that first shows the halt)
        if (timing && c % kSpmvTimingStride == 0) add_timing(slot);
        if (last.halt) break;
        launch(c + 2);
      }
      it = last.iterations;
      replacements = last.replacements;
      rr = last.rr;
    }
  }
  *iterations_out = it;
  *replacements_out = replacements;
  *rr_out = rr;
  return rr <= target2 ? 1 : 0;
}

// Batched IDR(s): every segment (independent system) runs its own recurrences
// and convergence test through shared launches; converged segments freeze.
// Per-segment dot products come from seg_reduce_kernel (the SpMVs run without
// fused dots); reliable updates recompute the true residual of all rows at once.
template <typename T, typename V>
void idr_solve_segmented(Context& ctx, PpeSolver& S, size_t n, const SellMatrix& M,
                         const V* vals, const double* dinv, double* x, double* r, T* work,
                         size_t ld, int s, double tolerance, int max_iterations,
                         const lsmps_csr& A64, const double* b64, double replace_factor,
                         lsmps_solve_report& rep) {
  const SolverSegments& g = S.segments;
  const int nseg = g.count;
  T* v = work;
  T* t = work + ld;
  T* p0 = work + 2 * ld;
  T* G = work + 3 * ld;
  T* U = work + (3 + s) * ld;
  const bool schwarz = S.schwarz.blocks > 0 && S.schwarz.n == n;
  // per-tank coarse space (block-diagonal Ac): applied like the single
  // system's, M^-1 w = Schwarz/Jacobi(w) + P Ac^-1 P^T w
  const bool coarse = schwarz && S.coarse.ready && S.coarse.n == n && S.coarse.tanks == nseg;
  const V* dia = sizeof(V) == sizeof(double) ? reinterpret_cast<const V*>(M.dia64.data())
                                             : reinterpret_cast<const V*>(M.dia32.data());
  // the same fused kernels as one system (b + next a1; a2 with the Schwarz
  // completion and the coarse term), each row reading its tank's scalars
  IdrVec<T> iv{n,
               ld,
               s,
               r,
               x,
               v,
               t,
               G,
               U,
               p0,
               S.scalars.data(),
               S.part_b.data(),
               schwarz ? S.schwarz.pdiag.data() : dinv,
               g.row_seg,
               schwarz ? S.schwarz.cover_start.data() : nullptr,
               schwarz ? S.schwarz.ybuf.data() : nullptr,
               schwarz ? S.schwarz.ybuf_stride : 0,
               coarse ? S.coarse.agg_of_row.data() : nullptr,
               coarse ? S.coarse.yc.data() : nullptr};
  const unsigned vb = vec_blocks(n);
  const unsigned sb = sell_blocks(M.groups);
  const unsigned rb = static_cast<unsigned>(nseg * kSegBlocks);
  const unsigned fb = static_cast<unsigned>((nseg + kBlock - 1) / kBlock);
  double* part = S.part_a.data();
  S.seg_status.reserve(2 * static_cast<size_t>(nseg));
  std::vector<int32_t> flags(nseg, 0), replace(nseg, 0);
  DeviceArray<int32_t>& dflags = S.seg_flags;
  dflags.reserve(nseg);
  auto read_status = [&](std::vector<double>& rr, std::vector<int>& brk) {
    seg_status_kernel<<<fb, kBlock, 0, ctx.stream>>>(S.scalars.data(), nseg, S.seg_status.data());
    LSMPS_HIP(hipMemcpyAsync(S.seg_host, S.seg_status.data(), 2 * nseg * sizeof(double),
                             hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    for (int k = 0; k < nseg; ++k) {
      rr[k] = S.seg_host[2 * k];
      brk[k] = S.seg_host[2 * k + 1] != 0.0 ? 1 : 0;
    }
  };
  auto push_flags = [&]() {
    LSMPS_HIP(hipMemcpyAsync(dflags.data(), flags.data(), nseg * sizeof(int32_t),
                             hipMemcpyHostToDevice, ctx.stream));
  };
  auto segment_norms = [&]() {  // sc[k].rr = |r_k|^2
    seg_reduce_kernel<T, 3><<<rb, kBlock, 0, ctx.stream>>>(g.start, s, p0, nullptr, r, part);
    scalar_norm<<<nseg, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), part, kSegBlocks);
    LSMPS_LAUNCH_CHECK();
  };
  // SpMV timing as in the single-system solve: events around each launch of
  // every kSpmvTimingStride-th cycle (slots 2 k, 2 k + 1), read after the
  // cycle's status copy has waited
  const bool timing = spmv_timing_ready(ctx);
  const double spmv_bytes = spmv_algorithmic_bytes(M, n, sizeof(V), sizeof(T));
  int timed = 0;        // SpMVs recorded in the current cycle
  bool cycle_timed = timing;  // the first cycle is timed
  auto spmv = [&](const T* in, T* out) {
    SellArgs<T, V> a{n, M.chunks, M.groups, M.dia_runs.data(), M.dia_plan.data(),
                     M.dia_start.data(),
                     dia, M.chunk_start.data(), M.chunk_width.data(), M.col_start.data(),
                     M.nwin.data(), M.windows.data(), M.cols.data(), vals, in, out, p0, s,
                     nullptr, nullptr};
    const bool mark = cycle_timed && 2 * timed + 1 < Context::kTimingSet;
    if (mark) LSMPS_HIP(hipEventRecord(ctx.timing_events[2 * timed], ctx.stream));
    sell_spmv_kernel<T, V, 0><<<sb, kSellBlock, 0, ctx.stream>>>(a);
    if (mark) {
      LSMPS_HIP(hipEventRecord(ctx.timing_events[2 * timed + 1], ctx.stream));
      ++timed;
    }
  };
  auto add_timing = [&]() {  // after a host wait that followed the cycle's SpMVs
    for (int k = 0; k < timed; ++k) {
      float ms = 0.0f;
      LSMPS_HIP(hipEventElapsedTime(&ms, ctx.timing_events[2 * k], ctx.timing_events[2 * k + 1]));
      ctx.spmv_ms += ms;
      ctx.spmv_bytes += spmv_bytes;
      ++ctx.spmv_launches;
    }
    timed = 0;
  };
  // |b_k|^2 per segment (r = b on entry, x = 0)
  std::vector<double> bnorm2(nseg), rr(nseg), rr_rep(nseg), target2(nseg);
  std::vector<int> brk(nseg), done(nseg, 0);
  fill_kernel<T><<<blocks_for(2 * s * ld), kBlock, 0, ctx.stream>>>(2 * s * ld, G, T(0));
  copy_kernel<double, T><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, r, p0);
  idr_init_scalars<<<nseg, kBlock, 0, ctx.stream>>>(S.scalars.data(), s);
  segment_norms();
  read_status(bnorm2, brk);
  for (int k = 0; k < nseg; ++k) {
    target2[k] = tolerance * tolerance * bnorm2[k];
    rr[k] = rr_rep[k] = bnorm2[k];
    done[k] = !(bnorm2[k] > 0.0);
    flags[k] = done[k];
  }
  push_flags();
  seg_flags_kernel<<<fb, kBlock, 0, ctx.stream>>>(S.scalars.data(), nseg, dflags.data());
  const double factor2 = replace_factor * replace_factor;
  int it = 0, replacements = 0;
  auto active = [&]() {
    for (int k = 0; k < nseg; ++k)
      if (!done[k]) return true;
    return false;
  };
  for (int cycle = 0; active() && it < max_iterations; ++cycle) {
    cycle_timed = timing && cycle % kSpmvTimingStride == 0;
    seg_reduce_kernel<T, 4><<<rb, kBlock, 0, ctx.stream>>>(g.start, s, p0, nullptr, r, part);
    idr_scalar_f<<<nseg, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), part, kSegBlocks, s);
    // step 0's first half; later steps get it fused into the previous step's b
    if (schwarz) idr_a1_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv, 0);
    for (int k = 0; k < s; ++k) {
      if (schwarz) {
        if (coarse) {  // block products and restriction in one launch, then yc
          products_and_restrict<T>(ctx, S.schwarz, S.coarse, v, v, true);
          coarse_product(ctx, S.coarse);
        } else {
          schwarz_products<T>(ctx, S.schwarz, v);
        }
        idr_a2g_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
      } else {
        idr_a_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
      }
      spmv(U + k * ld, G + k * ld);
      seg_reduce_kernel<T, 1><<<rb, kBlock, 0, ctx.stream>>>(g.start, s, p0, G + k * ld, r, part);
      idr_scalar_after<<<nseg, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), part,
                                                                kSegBlocks, s, k);
      if (schwarz && k < s - 1) {
        idr_ba_kernel<T, true, true><<<vb, kBlock, 0, ctx.stream>>>(iv, k);
      } else {
        idr_b_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv, k, k == s - 1 ? 1 : 0);
      }
      ++it;
    }
    if (schwarz) {
      if (coarse) {  // products of v and restriction of r in one launch
        products_and_restrict<T>(ctx, S.schwarz, S.coarse, v, r, false);
        coarse_product(ctx, S.coarse);
        schwarz_complete_prolong_add<T>(ctx, S.schwarz, S.coarse, v);
      } else {
        schwarz_apply<T>(ctx, S.schwarz, v);
      }
    }
    spmv(v, t);
    seg_reduce_kernel<T, 2><<<rb, kBlock, 0, ctx.stream>>>(g.start, s, p0, t, r, part);
    idr_scalar_omega<<<nseg, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), part, part,
                                                              kSegBlocks, kSegBlocks, 2);
    idr_c_kernel<T, true><<<vb, kBlock, 0, ctx.stream>>>(iv);
    segment_norms();
    ++it;
    read_status(rr, brk);
    add_timing();
    // reliable updates of the segments whose residual met the target or fell
    // by the replacement factor (the other segments' rows are not touched)
    bool need_true = false;
    for (int k = 0; k < nseg; ++k) {
      replace[k] = 0;
      if (done[k]) continue;
      if (!std::isfinite(rr[k])) raise(LSMPS_ERR_NUMERICAL, "PPE solve produced a non-finite residual");
      if (rr[k] <= target2[k] || rr[k] <= factor2 * rr_rep[k]) {
        replace[k] = 1;
        need_true = true;
      }
    }
    if (need_true) {
      LSMPS_HIP(hipMemcpyAsync(dflags.data(), replace.data(), nseg * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx.stream));
      const unsigned nb = spmv_blocks(n);
      residual_kernel<<<nb, kBlock, 0, ctx.stream>>>(n, A64.row_offsets, A64.col_indices,
                                                     A64.values, x, b64, r, nullptr,
                                                     S.part_b.data(), nullptr, g.row_seg,
                                                     dflags.data());
      segment_norms();
      read_status(rr, brk);
      for (int k = 0; k < nseg; ++k)
        if (replace[k]) rr_rep[k] = rr[k];
      ++replacements;
    }
    bool restart = false;
    for (int k = 0; k < nseg; ++k) {
      flags[k] = 0;
      if (done[k]) continue;
      if (rr[k] <= target2[k]) {
        done[k] = 1;
        flags[k] = 1;
      } else if (brk[k]) {
        flags[k] = 2;
        restart = true;
      }
    }
    push_flags();
    seg_flags_kernel<<<fb, kBlock, 0, ctx.stream>>>(S.scalars.data(), nseg, dflags.data());
    if (restart) {
      seg_restart_kernel<T><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, g.row_seg, dflags.data(),
                                                                     ld, s, r, p0, G, U);
      idr_init_scalars<<<nseg, kBlock, 0, ctx.stream>>>(S.scalars.data(), s, dflags.data());
    }
    LSMPS_LAUNCH_CHECK();
  }
  double total = 0.0, bn = 0.0, worst = 0.0;
  bool all = true;
  for (int k = 0; k < nseg; ++k) {
    total += rr[k];
    bn += bnorm2[k];
    if (bnorm2[k] > 0.0) worst = std::max(worst, std::sqrt(rr[k] / bnorm2[k]));
    all = all && done[k];
  }
  rep.iterations = it;
  rep.refinements = replacements;
  rep.rhs_norm = std::sqrt(bn);
  rep.relative_residual = worst;  // the largest per-segment relative residual
  if (coarse && coarse_usable(ctx, S.coarse)) rep.coarse_unknowns = S.coarse.count;
  rep.converged = all ? 1 : 0;
  (void)total;
}

void pcg_solve(Context& ctx, PpeSolver& S, const lsmps_csr& A, const double* b, double* x,
               double target, int max_iterations, int check, lsmps_solve_report& rep) {
  const size_t n = A.rows;
  const size_t ld = (n + 63) / 64 * 64;
  S.vec64.reserve(3 * ld);
  double* r = S.vec64.data();
  double* p = r + ld;
  double* q = p + ld;
  const unsigned vb = vec_blocks(n), sb = spmv_blocks(n);
  const double* dinv = S.dinv64.data();
  double rr = true_residual(ctx, S, A, x, b, r, nullptr);
  idr_init_scalars<<<1, kBlock, 0, ctx.stream>>>(S.scalars.data(), 1);
  rz_kernel<<<vb, kBlock, 0, ctx.stream>>>(n, r, dinv, S.part_b.data());
  cg_scalar_rz<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(), static_cast<int>(vb));
  cg_direction_kernel<<<vb, kBlock, 0, ctx.stream>>>(n, p, r, dinv, S.scalars.data(), 1);
  LSMPS_LAUNCH_CHECK();
  int it = 0;
  const double target2 = target * target;
  while (rr > target2 && it < max_iterations) {
    SpmvArgs<double, double> sa{n, A.row_offsets, A.col_indices, A.values, p, q, nullptr, 0,
                                nullptr, S.part_a.data()};
    spmv_kernel<double, double, 3><<<sb, kBlock, 0, ctx.stream>>>(sa);
    cg_scalar_alpha<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_a.data(), static_cast<int>(sb));
    cg_update_kernel<<<vb, kBlock, 0, ctx.stream>>>(n, x, r, p, q, dinv, S.scalars.data(),
                                                    S.part_b.data());
    cg_scalar_rz<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(), static_cast<int>(vb));
    cg_direction_kernel<<<vb, kBlock, 0, ctx.stream>>>(n, p, r, dinv, S.scalars.data(), 0);
    LSMPS_LAUNCH_CHECK();
    ++it;
    if (it % check == 0 || it >= max_iterations) {
      rr = read_rr(ctx, S);
      if (!std::isfinite(rr)) break;
      if (rr <= target2) rr = true_residual(ctx, S, A, x, b, r, nullptr);
    }
  }
  rep.iterations = it;
  rep.relative_residual = std::sqrt(rr);
}

}  // namespace

// y = A x through the Krylov layout and SpMV kernel, for direct checks of the
// layout against a CSR product: FP32 values (fp32_values) or FP64.
void sell_spmv(Context& ctx, const lsmps_csr& A, const double* x, double* y, bool fp32_values) {
  require(A.rows == A.cols, "SpMV check: the matrix must be square");
  require(x != nullptr && y != nullptr, "SpMV check: null vector");
  const size_t n = A.rows;
  if (n == 0) return;
  SellMatrix M;
  const unsigned sb = [&] {
    if (fp32_values) build_sell<float>(ctx, M, A, nullptr); else build_sell<double>(ctx, M, A, nullptr);
    return sell_blocks((n + kGroupRows - 1) / kGroupRows);
  }();
  if (fp32_values) {
    SellArgs<double, float> a{n, M.chunks, M.groups, M.dia_runs.data(), M.dia_plan.data(),
                              M.dia_start.data(), M.dia32.data(), M.chunk_start.data(),
                              M.chunk_width.data(), M.col_start.data(), M.nwin.data(),
                              M.windows.data(), M.cols.data(), M.vals32.data(), x, y, nullptr, 0,
                              nullptr, nullptr};
    sell_spmv_kernel<double, float, 0><<<sb, kSellBlock, 0, ctx.stream>>>(a);
  } else {
    SellArgs<double, double> a{n, M.chunks, M.groups, M.dia_runs.data(), M.dia_plan.data(),
                               M.dia_start.data(), M.dia64.data(), M.chunk_start.data(),
                               M.chunk_width.data(), M.col_start.data(), M.nwin.data(),
                               M.windows.data(), M.cols.data(), M.vals64.data(), x, y, nullptr, 0,
                               nullptr, nullptr};
    sell_spmv_kernel<double, double, 0><<<sb, kSellBlock, 0, ctx.stream>>>(a);
  }
  LSMPS_LAUNCH_CHECK();
  ctx.sync();  // M's buffers are freed on return
}

void ppe_solve(Context& ctx, PpeSolver& S, const lsmps_csr& A, const double* b, double* x,
               const lsmps_solver_options& opt, lsmps_solve_report* report) {
  lsmps_solve_report rep{};
  rep.method = opt.method;
  const size_t n = A.rows;
  require(A.rows == A.cols, "PPE solve requires a square CSR matrix, rhs, and solution buffer");
  if (n == 0) {
    rep.converged = 1;
    if (report) *report = rep;
    return;
  }
  require(A.row_offsets && A.col_indices && A.values && b && x,
          "PPE solve requires a square CSR matrix, rhs, and solution buffer");
  require(opt.tolerance > 0.0, "PPE solve tolerance must be positive");
  require(n <= static_cast<size_t>(INT32_MAX), "PPE solve row count exceeds 32-bit indices");
  if (!S.host) LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&S.host), 8 * sizeof(double)));
  S.scalars.reserve(1);
  S.part_a.reserve(static_cast<size_t>(kMaxReduceBlocks) * kPartialStride);
  S.part_b.reserve(static_cast<size_t>(kMaxReduceBlocks) * kPartialStride);
  S.dinv64.reserve(n);
  S.diag64.reserve(n);
  S.offabs.reserve(n);
  row_stats_kernel<<<blocks_for(n * kRowLanes), kBlock, 0, ctx.stream>>>(
      n, A.row_offsets, A.col_indices, A.values, S.diag64.data(), S.dinv64.data(),
      S.offabs.data());
  LSMPS_LAUNCH_CHECK();
  // ||b||
  dot_self_kernel<<<vec_blocks(n), kBlock, 0, ctx.stream>>>(n, b, S.part_b.data());
  scalar_norm<<<1, kReduceThreads, 0, ctx.stream>>>(S.scalars.data(), S.part_b.data(),
                                            static_cast<int>(vec_blocks(n)));
  LSMPS_LAUNCH_CHECK();
  const double bnorm = std::sqrt(read_rr(ctx, S));
  rep.rhs_norm = bnorm;
  fill_kernel<double><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, x, 0.0);
  LSMPS_LAUNCH_CHECK();
  if (bnorm == 0.0) {
    rep.converged = 1;
    if (report) *report = rep;
    ctx.finish();
    return;
  }
  const double target = opt.tolerance * bnorm;
  const int max_it = opt.max_iterations > 0 ? opt.max_iterations : 20000;
  const size_t ld = (n + 63) / 64 * 64;
  int shadow_dim = opt.method == LSMPS_SOLVER_BICGSTAB ? 1 : opt.shadow_dimension;
  // s = 0 picks 14 when the coarse space is requested and 16 otherwise: with
  // the smooth modes removed by the coarse term, s = 12..16 cost about the same
  // (fewer iterations against more vector traffic per step).
  // (batched tanks: the caller's aggregation is per tank, one block each)
  const bool coarse_requested =
      opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE && S.coarse.n == n &&
      (S.segments.count == 0 ? S.coarse.tanks == 0 && S.coarse.count > 1
                             : S.coarse.tanks == S.segments.count && S.coarse.count > 0);
  if (shadow_dim <= 0) shadow_dim = coarse_requested ? 14 : 16;
  if (shadow_dim > kMaxShadow) shadow_dim = kMaxShadow;
  if (static_cast<size_t>(shadow_dim) > n) shadow_dim = static_cast<int>(n);

  if (opt.method == LSMPS_SOLVER_PCG) {
    pcg_solve(ctx, S, A, b, x, target, max_it, opt.check_interval > 0 ? opt.check_interval : 16,
              rep);
    rep.relative_residual /= bnorm;
    rep.converged = rep.relative_residual <= opt.tolerance ? 1 : 0;
  } else {
    // r, then the basis: v, t, P_0, G[s], U[s]
    S.vec64.reserve((4 + 2 * static_cast<size_t>(shadow_dim)) * ld);
    double* r = S.vec64.data();
    copy_kernel<double, double><<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, b, r);
    LSMPS_LAUNCH_CHECK();
    int it = 0, replacements = 0;
    double rr = 0.0;
    S.schwarz.blocks = 0;
    S.coarse.ready = false;
    if (opt.preconditioner == LSMPS_PRECOND_SCHWARZ ||
        opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE) {
      if (S.schwarz_transpose_source && S.schwarz_transpose_source->n == n) {
        schwarz_transpose_of(ctx, S.schwarz, *S.schwarz_transpose_source, ctx.schwarz_half());
      } else {
        schwarz_setup(ctx, S.schwarz, A, S.diag64.data(), S.dinv64.data(), S.offabs.data(),
                      ctx.precision == LSMPS_PRECISION_MIXED, ctx.schwarz_half());
      }
      int64_t blocks = 0, rows = 0;
      schwarz_stats(ctx, S.schwarz, &blocks, &rows);
      rep.schwarz_blocks = static_cast<int32_t>(blocks);
      rep.schwarz_rows = static_cast<int32_t>(std::min<int64_t>(rows, INT32_MAX));
    }
    const bool mixed = ctx.precision == LSMPS_PRECISION_MIXED;
    // mixed: FP32 matrix values in the Krylov SpMVs; FP64 basis, iterate,
    // residual, dots and reliable updates against the FP64 matrix.  (An FP32
    // basis stagnates: r -= beta G_k with a rounded G_k leaves an error of
    // 1e-7 |beta G_k|, far above the target when the recurrences grow.)
    const double factor = mixed ? (opt.inner_tolerance > 0.0 ? opt.inner_tolerance : 1.0e-3) : 0.0;
    if (mixed) {
      build_sell<float>(ctx, S.sell, A, nullptr);
    } else {
      build_sell<double>(ctx, S.sell, A, nullptr);
    }
    // The coarse space (it needs the caller's aggregation of these rows: one
    // system, or one block per tank) is set up on the library stream after the Schwarz blocks and the
    // SpMV layout: every buffer it touches is reserved first, on this thread,
    // and nothing is enqueued on another stream.
    if (coarse_requested && S.schwarz.n == n) {
      coarse_reserve(S.coarse);
      const AllocationFreeze frozen;  // no device allocation until the loop starts
      coarse_setup(ctx, S.coarse, A, S.diag64.data(), S.offabs.data(), ctx.stream);
      if (S.coarse.ready) coarse_set_weights(ctx, S.coarse, S.schwarz.pdiag.data());
    }
    if (S.segments.count > 0) {
      // independent systems stacked in one matrix: per-segment recurrences
      if (!S.seg_host || S.seg_host_count < static_cast<size_t>(S.segments.count)) {
        if (S.seg_host) LSMPS_HIP(hipHostFree(S.seg_host));
        S.seg_host = nullptr;
        LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&S.seg_host),
                                2 * S.segments.count * sizeof(double)));
        S.seg_host_count = static_cast<size_t>(S.segments.count);
      }
      S.scalars.reserve(static_cast<size_t>(S.segments.count));
      S.part_a.reserve(static_cast<size_t>(S.segments.count) * kSegBlocks * kPartialStride);
      if (mixed) {
        idr_solve_segmented<double, float>(ctx, S, n, S.sell, S.sell.vals32.data(),
                                           S.dinv64.data(), x, r, r + ld, ld, shadow_dim,
                                           opt.tolerance, max_it, A, b, factor, rep);
      } else {
        idr_solve_segmented<double, double>(ctx, S, n, S.sell, S.sell.vals64.data(),
                                            S.dinv64.data(), x, r, r + ld, ld, shadow_dim,
                                            opt.tolerance, max_it, A, b, factor, rep);
      }
    } else {
      S.graph_requested = ctx.graph();
      if (mixed) {
        idr_solve<double, float>(ctx, S, n, S.sell, S.sell.vals32.data(), S.dinv64.data(), x,
                                 r, r + ld, ld, shadow_dim, bnorm * bnorm, target, max_it, A, b,
                                 factor, &it, &replacements, &rr);
      } else {
        idr_solve<double, double>(ctx, S, n, S.sell, S.sell.vals64.data(), S.dinv64.data(), x,
                                  r, r + ld, ld, shadow_dim, bnorm * bnorm, target, max_it, A,
                                  b, factor, &it, &replacements, &rr);
      }
      rep.iterations = it;
      rep.refinements = replacements;
      rep.graph_cycles = S.last_graph_cycles;
      if (S.coarse.ready && coarse_usable(ctx, S.coarse)) rep.coarse_unknowns = S.coarse.count;
      rep.relative_residual = std::sqrt(rr) / bnorm;
      rep.converged = rep.relative_residual <= opt.tolerance ? 1 : 0;
    }
  }
  if (opt.clamp_negative) {
    // clamp_pressure_kernel (ppe_matrix.cu:373-379)
    clamp_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, x);
    LSMPS_LAUNCH_CHECK();
  }
  if (report) *report = rep;
  ctx.finish();
  if (!rep.converged)
    raise(LSMPS_ERR_NOT_CONVERGED, "PPE solver did not reach the tolerance (relative residual " +
                                       std::to_string(rep.relative_residual) + ")");
}

}  // namespace lsmps
