// Restricted additive Schwarz correction for the PPE's wrong-sign rows.
//
// The reference's PPE (src/ppe/ppe_matrix.cu:243-370) has a few hundred rows
// with a positive diagonal: NearSurface particles in wall corners whose
// truncated, wall-dominated neighbourhoods give the LSMPS Laplacian row the
// wrong sign.  They make the system indefinite and dominate Krylov
// convergence (Jacobi-IDR(8) on the 250,000-particle dam-break: ~4000 SpMVs).
// This preconditioner keeps Jacobi on every other row and, around each
// wrong-sign row, solves the local problem exactly: the block of its 2-hop
// neighbourhood (1-hop if that exceeds kMaxBlock rows) is factored by LU with
// partial pivoting into an explicit inverse; rows covered by several blocks take
// the average of their blocks' values, summed in block order (restricted
// additive Schwarz with averaging).  The
// matrix itself is not modified, so the solution is that of the reference system.
//
// Setup, per solve: one workgroup per seed row builds the member set in an LDS
// hash table, sorts it (deterministic membership and order), assembles the dense
// block, factors it and forms the inverse (column-major).  The application
// multiplies every block's inverse with its input slice (kApplyParts workgroups
// per block) into partial products, and each covered row then takes the mean of
// its blocks' values, summed in block order, in place.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "schwarz.hpp"

namespace lsmps {
namespace {

constexpr int kTable = 4096;  // LDS hash table entries (power of two)

// Per row: 1 when eligible to join a block (has off-diagonal couplings),
// 2 when it is also a seed (positive diagonal).  Dirichlet rows (identity rows
// of Surface / Splash particles) have no off-diagonal values and stay Jacobi.
__global__ void classify_rows_kernel(size_t n, const double* diag, const double* offabs,
                                     int32_t* eligible, int32_t* seed) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double off = offabs[i];
  eligible[i] = off > 0.0 ? 1 : 0;
  seed[i] = (off > 0.0 && diag[i] > 0.0) ? 1 : 0;
}

__global__ void compact_seeds_kernel(size_t n, const int32_t* flag, const int32_t* pos,
                                     int32_t* seeds) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) seeds[pos[i]] = static_cast<int32_t>(i);
}

// Linear probing, bounded by the table size: the callers stop inserting once
// more than kMaxBlock keys are present, far below kTable, so a probe sequence
// always finds its key or a free slot; the bound only guarantees termination.
__device__ void table_insert(int32_t* table, int32_t key, int32_t* count) {
  uint32_t h = (static_cast<uint32_t>(key) * 2654435761u) & (kTable - 1);
  for (int probe = 0; probe < kTable; ++probe) {
    const int32_t prev = atomicCAS(&table[h], -1, key);
    if (prev == -1) {
      atomicAdd(count, 1);
      return;
    }
    if (prev == key) return;
    h = (h + 1) & (kTable - 1);
  }
}

// Member set of block b (1 workgroup of kMaxBlock threads per block).
__global__ void __launch_bounds__(kMaxBlock) members_kernel(
    const int32_t* seeds, const int32_t* rows, const int32_t* cols, const int32_t* eligible,
    int32_t* members, int32_t* sizes) {
  __shared__ int32_t table[kTable];
  __shared__ int32_t list[kMaxBlock];
  __shared__ int32_t count;
  __shared__ int32_t filled;
  const int t = threadIdx.x;
  const int32_t seed = seeds[blockIdx.x];
  for (int e = t; e < kTable; e += blockDim.x) table[e] = -1;
  if (t == 0) count = 0;
  __syncthreads();
  // 1-hop: the seed row's columns (the seed itself is its diagonal entry)
  const int32_t sb = rows[seed], se = rows[seed + 1];
  for (int32_t c = sb + t; c < se; c += blockDim.x) {
    const int32_t j = cols[c];
    if (eligible[j]) table_insert(table, j, &count);
  }
  __syncthreads();
  // 2-hop: columns of every eligible 1-hop member.  `count` is read between two
  // barriers so that every thread sees the same value (no insert of the next
  // row can overtake a slower wave's read) and the loop exits uniformly.
  bool two_hop = true;
  for (int32_t m = sb; m < se; ++m) {
    const int32_t j = cols[m];
    if (!eligible[j]) continue;
    for (int32_t c = rows[j] + t; c < rows[j + 1]; c += blockDim.x) {
      const int32_t k = cols[c];
      if (eligible[k]) table_insert(table, k, &count);
    }
    __syncthreads();
    const int32_t now = count;
    __syncthreads();
    if (now > kMaxBlock) {
      two_hop = false;
      break;
    }
  }
  if (!two_hop) {  // fall back to the 1-hop set
    for (int e = t; e < kTable; e += blockDim.x) table[e] = -1;
    if (t == 0) count = 0;
    __syncthreads();
    for (int32_t c = sb + t; c < se; c += blockDim.x) {
      const int32_t j = cols[c];
      if (eligible[j]) table_insert(table, j, &count);
    }
    __syncthreads();
  }
  // compact and sort ascending (bitonic sort over kMaxBlock slots)
  if (t == 0) filled = 0;
  list[t] = INT32_MAX;
  __syncthreads();
  for (int e = t; e < kTable; e += blockDim.x) {
    const int32_t key = table[e];
    if (key >= 0) {
      const int32_t slot = atomicAdd(&filled, 1);
      if (slot < kMaxBlock) list[slot] = key;  // a >kMaxBlock 1-hop row is truncated
    }
  }
  __syncthreads();
  if (t == 0 && filled > kMaxBlock) filled = kMaxBlock;
  for (int size = 2; size <= kMaxBlock; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int partner = t ^ stride;
      if (partner > t) {
        const bool up = (t & size) == 0;
        const int32_t a = list[t], b = list[partner];
        if ((a > b) == up) {
          list[t] = b;
          list[partner] = a;
        }
      }
      __syncthreads();
    }
  }
  members[static_cast<size_t>(blockIdx.x) * kMaxBlock + t] = list[t];
  if (t == 0) sizes[blockIdx.x] = filled;
}

__device__ int find_member(const int32_t* list, int n, int32_t key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int32_t v = list[mid];
    if (v == key) return mid;
    if (v < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// Dense block, LU with partial pivoting, explicit inverse (column-major).
template <typename V>
__global__ void __launch_bounds__(kMaxBlock) factor_kernel(
    const int32_t* rows, const int32_t* cols, const double* vals, const int32_t* members,
    const int32_t* sizes, double* scratch, V* inverse) {
  __shared__ int32_t list[kMaxBlock];
  __shared__ int32_t piv[kMaxBlock];
  __shared__ double red_v[kMaxBlock];
  __shared__ int32_t red_i[kMaxBlock];
  const int t = threadIdx.x;
  const int n = sizes[blockIdx.x];
  constexpr int ld = kMaxBlock;
  double* a = scratch + static_cast<size_t>(blockIdx.x) * ld * ld;  // row-major
  V* inv = inverse + static_cast<size_t>(blockIdx.x) * ld * ld;     // column-major
  list[t] = members[static_cast<size_t>(blockIdx.x) * kMaxBlock + t];
  for (int e = t; e < ld * ld; e += blockDim.x) a[e] = 0.0;
  __syncthreads();
  if (t < n) {
    const int32_t row = list[t];
    for (int32_t c = rows[row]; c < rows[row + 1]; ++c) {
      const int p = find_member(list, n, cols[c]);
      if (p >= 0) a[t * ld + p] += vals[c];
    }
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // pivot: largest |a[r][k]|, r >= k (ties -> smallest r)
    const double v = (t >= k && t < n) ? fabs(a[t * ld + k]) : -1.0;
    red_v[t] = v;
    red_i[t] = t;
    __syncthreads();
    for (int w = kMaxBlock / 2; w > 0; w >>= 1) {
      if (t < w) {
        const double o = red_v[t + w];
        const int32_t oi = red_i[t + w];
        if (o > red_v[t] || (o == red_v[t] && oi < red_i[t])) {
          red_v[t] = o;
          red_i[t] = oi;
        }
      }
      __syncthreads();
    }
    const int p = red_i[0];
    if (t == 0) piv[k] = p;
    if (p != k) {
      for (int j = t; j < n; j += blockDim.x) {
        const double x = a[k * ld + j];
        a[k * ld + j] = a[p * ld + j];
        a[p * ld + j] = x;
      }
    }
    __syncthreads();
    double akk = a[k * ld + k];
    if (akk == 0.0) akk = 1.0e-300;  // singular block: keep finite, Krylov absorbs it
    for (int r = k + 1 + t; r < n; r += blockDim.x) a[r * ld + k] /= akk;
    __syncthreads();
    const int m = n - k - 1;
    for (int e = t; e < m * m; e += blockDim.x) {
      const int r = k + 1 + e / m, j = k + 1 + e % m;
      a[r * ld + j] -= a[r * ld + k] * a[k * ld + j];
    }
    __syncthreads();
  }
  // Inverse column t (thread t): solve P A x = e_t, i.e. L U x = P e_t.  The
  // right-hand sides are stored X[i * ld + t] so that the substitutions, which
  // advance i in lockstep over all threads, read and write coalesced rows while
  // the factor entries a[i][j] are broadcast.
  if (t < n) {
    double* X = scratch + static_cast<size_t>(gridDim.x) * ld * ld +
                static_cast<size_t>(blockIdx.x) * ld * ld;
    for (int i = 0; i < n; ++i) X[i * ld + t] = (i == t) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) {
        const double x = X[k * ld + t];
        X[k * ld + t] = X[p * ld + t];
        X[p * ld + t] = x;
      }
    }
    for (int i = 0; i < n; ++i) {  // unit lower
      double s = X[i * ld + t];
      for (int j = 0; j < i; ++j) s -= a[i * ld + j] * X[j * ld + t];
      X[i * ld + t] = s;
    }
    for (int i = n - 1; i >= 0; --i) {  // upper
      double s = X[i * ld + t];
      for (int j = i + 1; j < n; ++j) s -= a[i * ld + j] * X[j * ld + t];
      double d = a[i * ld + i];
      if (d == 0.0) d = 1.0e-300;
      X[i * ld + t] = s / d;
    }
    // inverse(i, t) at t * ld + i (column-major): the application reads column j
    // as inv[j * ld + row], coalesced over its threads.
    for (int i = 0; i < n; ++i) inv[static_cast<size_t>(t) * ld + i] = static_cast<V>(X[i * ld + t]);
  }
}

// Mixed mode: the block inverse by Gauss-Jordan elimination with partial
// pivoting in FP32, the whole block held in registers.  1024 threads; thread
// (ty, tx) = (t % 32, t / 32) owns rows ty + 32 a and columns tx + 32 b
// (a, b < 8), 64 entries.  Step k: the half-wave owning column k picks the
// pivot row p (largest |a_rk|, r >= k, ties to the smaller row) and publishes
// column k, the owners of rows k and p publish those rows (LDS, double
// buffered by k parity, so two barriers per step), then every thread updates
// its 64 entries: row k becomes row p / pivot, every other row r loses
// f_r times it, and column k takes -f_r / pivot (in-place inversion).  The
// row interchanges become column interchanges of the result, applied when the
// inverse is written out (column-major, like the FP64 path).  The block is
// assembled into the output buffer first; nothing else is allocated.
constexpr int kGjThreads = 1024;

__global__ void __launch_bounds__(kGjThreads) gj_inverse_kernel(
    const int32_t* rows, const int32_t* cols, const double* vals, const int32_t* members,
    const int32_t* sizes, float* inverse) {
  __shared__ int32_t list[kMaxBlock];
  __shared__ int32_t piv[kMaxBlock];
  __shared__ int32_t pos[kMaxBlock];
  __shared__ float row_k[2][kMaxBlock];
  __shared__ float row_p[2][kMaxBlock];
  __shared__ float col_k[2][kMaxBlock];
  __shared__ int32_t pivot_row[2];
  constexpr int ld = kMaxBlock;
  const int t = threadIdx.x;
  const int n = sizes[blockIdx.x];
  float* buf = inverse + static_cast<size_t>(blockIdx.x) * ld * ld;
  if (t < kMaxBlock) list[t] = members[static_cast<size_t>(blockIdx.x) * kMaxBlock + t];
  for (int e = t; e < ld * ld; e += kGjThreads) buf[e] = 0.0f;
  __syncthreads();
  {
    const int r = t >> 2, q = t & 3;  // 4 threads per block row
    if (r < n) {
      const int32_t row = list[r];
      for (int32_t c = rows[row] + q; c < rows[row + 1]; c += 4) {
        const int p = find_member(list, n, cols[c]);
        if (p >= 0) buf[static_cast<size_t>(p) * ld + r] = static_cast<float>(vals[c]);
      }
    }
  }
  __syncthreads();
  const int ty = t & 31, tx = t >> 5;
  float a[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) a[i][j] = buf[static_cast<size_t>(tx + 32 * j) * ld + ty + 32 * i];
  for (int k = 0; k < n; ++k) {
    const int s = k & 1;
    const int kb = k >> 5, kt = k & 31;
    if (tx == kt) {
      float best = -1.0f;
      int best_r = INT32_MAX;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) v = (j == kb) ? a[i][j] : v;
        const int r = ty + 32 * i;
        col_k[s][r] = v;
        const float m = fabsf(v);
        if (r >= k && r < n && m > best) {
          best = m;
          best_r = r;
        }
      }
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off, 32);
        const int orow = __shfl_xor(best_r, off, 32);
        if (ob > best || (ob == best && orow < best_r)) {
          best = ob;
          best_r = orow;
        }
      }
      if (ty == 0) {
        pivot_row[s] = best_r < n ? best_r : k;
        piv[k] = best_r < n ? best_r : k;
      }
    }
    if (ty == kt) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float v = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) v = (i == kb) ? a[i][j] : v;
        row_k[s][tx + 32 * j] = v;
      }
    }
    __syncthreads();
    const int p = pivot_row[s];
    if (ty == (p & 31)) {
      const int pb = p >> 5;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float v = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) v = (i == pb) ? a[i][j] : v;
        row_p[s][tx + 32 * j] = v;
      }
    }
    __syncthreads();
    float pivot = row_p[s][k];
    if (pivot == 0.0f) pivot = 1.0e-30f;  // singular block: keep finite, Krylov absorbs it
    const float ip = 1.0f / pivot;
    float u[8], rk[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int col = tx + 32 * j;
      u[j] = col == k ? ip : row_p[s][col] * ip;
      rk[j] = row_k[s][col];
    }
    const float fk = row_k[s][k];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = ty + 32 * i;
      const bool is_k = r == k, is_p = r == p;
      const float f = is_p ? fk : col_k[s][r];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int col = tx + 32 * j;
        const float src = is_p ? rk[j] : a[i][j];
        const float v = col == k ? -f * ip : src - f * u[j];
        a[i][j] = is_k ? u[j] : v;
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    // column order of the inverse: undo the row interchanges in reverse
    for (int c = 0; c < kMaxBlock; ++c) list[c] = c;
    for (int k = n - 1; k >= 0; --k) {
      const int q = piv[k];
      const int32_t x = list[k];
      list[k] = list[q];
      list[q] = x;
    }
    for (int c = 0; c < kMaxBlock; ++c) pos[list[c]] = c;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j)
      buf[static_cast<size_t>(pos[tx + 32 * j]) * ld + ty + 32 * i] = a[i][j];
}

// Number of blocks covering each row.
__global__ void cover_count_kernel(const int32_t* members, const int32_t* sizes, int32_t* cover) {
  const int t = threadIdx.x;
  if (t < sizes[blockIdx.x]) atomicAdd(&cover[members[static_cast<size_t>(blockIdx.x) * kMaxBlock + t]], 1);
}

// Row -> covering block slots (b * kMaxBlock + position), unordered.
__global__ void cover_fill_kernel(const int32_t* members, const int32_t* sizes,
                                  const int32_t* start, int32_t* cursor, int32_t* entries) {
  const int t = threadIdx.x;
  if (t >= sizes[blockIdx.x]) return;
  const int32_t row = members[static_cast<size_t>(blockIdx.x) * kMaxBlock + t];
  const int32_t k = atomicAdd(&cursor[row], 1);
  entries[start[row] + k] = static_cast<int32_t>(blockIdx.x) * kMaxBlock + t;
}

// Ascending slot order per row, so the averaged value is summed in block order
// and is the same on every run.
__global__ void cover_sort_kernel(size_t n, const int32_t* start, int32_t* entries) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t b = start[i], e = start[i + 1];
  for (int32_t p = b + 1; p < e; ++p) {
    const int32_t v = entries[p];
    int32_t q = p - 1;
    while (q >= b && entries[q] > v) {
      entries[q + 1] = entries[q];
      --q;
    }
    entries[q + 1] = v;
  }
}

template <typename T, typename V>
__global__ void __launch_bounds__(kMaxBlock) apply_kernel(const int32_t* members,
                                                          const int32_t* sizes, const V* inverse,
                                                          const float* scale, const T* w,
                                                          const int32_t* entry_of, double* ybuf,
                                                          size_t stride, const int32_t* skip) {
  if (launch_skipped(skip)) return;
  schwarz_block_product<T, V>(blockIdx.x, members, sizes, inverse, scale, w, entry_of, ybuf,
                              stride);
}

// Covered rows take the mean of their blocks' values (each the sum of its
// kApplyParts partials), in ascending block order.
template <typename T>
__global__ void gather_kernel(size_t count, const int32_t* covered, const int32_t* start,
                              const double* ybuf, size_t stride, T* v, const int32_t* skip) {
  if (launch_skipped(skip)) return;
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int32_t i = covered[k];
  v[i] = static_cast<T>(schwarz_row_mean(start, ybuf, stride, i));
}

// entry_of[slot] = e for every cover entry e (one thread per row).
__global__ void entry_of_kernel(size_t n, const int32_t* start, const int32_t* entries,
                                int32_t* entry_of) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t e = start[i]; e < start[i + 1]; ++e) entry_of[entries[e]] = e;
}

// FP16 copies of the FP32 inverses: one wavefront per block column c < n;
// e_c puts the column's largest magnitude m in [2^14, 2^15) (e = 14 - ilogb m),
// so every entry keeps FP16's 11 significant bits relative to m and nothing
// overflows; scale[c] = 2^-e_c undoes it exactly in the application.
__global__ void __launch_bounds__(kBlock) quantize_kernel(const int32_t* sizes, const float* inv32,
                                                          _Float16* inv16, float* scale) {
  constexpr int ld = kMaxBlock;
  constexpr int kWaves = kBlock / 64;
  const size_t b = blockIdx.x;
  const int n = sizes[b];
  const int lane = threadIdx.x & 63;
  for (int c = threadIdx.x >> 6; c < n; c += kWaves) {
    const size_t base = (b * ld + c) * ld;
    float m = 0.0f;
    for (int r = lane; r < n; r += 64) m = fmaxf(m, fabsf(inv32[base + r]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const int e = m > 0.0f && isfinite(m) ? 14 - ilogbf(m) : 0;
    const float up = ldexpf(1.0f, e);
    for (int r = lane; r < n; r += 64)
      inv16[half_index(b, r, c)] = static_cast<_Float16>(inv32[base + r] * up);
    if (lane == 0) scale[b * ld + c] = ldexpf(1.0f, -e);
  }
}

// pdiag and the flags of covered rows.
__global__ void pdiag_kernel(size_t n, const int32_t* start, const double* dinv, double* pdiag,
                             int32_t* flag) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool covered = start[i + 1] > start[i];
  pdiag[i] = covered ? 1.0 : dinv[i];
  flag[i] = covered ? 1 : 0;
}

__global__ void compact_kernel(size_t n, const int32_t* flag, const int32_t* pos, int32_t* out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = static_cast<int32_t>(i);
}

}  // namespace

namespace {
void quantize_blocks(Context& ctx, SchwarzPrec& P) {
  const size_t dense = P.blocks * kMaxBlock * kMaxBlock;
  P.inv16.reserve(dense);
  P.col_scale.reserve(P.blocks * kMaxBlock);
  quantize_kernel<<<static_cast<unsigned>(P.blocks), kBlock, 0, ctx.stream>>>(
      P.sizes.data(), P.inv32.data(), reinterpret_cast<_Float16*>(P.inv16.data()),
      P.col_scale.data());
  LSMPS_LAUNCH_CHECK();
}
}  // namespace

void schwarz_setup(Context& ctx, SchwarzPrec& P, const lsmps_csr& A, const double* diag,
                   const double* dinv, const double* offabs, bool fp32_inverse,
                   bool half_application) {
  const size_t n = A.rows;
  P.n = n;
  P.blocks = 0;
  P.fp32 = fp32_inverse;
  P.half = fp32_inverse && half_application;
  P.eligible.reserve(n);
  P.seed_flag.reserve(n);
  P.seed_pos.reserve(n + 1);
  classify_rows_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(
      n, diag, offabs, P.eligible.data(), P.seed_flag.data());
  LSMPS_LAUNCH_CHECK();
  exclusive_scan_i32(ctx, P.seed_flag.data(), P.seed_pos.data(), n);
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, P.seed_pos.data() + n, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  const size_t nb = static_cast<size_t>(ctx.host_i32[0]);
  P.blocks = nb;
  P.covered_count = 0;
  if (nb == 0) return;
  P.seeds.reserve(nb);
  compact_seeds_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, P.seed_flag.data(),
                                                                 P.seed_pos.data(), P.seeds.data());
  P.members.reserve(nb * kMaxBlock);
  P.sizes.reserve(nb);
  members_kernel<<<static_cast<unsigned>(nb), kMaxBlock, 0, ctx.stream>>>(
      P.seeds.data(), A.row_offsets, A.col_indices, P.eligible.data(), P.members.data(),
      P.sizes.data());
  LSMPS_LAUNCH_CHECK();
  const size_t dense = nb * kMaxBlock * kMaxBlock;
  if (fp32_inverse) {
    P.inv32.reserve(dense);
    gj_inverse_kernel<<<static_cast<unsigned>(nb), kGjThreads, 0, ctx.stream>>>(
        A.row_offsets, A.col_indices, A.values, P.members.data(), P.sizes.data(),
        P.inv32.data());
    LSMPS_LAUNCH_CHECK();
    if (P.half) quantize_blocks(ctx, P);
  } else {
    P.scratch.reserve(2 * dense);
    P.inv64.reserve(dense);
    factor_kernel<double><<<static_cast<unsigned>(nb), kMaxBlock, 0, ctx.stream>>>(
        A.row_offsets, A.col_indices, A.values, P.members.data(), P.sizes.data(),
        P.scratch.data(), P.inv64.data());
  }
  LSMPS_LAUNCH_CHECK();
  // Row -> covering slots (CSR over rows), sorted per row.
  P.cover.reserve(n);
  P.cover_start.reserve(n + 1);
  P.ybuf_stride = nb * kMaxBlock;
  P.ybuf.reserve(nb * kApplyParts * kMaxBlock);
  fill_i32(ctx.stream, P.cover.data(), n, 0);
  cover_count_kernel<<<static_cast<unsigned>(nb), kMaxBlock, 0, ctx.stream>>>(
      P.members.data(), P.sizes.data(), P.cover.data());
  LSMPS_LAUNCH_CHECK();
  exclusive_scan_i32(ctx, P.cover.data(), P.cover_start.data(), n);
  P.cover_entries.reserve(nb * kMaxBlock);
  fill_i32(ctx.stream, P.cover.data(), n, 0);
  cover_fill_kernel<<<static_cast<unsigned>(nb), kMaxBlock, 0, ctx.stream>>>(
      P.members.data(), P.sizes.data(), P.cover_start.data(), P.cover.data(),
      P.cover_entries.data());
  cover_sort_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, P.cover_start.data(),
                                                             P.cover_entries.data());
  LSMPS_LAUNCH_CHECK();
  // every slot b * kMaxBlock + t with t < sizes[b] is one row's entry; the
  // others are never written by the products
  P.entry_of.reserve(nb * kMaxBlock);
  entry_of_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, P.cover_start.data(),
                                                           P.cover_entries.data(),
                                                           P.entry_of.data());
  // preconditioner diagonal and the list of covered rows (flags and scan reuse
  // the seed arrays, which are no longer needed)
  P.pdiag.reserve(n);
  pdiag_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, P.cover_start.data(), dinv,
                                                         P.pdiag.data(), P.seed_flag.data());
  LSMPS_LAUNCH_CHECK();
  exclusive_scan_i32(ctx, P.seed_flag.data(), P.seed_pos.data(), n);
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, P.seed_pos.data() + n, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  P.covered_count = static_cast<size_t>(ctx.host_i32[0]);
  P.covered.reserve(std::max<size_t>(P.covered_count, 1));
  compact_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, P.seed_flag.data(),
                                                           P.seed_pos.data(), P.covered.data());
  LSMPS_LAUNCH_CHECK();
}

namespace {
// dst block b = transpose of src block b (column-major kMaxBlock^2 each).
template <typename V>
__global__ void __launch_bounds__(kBlock) transpose_blocks_kernel(size_t blocks, const V* src,
                                                                  V* dst) {
  constexpr size_t nn = static_cast<size_t>(kMaxBlock) * kMaxBlock;
  const size_t total = blocks * nn;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t b = e / nn, k = e % nn;
    const size_t r = k % kMaxBlock, c = k / kMaxBlock;
    dst[b * nn + c + r * kMaxBlock] = src[e];
  }
}

template <typename T>
void copy_array(Context& ctx, DeviceArray<T>& dst, const DeviceArray<T>& src, size_t count) {
  dst.reserve(std::max<size_t>(count, 1));
  if (count)
    LSMPS_HIP(hipMemcpyAsync(dst.data(), src.data(), count * sizeof(T), hipMemcpyDeviceToDevice,
                             ctx.stream));
}
}  // namespace

void schwarz_transpose_of(Context& ctx, SchwarzPrec& P, const SchwarzPrec& src,
                          bool half_application) {
  const size_t n = src.n, nb = src.blocks;
  P.n = n;
  P.blocks = nb;
  P.fp32 = src.fp32;
  P.half = src.fp32 && half_application;
  P.covered_count = src.covered_count;
  if (nb == 0) return;
  const size_t dense = nb * kMaxBlock * kMaxBlock;
  copy_array(ctx, P.members, src.members, nb * kMaxBlock);
  copy_array(ctx, P.sizes, src.sizes, nb);
  copy_array(ctx, P.cover_start, src.cover_start, n + 1);
  copy_array(ctx, P.cover_entries, src.cover_entries, nb * kMaxBlock);
  copy_array(ctx, P.entry_of, src.entry_of, nb * kMaxBlock);
  copy_array(ctx, P.covered, src.covered, std::max<size_t>(src.covered_count, 1));
  copy_array(ctx, P.pdiag, src.pdiag, n);
  P.ybuf_stride = nb * kMaxBlock;
  P.ybuf.reserve(nb * kApplyParts * kMaxBlock);
  const unsigned grid = static_cast<unsigned>(std::min<size_t>(blocks_for(dense), 8192));
  if (src.fp32) {
    P.inv32.reserve(dense);
    transpose_blocks_kernel<float><<<grid, kBlock, 0, ctx.stream>>>(nb, src.inv32.data(),
                                                                     P.inv32.data());
    LSMPS_LAUNCH_CHECK();
    // the transposed columns are the source's rows: their scales are new
    if (P.half) quantize_blocks(ctx, P);
  } else {
    P.inv64.reserve(dense);
    transpose_blocks_kernel<double><<<grid, kBlock, 0, ctx.stream>>>(nb, src.inv64.data(),
                                                                      P.inv64.data());
  }
  LSMPS_LAUNCH_CHECK();
}

template <typename T>
void schwarz_products(Context& ctx, const SchwarzPrec& P, const T* v, const int32_t* skip) {
  if (P.blocks == 0) return;
  const unsigned grid = static_cast<unsigned>(P.blocks * kApplyParts);
  if (P.half) {
    apply_kernel<T, _Float16><<<grid, kMaxBlock, 0, ctx.stream>>>(
        P.members.data(), P.sizes.data(), reinterpret_cast<const _Float16*>(P.inv16.data()),
        P.col_scale.data(), v, P.entry_of.data(), P.ybuf.data(), P.ybuf_stride, skip);
  } else if (P.fp32) {
    apply_kernel<T, float><<<grid, kMaxBlock, 0, ctx.stream>>>(P.members.data(), P.sizes.data(),
                                                               P.inv32.data(), nullptr, v,
                                                               P.entry_of.data(), P.ybuf.data(),
                                                               P.ybuf_stride, skip);
  } else {
    apply_kernel<T, double><<<grid, kMaxBlock, 0, ctx.stream>>>(P.members.data(), P.sizes.data(),
                                                                P.inv64.data(), nullptr, v,
                                                                P.entry_of.data(), P.ybuf.data(),
                                                                P.ybuf_stride, skip);
  }
  LSMPS_LAUNCH_CHECK();
}

template <typename T>
void schwarz_complete(Context& ctx, const SchwarzPrec& P, T* v, const int32_t* skip) {
  if (P.blocks == 0) return;
  gather_kernel<T><<<blocks_for(std::max<size_t>(P.covered_count, 1)), kBlock, 0, ctx.stream>>>(
      P.covered_count, P.covered.data(), P.cover_start.data(), P.ybuf.data(), P.ybuf_stride, v,
      skip);
  LSMPS_LAUNCH_CHECK();
}

template <typename T>
void schwarz_apply(Context& ctx, const SchwarzPrec& P, T* v, const int32_t* skip) {
  if (P.blocks == 0) return;
  schwarz_products<T>(ctx, P, v, skip);
  schwarz_complete<T>(ctx, P, v, skip);
}

template void schwarz_complete<double>(Context&, const SchwarzPrec&, double*, const int32_t*);
template void schwarz_complete<float>(Context&, const SchwarzPrec&, float*, const int32_t*);
template void schwarz_apply<double>(Context&, const SchwarzPrec&, double*, const int32_t*);
template void schwarz_apply<float>(Context&, const SchwarzPrec&, float*, const int32_t*);
template void schwarz_products<double>(Context&, const SchwarzPrec&, const double*,
                                       const int32_t*);
template void schwarz_products<float>(Context&, const SchwarzPrec&, const float*,
                                      const int32_t*);

void schwarz_stats(Context& ctx, const SchwarzPrec& P, int64_t* blocks, int64_t* members) {
  *blocks = static_cast<int64_t>(P.blocks);
  *members = 0;
  if (P.blocks == 0) return;
  std::vector<int32_t> sizes(P.blocks);
  LSMPS_HIP(hipMemcpyAsync(sizes.data(), P.sizes.data(), P.blocks * sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  for (int32_t s : sizes) *members += s;
}

}  // namespace lsmps
