// Coarse-space correction of the PPE preconditioner: aggregation over blocks of
// cells, Galerkin coarse matrix in exact fixed point, Gauss-Jordan inverse, and
// the per-application restriction / coarse product / prolongation.
//
// Kernels that reduce over an aggregate or a coarse row give each one a
// wavefront: lanes take the items in a fixed stride and the partial sums meet
// in a fixed butterfly, so the sums are reproducible.
#include <algorithm>
#include <cmath>

#include "coarse.hpp"
#include "schwarz.hpp"
#include "stages.hpp"

namespace lsmps {
namespace {

constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ bool aggregated_label(int32_t label) {
  return label == LSMPS_SURFACE_INNER || label == LSMPS_SURFACE_NEAR_SURFACE;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct CoarseGrid {
  int32_t nx, ny, nz;     // fine cells
  int32_t cx, cy, cz;     // coarse cells
  int32_t f;
};

// Calls visit(row) for the members of coarse cell cc's fine cells in (z, y, x)
// order, 64 candidates at a time (lane l gets candidate base + l, or -1).
template <typename F>
__device__ void for_each_candidate(const CoarseGrid& g, size_t cc, const int32_t* sorted,
                                   const int32_t* begin, const int32_t* end, int lane, F visit) {
  const int32_t ccx = static_cast<int32_t>(cc % g.cx);
  const int32_t ccy = static_cast<int32_t>((cc / g.cx) % g.cy);
  const int32_t ccz = static_cast<int32_t>(cc / (static_cast<size_t>(g.cx) * g.cy));
  for (int32_t dz = 0; dz < g.f; ++dz) {
    const int32_t z = ccz * g.f + dz;
    if (z >= g.nz) break;
    for (int32_t dy = 0; dy < g.f; ++dy) {
      const int32_t y = ccy * g.f + dy;
      if (y >= g.ny) break;
      for (int32_t dx = 0; dx < g.f; ++dx) {
        const int32_t x = ccx * g.f + dx;
        if (x >= g.nx) break;
        const size_t cell = (static_cast<size_t>(z) * g.ny + y) * g.nx + x;
        const int32_t b = begin[cell];
        if (b < 0) continue;
        const int32_t e = end[cell];
        for (int32_t base = b; base < e; base += 64) {
          const int32_t k = base + lane;
          visit(k < e ? sorted[k] : -1);
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock) count_members_kernel(
    CoarseGrid g, size_t ccells, const int32_t* labels, const int32_t* sorted,
    const int32_t* begin, const int32_t* end, int32_t* count, int32_t* flag) {
  const size_t cc = static_cast<size_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (cc >= ccells) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  int c = 0;
  for_each_candidate(g, cc, sorted, begin, end, lane, [&](int32_t row) {
    if (row >= 0 && aggregated_label(labels[row])) ++c;
  });
  c = wave_sum_i(c);
  if (lane == 0) {
    count[cc] = c;
    flag[cc] = c > 0 ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kBlock) fill_members_kernel(
    CoarseGrid g, size_t ccells, const int32_t* labels, const int32_t* sorted,
    const int32_t* begin, const int32_t* end, const int32_t* count, const int32_t* start,
    const int32_t* id, int32_t* agg_start, int32_t* agg_rows, int32_t* agg_of_row) {
  const size_t cc = static_cast<size_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (cc >= ccells) return;
  if (count[cc] == 0) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const int32_t a = id[cc];
  int32_t pos = start[cc];
  if (lane == 0) agg_start[a] = pos;
  for_each_candidate(g, cc, sorted, begin, end, lane, [&](int32_t row) {
    const bool take = row >= 0 && aggregated_label(labels[row]);
    const unsigned long long mask = __ballot(take);
    if (take) {
      const int before = __popcll(mask & ((1ull << lane) - 1ull));
      agg_rows[pos + before] = row;
      agg_of_row[row] = a;
    }
    pos += __popcll(mask);
  });
}

__global__ void set_last_kernel(int32_t* agg_start, int count, const int32_t* counts,
                                const int32_t* starts, size_t ccells) {
  agg_start[count] = starts[ccells - 1] + counts[ccells - 1];
}

__global__ void fill_i32(int32_t* p, size_t n, int32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- batched tanks: aggregation over each tank's own grid -----------------

// Largest k in [0, tanks) with start[k] <= v (start ascending).
__device__ __forceinline__ int tank_of(const int32_t* start, int tanks, int64_t v) {
  int lo = 0, hi = tanks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= v) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

__device__ __forceinline__ CoarseGrid grid_of(const TankCoarseGrid& t) {
  return CoarseGrid{t.nx, t.ny, t.nz, t.cx, t.cy, t.cz, t.f};
}

// count_members_kernel over the coarse cells of every tank (coarse cell cc
// belongs to the tank whose range of tank_cc holds it; its fine cells are
// that tank's, offset by the tank's first cell).
__global__ void __launch_bounds__(kBlock) count_tank_members_kernel(
    const TankCoarseGrid* grids, const int32_t* tank_cc, int tanks, size_t ccells,
    const int32_t* labels, const int32_t* sorted, const int32_t* begin, const int32_t* end,
    int32_t* count, int32_t* flag) {
  const size_t cc = static_cast<size_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (cc >= ccells) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const TankCoarseGrid t = grids[tank_of(tank_cc, tanks, static_cast<int64_t>(cc))];
  int c = 0;
  for_each_candidate(grid_of(t), cc - static_cast<size_t>(t.cc_base), sorted, begin + t.base,
                     end + t.base, lane, [&](int32_t row) {
                       if (row >= 0 && aggregated_label(labels[row])) ++c;
                     });
  c = wave_sum_i(c);
  if (lane == 0) {
    count[cc] = c;
    flag[cc] = c > 0 ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kBlock) fill_tank_members_kernel(
    const TankCoarseGrid* grids, const int32_t* tank_cc, int tanks, size_t ccells,
    const int32_t* labels, const int32_t* sorted, const int32_t* begin, const int32_t* end,
    const int32_t* count, const int32_t* start, const int32_t* id, int32_t* agg_start,
    int32_t* agg_rows, int32_t* agg_of_row) {
  const size_t cc = static_cast<size_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (cc >= ccells) return;
  if (count[cc] == 0) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const TankCoarseGrid t = grids[tank_of(tank_cc, tanks, static_cast<int64_t>(cc))];
  const int32_t a = id[cc];
  int32_t pos = start[cc];
  if (lane == 0) agg_start[a] = pos;
  for_each_candidate(grid_of(t), cc - static_cast<size_t>(t.cc_base), sorted, begin + t.base,
                     end + t.base, lane, [&](int32_t row) {
                       const bool take = row >= 0 && aggregated_label(labels[row]);
                       const unsigned long long mask = __ballot(take);
                       if (take) {
                         const int before = __popcll(mask & ((1ull << lane) - 1ull));
                         agg_rows[pos + before] = row;
                         agg_of_row[row] = a;
                       }
                       pos += __popcll(mask);
                     });
}

// tank_agg[k] = first aggregate of tank k (k < tanks), tank_agg[tanks] = the
// aggregate count and tank_agg[tanks + 1] = the member count.
__global__ void tank_ranges_kernel(const int32_t* tank_cc, int tanks, size_t ccells,
                                   const int32_t* cell_count, const int32_t* cell_flag,
                                   const int32_t* cell_start, const int32_t* cell_id,
                                   int32_t* tank_agg) {
  const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < tanks) tank_agg[k] = cell_id[tank_cc[k]];
  if (k == tanks) tank_agg[k] = cell_id[ccells - 1] + cell_flag[ccells - 1];
  if (k == tanks + 1) tank_agg[k] = cell_start[ccells - 1] + cell_count[ccells - 1];
}

// agg_tank[I] = k for the aggregates of tank k (one block per tank).
__global__ void __launch_bounds__(kBlock) agg_tank_kernel(const int32_t* tank_agg,
                                                          int32_t* agg_tank) {
  const int k = blockIdx.x;
  for (int32_t I = tank_agg[k] + threadIdx.x; I < tank_agg[k + 1]; I += kBlock) agg_tank[I] = k;
}

// Ac = P^T A P in exact 64-bit fixed point.  Row I's scale 2^s_I is set from a
// bound B_I >= |Ac_IJ| (the sum over its members of sum_j |a_ij|) so that
// 2^s_I B_I < 2^61: every a_ij enters as round(a_ij 2^s_I) and all additions
// are exact integer additions, whose order does not matter.

// Row bound |a_ii| + sum |a_ij|, j != i, for the aggregated rows (others: 0).
__global__ void __launch_bounds__(kBlock) row_bound_kernel(size_t n, const int32_t* agg_of_row,
                                                           const double* diag,
                                                           const double* offabs, double* out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = agg_of_row[i] >= 0 ? fabs(diag[i]) + offabs[i] : 0.0;
}

__device__ __forceinline__ double row_scale(double B) {
  return B > 0.0 ? ldexp(1.0, 60 - ilogb(B)) : 0.0;
}

// Block (I, part) adds the entries of its share of aggregate I's members into
// an LDS row, then the nonzero LDS entries into row I of the integer matrix.
// Sixteen lanes share a member row and stride over its entries (coalesced CSR
// reads); the LDS atomics are integer additions, so their order is free.
constexpr int kMatrixParts = 8;
constexpr int kMatrixLanes = 16;
__global__ void __launch_bounds__(kBlock) coarse_matrix_kernel(
    int count, const int32_t* agg_start, const int32_t* agg_rows, const int32_t* agg_of_row,
    const int32_t* rows, const int32_t* cols, const double* vals, const double* bound,
    unsigned long long* out) {
  __shared__ unsigned long long acc[kMaxCoarse];
  constexpr int kGroups = kBlock / kMatrixLanes;
  const int I = blockIdx.x / kMatrixParts;
  const int part = blockIdx.x % kMatrixParts;
  const int group = threadIdx.x / kMatrixLanes;
  const int l = threadIdx.x % kMatrixLanes;
  const double scale = row_scale(bound[I]);
  for (int J = threadIdx.x; J < count; J += kBlock) acc[J] = 0ull;
  __syncthreads();
  if (scale > 0.0) {
    const int32_t me = agg_start[I + 1];
    for (int32_t m = agg_start[I] + part * kGroups + group; m < me;
         m += kMatrixParts * kGroups) {
      const int32_t i = agg_rows[m];
      for (int32_t e = rows[i] + l; e < rows[i + 1]; e += kMatrixLanes) {
        const int32_t J = agg_of_row[cols[e]];
        if (J >= 0)
          atomicAdd(&acc[J], static_cast<unsigned long long>(llrint(vals[e] * scale)));
      }
    }
  }
  __syncthreads();
  unsigned long long* row = out + static_cast<size_t>(I) * count;
  for (int J = threadIdx.x; J < count; J += kBlock)
    if (acc[J] != 0ull) atomicAdd(&row[J], acc[J]);
}

__global__ void __launch_bounds__(kBlock) coarse_convert_kernel(int count, const double* bound,
                                                                const unsigned long long* in,
                                                                double* out) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<size_t>(count) * count) return;
  const double scale = row_scale(bound[e / count]);
  out[e] = scale > 0.0 ? static_cast<double>(static_cast<long long>(in[e])) / scale : 0.0;
}

__global__ void __launch_bounds__(kBlock) fill_u64(unsigned long long* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// coarse_matrix_kernel for batched tanks: row I of its tank's block (columns
// local to the tank; the tanks' rows never couple, and an entry outside the
// tank's range is not added).
__global__ void __launch_bounds__(kBlock) tank_matrix_kernel(
    const int32_t* agg_tank, const int32_t* tank_agg, const int64_t* block_offset,
    const int32_t* agg_start, const int32_t* agg_rows, const int32_t* agg_of_row,
    const int32_t* rows, const int32_t* cols, const double* vals, const double* bound,
    unsigned long long* out) {
  __shared__ unsigned long long acc[kMaxTankCoarse];
  constexpr int kGroups = kBlock / kMatrixLanes;
  const int I = blockIdx.x / kMatrixParts;
  const int part = blockIdx.x % kMatrixParts;
  const int group = threadIdx.x / kMatrixLanes;
  const int l = threadIdx.x % kMatrixLanes;
  const int tank = agg_tank[I];
  const int32_t lo = tank_agg[tank];
  const int32_t c = tank_agg[tank + 1] - lo;  // <= kMaxTankCoarse
  const double scale = row_scale(bound[I]);
  for (int J = threadIdx.x; J < c; J += kBlock) acc[J] = 0ull;
  __syncthreads();
  if (scale > 0.0) {
    const int32_t me = agg_start[I + 1];
    for (int32_t m = agg_start[I] + part * kGroups + group; m < me;
         m += kMatrixParts * kGroups) {
      const int32_t i = agg_rows[m];
      for (int32_t e = rows[i] + l; e < rows[i + 1]; e += kMatrixLanes) {
        const int32_t J = agg_of_row[cols[e]] - lo;
        if (J >= 0 && J < c)
          atomicAdd(&acc[J], static_cast<unsigned long long>(llrint(vals[e] * scale)));
      }
    }
  }
  __syncthreads();
  unsigned long long* row = out + block_offset[tank] + static_cast<size_t>(I - lo) * c;
  for (int J = threadIdx.x; J < c; J += kBlock)
    if (acc[J] != 0ull) atomicAdd(&row[J], acc[J]);
}

// Integer row I of its tank's block to doubles (one block per aggregate).
__global__ void __launch_bounds__(kBlock) tank_convert_kernel(const int32_t* agg_tank,
                                                              const int32_t* tank_agg,
                                                              const int64_t* block_offset,
                                                              const double* bound,
                                                              const unsigned long long* in,
                                                              double* out) {
  const int I = blockIdx.x;
  const int tank = agg_tank[I];
  const int32_t lo = tank_agg[tank];
  const int32_t c = tank_agg[tank + 1] - lo;
  const double scale = row_scale(bound[I]);
  const size_t base = block_offset[tank] + static_cast<size_t>(I - lo) * c;
  for (int J = threadIdx.x; J < c; J += kBlock)
    out[base + J] = scale > 0.0 ? static_cast<double>(static_cast<long long>(in[base + J])) / scale
                                : 0.0;
}

// Blocked Gauss-Jordan inverse with partial pivoting.  Step k: pivot row p >= k
// of largest |a_rk| (lowest index on ties), swap rows k and p, scale row k by
// inv = 1 / a_kk with a_kk := 1 first (the in-place inverse), then for every
// other row r, with f = a_rk before the step: a_rk := -f inv and
// a_rj -= f a_kj (j != k).
//
// On the columns a step does not eliminate it is A := (I + (z - e_k) e_k^T) A,
// z being column k after the step (z_k = inv, z_r = -f_r inv).  b consecutive
// steps compose to A := (I + (W - E) E^T) A, where W holds the panel's b
// columns after its b steps and E the identity's columns k0 .. k0 + b - 1.
// So the steps run in panels of b columns:
//  1. gj_panel_kernel (one workgroup, the panel's rows in registers) applies
//     the b steps to the panel's columns and records their pivot rows;
//  2. gj_swap_kernel applies the b row interchanges to every other column and
//     copies the interchanged panel rows R = A[k0 .. k0 + b) of those columns
//     aside;
//  3. gj_rank_kernel adds (W - E) R to every other column: a rank-b update,
//     count x b times b x count, on the FP64 matrix cores
//     (v_mfma_f64_16x16x4f64, one wavefront per 16 x 16 tile).
// Row k0 + i of the update is W[k0 + i] R (its identity term cancels the
// interchanged row), every other row r adds W[r] R; each sum runs in the
// matrix cores' fixed order, so the inverse is reproducible run to run.
constexpr int kPanelThreads = 1024;
constexpr int kPanelWaves = kPanelThreads / 64;
static_assert(kPanelWaves <= 64 && (kPanelWaves & (kPanelWaves - 1)) == 0, "one partial per lane");
static_assert(kMaxCoarse <= 4 * kPanelThreads, "panel rows per thread");
constexpr int kMaxPanel = 16;  // widest panel (the R buffer holds kMaxPanel rows)

// Panel width: at most 64 register doubles per thread (rows per thread x
// width) and 16 unrolled steps (a 32-step unroll is not taken by the compiler,
// which then keeps the panel in scratch memory).
int panel_width(int count) { return count <= 2 * kPanelThreads ? 16 : 8; }

__device__ __forceinline__ void better_pivot(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

// The panel's w <= B steps on its columns [k0, k0 + w); perm[k] receives the
// pivot row of step k.  Thread t of T holds rows t, t + T, ... (R of
// them, covering count).  Per step the pivot partials, the old row k and the
// new row k (row p scaled by its owner) pass through LDS, double-buffered by
// step parity: two barriers per step.  The step loop is unrolled, so that
// every register index is a constant (a run-time column index would send the
// panel to scratch memory).
template <int R, int B, int T>
__global__ void __launch_bounds__(T) gj_panel_kernel(double* a, int count, int k0, int w,
                                                     int32_t* perm, int32_t* status) {
  constexpr int kWavesT = T / 64;
  static_assert(kWavesT <= 64 && (kWavesT & (kWavesT - 1)) == 0, "one partial per lane");
  __shared__ double wv[2][kWavesT];
  __shared__ int wi[2][kWavesT];
  __shared__ double xk[2][B], xr[2][B];
  if (status[0]) return;  // an earlier panel met a zero pivot (uniform)
  const int t = threadIdx.x;
  double P[R][B];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = t + q * T;
#pragma unroll
    for (int c = 0; c < B; ++c)
      P[q][c] = r < count && c < w ? a[static_cast<size_t>(r) * count + k0 + c] : 0.0;
  }
#pragma unroll
  for (int s = 0; s < B; ++s) {
    if (s < w) {  // uniform; a constant trip count lets the loop unroll
      const int k = k0 + s;
      const int buf = s & 1;
      double best = -1.0;
      int bidx = count;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
        const double v = fabs(P[q][s]);
        if (r >= k && r < count && v > best) {
          best = v;
          bidx = r;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(best, off, 64);
        const int i2 = __shfl_xor(bidx, off, 64);
        better_pivot(best, bidx, v2, i2);
      }
      if ((t & 63) == 0) {
        wv[buf][t >> 6] = best;
        wi[buf][t >> 6] = bidx;
      }
      __syncthreads();
      {  // every wavefront reduces the kWavesT partials the same way
        const int l = t & 63;
        best = l < kWavesT ? wv[buf][l] : -1.0;
        bidx = l < kWavesT ? wi[buf][l] : count;
#pragma unroll
        for (int off = kWavesT / 2; off > 0; off >>= 1) {
          const double v2 = __shfl_xor(best, off, 64);
          const int i2 = __shfl_xor(bidx, off, 64);
          better_pivot(best, bidx, v2, i2);
        }
        best = __shfl(best, 0, 64);
        bidx = __shfl(bidx, 0, 64);
      }
      const bool singular = !(best > 0.0) || !isfinite(best) || bidx >= count;
      if (singular) {
        if (t == 0) {
          perm[k] = k;
          status[0] = 1;
        }
        return;  // uniform: every thread reduced the same partials
      }
      const int p = bidx;
      if (t == 0) perm[k] = p;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
        if (r == k) {
#pragma unroll
          for (int c = 0; c < B; ++c) xk[buf][c] = P[q][c];
        }
        if (r == p) {
          const double inv = 1.0 / P[q][s];
#pragma unroll
          for (int c = 0; c < B; ++c) xr[buf][c] = (c == s ? 1.0 : P[q][c]) * inv;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
        if (r >= count) continue;
        if (r == k) {
#pragma unroll
          for (int c = 0; c < B; ++c) P[q][c] = xr[buf][c];
          continue;
        }
        const bool is_p = r == p;  // row p now holds the old row k
        const double f = is_p ? xk[buf][s] : P[q][s];
#pragma unroll
        for (int c = 0; c < B; ++c) {
          const double src = is_p ? xk[buf][c] : P[q][c];
          P[q][c] = (c == s ? 0.0 : src) - f * xr[buf][c];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = t + q * T;
    if (r >= count) continue;
#pragma unroll
    for (int c = 0; c < B; ++c)
      if (c < w) a[static_cast<size_t>(r) * count + k0 + c] = P[q][c];
  }
}

// The panel's row interchanges on every column outside it (one thread per
// column, the interchanges in step order), then its rows of those columns
// copied to panel_rows (w x count).
__global__ void __launch_bounds__(kBlock) gj_swap_kernel(double* a, int count, int k0, int w,
                                                         const int32_t* perm,
                                                         const int32_t* status,
                                                         double* panel_rows) {
  if (status[0]) return;
  const int j = static_cast<int>(blockIdx.x) * kBlock + threadIdx.x;
  if (j >= count || (j >= k0 && j < k0 + w)) return;
  for (int s = 0; s < w; ++s) {
    const int k = k0 + s;
    const int p = perm[k];
    if (p != k) {
      const double x = a[static_cast<size_t>(k) * count + j];
      a[static_cast<size_t>(k) * count + j] = a[static_cast<size_t>(p) * count + j];
      a[static_cast<size_t>(p) * count + j] = x;
    }
  }
  for (int s = 0; s < w; ++s)
    panel_rows[static_cast<size_t>(s) * count + j] = a[static_cast<size_t>(k0 + s) * count + j];
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// a[:, j] += (W - E) R[:, j] for the columns j outside the panel.  Wavefront
// per 16 x 16 tile (rows i0.., columns j0..); per K = 4 slice of the panel
// lane l supplies (W - E)[i0 + l % 16][kk + l / 16] and R[kk + l / 16][j0 + l % 16]
// and accumulates rows i0 + l / 16 + 4 v (v < 4) of column j0 + l % 16 (the
// instruction's result layout, checked on the device).
__global__ void __launch_bounds__(kBlock) gj_rank_kernel(double* a, int count, int k0, int w,
                                                         const double* panel_rows,
                                                         const int32_t* status) {
  if (status[0]) return;
  const int tiles = (count + 15) / 16;
  const int tile = static_cast<int>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (tile >= tiles * tiles) return;  // uniform per wavefront
  const int i0 = (tile / tiles) * 16, j0 = (tile % tiles) * 16;
  if (j0 >= k0 && j0 + 16 <= k0 + w) return;  // every column in the panel
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int jc = j0 + li;
  const bool col_ok = jc < count && !(jc >= k0 && jc < k0 + w);
  double4_t acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + lk + 4 * v;
    acc[v] = col_ok && i < count ? a[static_cast<size_t>(i) * count + jc] : 0.0;
  }
  const int ia = i0 + li;
  for (int kk = 0; kk < w; kk += 4) {
    const int ka = kk + lk;
    double av = 0.0, bv = 0.0;
    if (ia < count && ka < w)
      av = a[static_cast<size_t>(ia) * count + k0 + ka] - (ia == k0 + ka ? 1.0 : 0.0);
    if (col_ok && ka < w) bv = panel_rows[static_cast<size_t>(ka) * count + jc];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + lk + 4 * v;
    if (col_ok && i < count) a[static_cast<size_t>(i) * count + jc] = acc[v];
  }
}

// The row swaps of the pivoting become column swaps of the inverse, applied in
// reverse step order: out[:, j] = a[:, idx[j]] with idx the composed order.
// After a zero pivot (status set) the later steps never ran and their perm
// entries were never written: the order is then left as the identity (the
// inverse is unusable anyway and its products are zero).
// (One workgroup; the order is composed in LDS by one thread.)
__global__ void __launch_bounds__(kBlock) gj_order_kernel(int count, const int32_t* perm,
                                                          const int32_t* status, int32_t* idx) {
  __shared__ int32_t pv[kMaxCoarse], order[kMaxCoarse];
  const bool valid = status[0] == 0;  // uniform
  for (int j = threadIdx.x; j < count; j += kBlock) {
    pv[j] = valid ? perm[j] : j;
    order[j] = j;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = count - 1; k >= 0; --k) {
      const int p = pv[k];
      if (p < k || p >= count) continue;  // a pivot row is never above its step
      const int32_t tmp = order[k];
      order[k] = order[p];
      order[p] = tmp;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < count; j += kBlock) idx[j] = order[j];
}

__global__ void __launch_bounds__(kBlock) gj_permute_kernel(const double* a, int count,
                                                            const int32_t* idx, double* out) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(count) * count;
  if (e >= total) return;
  const size_t r = e / count;
  const int j = static_cast<int>(e % count);
  out[e] = a[r * count + idx[j]];
}

// Gauss-Jordan inverse of each tank's block (one workgroup per tank, c <=
// kMaxTankCoarse = kBlock rows: thread t owns row t in the pivot search), with
// the steps of the blocked elimination above applied one column at a time:
// pivot p >= k of largest |a_pk| (lowest index on ties), rows k and p swapped,
// row k scaled by inv = 1 / a_kk with a_kk := 1, then a_rj -= f a_kj for every
// other row r (f = a_rk before the step, a_rk taken as 0).  work holds the
// blocks on entry and is overwritten; ainv receives the column-permuted result.
__global__ void __launch_bounds__(kBlock) tank_gj_kernel(const int32_t* tank_agg,
                                                         const int64_t* block_offset,
                                                         double* work, double* ainv,
                                                         int32_t* status) {
  static_assert(kMaxTankCoarse <= kBlock, "one pivot candidate per thread");
  __shared__ double rowk[kMaxTankCoarse], colf[kMaxTankCoarse];
  __shared__ int32_t perm[kMaxTankCoarse], idx[kMaxTankCoarse];
  __shared__ double wv[kWaves];
  __shared__ int wi[kWaves];
  const int tank = blockIdx.x;
  const int c = tank_agg[tank + 1] - tank_agg[tank];
  if (c <= 0) return;  // uniform
  double* a = work + block_offset[tank];
  const int t = threadIdx.x;
  const size_t cc = static_cast<size_t>(c);
  for (int k = 0; k < c; ++k) {
    double best = t >= k && t < c ? fabs(a[static_cast<size_t>(t) * cc + k]) : -1.0;
    int bidx = t >= k && t < c ? t : c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double v2 = __shfl_xor(best, off, 64);
      const int i2 = __shfl_xor(bidx, off, 64);
      better_pivot(best, bidx, v2, i2);
    }
    if ((t & 63) == 0) {
      wv[t >> 6] = best;
      wi[t >> 6] = bidx;
    }
    __syncthreads();
    best = wv[0];
    bidx = wi[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) better_pivot(best, bidx, wv[w], wi[w]);
    const bool singular = !(best > 0.0) || !isfinite(best) || bidx >= c;
    if (singular) {
      if (t == 0) status[tank] = 1;
      return;  // uniform: every thread reduced the same partials
    }
    const int p = bidx;
    const double inv = 1.0 / a[static_cast<size_t>(p) * cc + k];
    if (t == 0) perm[k] = p;
    __syncthreads();  // every thread has read the pivot and the partials
    for (int j = t; j < c; j += kBlock) {
      const double xk = a[static_cast<size_t>(k) * cc + j];
      const double xp = a[static_cast<size_t>(p) * cc + j];
      if (p != k) a[static_cast<size_t>(p) * cc + j] = xk;
      const double v = (j == k ? 1.0 : xp) * inv;
      a[static_cast<size_t>(k) * cc + j] = v;
      rowk[j] = v;
    }
    __syncthreads();
    for (int r = t; r < c; r += kBlock) colf[r] = r == k ? 0.0 : a[static_cast<size_t>(r) * cc + k];
    __syncthreads();
    for (size_t e = t; e < cc * cc; e += kBlock) {
      const int r = static_cast<int>(e / cc);
      const int j = static_cast<int>(e % cc);
      if (r == k) continue;
      a[e] = (j == k ? 0.0 : a[e]) - colf[r] * rowk[j];
    }
    __syncthreads();
  }
  // the row swaps become column swaps of the inverse, in reverse step order
  if (t == 0) {
    for (int j = 0; j < c; ++j) idx[j] = j;
    for (int k = c - 1; k >= 0; --k) {
      const int p = perm[k];
      const int32_t tmp = idx[k];
      idx[k] = idx[p];
      idx[p] = tmp;
    }
  }
  __syncthreads();
  double* out = ainv + block_offset[tank];
  for (size_t e = t; e < cc * cc; e += kBlock)
    out[e] = a[(e / cc) * cc + idx[e % cc]];
}

// yc[I] = (row I of its tank's inverse block) . rc over the tank's aggregates;
// one wavefront per aggregate, lanes strided over the columns, fixed butterfly.
__global__ void __launch_bounds__(kBlock) tank_product_kernel(
    int count, const int32_t* agg_tank, const int32_t* tank_agg, const int64_t* block_offset,
    const double* ainv, const double* rc, const int32_t* status, double* yc,
    const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  const int I = static_cast<int>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (I >= count) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const int tank = agg_tank[I];
  const int32_t lo = tank_agg[tank];
  const int32_t c = tank_agg[tank + 1] - lo;
  double s = 0.0;
  if (!status[tank]) {  // a singular block gives its tank no correction
    const double* row = ainv + block_offset[tank] + static_cast<size_t>(I - lo) * c;
    for (int32_t J = lane; J < c; J += 64) s += row[J] * rc[lo + J];
  }
  s = wave_sum(s);
  if (lane == 0) yc[I] = s;
}

__global__ void __launch_bounds__(kBlock) inverse_kernel(size_t n, const double* pdiag,
                                                         double* wfac) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wfac[i] = 1.0 / pdiag[i];
}

// Block-wide sum with a fixed assignment (thread t takes items t, t + 256, ...)
// and a fixed tree, so the result does not depend on timing.
__device__ __forceinline__ double block_sum(double s, double* red) {
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += red[w];
  return t;
}

// rc[I] = sum over the members i of aggregate I of w_i (times wfac_i), by one
// kBlock-thread workgroup: up to kInFlight members per thread in flight (an
// aggregate has ~1400 rows, so usually one round), summed in member order, then
// the fixed block tree.
template <typename W>
__device__ __forceinline__ void restrict_block(int I, const int32_t* agg_start,
                                               const int32_t* agg_rows, const W* w,
                                               const double* wfac, double* rc) {
  __shared__ double red[kWaves];
  const int32_t mb = agg_start[I], me = agg_start[I + 1];
  constexpr int kInFlight = 8;
  double s = 0.0;
  for (int32_t m = mb + threadIdx.x; m < me; m += kInFlight * kBlock) {
    int32_t idx[kInFlight];
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      const int32_t mm = m + u * kBlock;
      idx[u] = mm < me ? agg_rows[mm] : -1;
    }
    double v[kInFlight], f[kInFlight];
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      v[u] = idx[u] >= 0 ? static_cast<double>(w[idx[u]]) : 0.0;
      f[u] = idx[u] >= 0 && wfac ? wfac[idx[u]] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      if (idx[u] < 0) continue;
      if (wfac) {
        s += v[u] * f[u];
      } else {
        s += v[u];
      }
    }
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) rc[I] = s;
}

// One block per aggregate.
template <typename W>
__global__ void __launch_bounds__(kBlock) restrict_kernel(int count, const int32_t* agg_start,
                                                          const int32_t* agg_rows,
                                                          const W* w, const double* wfac,
                                                          double* rc,
                                                          const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  (void)count;
  restrict_block<W>(blockIdx.x, agg_start, agg_rows, w, wfac, rc);
}

// The Schwarz block products of wa and the coarse restriction of wr in one
// launch: workgroups [0, product_groups) are the products' (b, part) groups,
// the rest one aggregate each.  Both read a finished vector and write disjoint
// buffers, so the two passes overlap instead of following each other.
template <typename T, typename V>
__global__ void __launch_bounds__(kBlock) products_restrict_kernel(
    unsigned product_groups, const int32_t* members, const int32_t* sizes, const V* inverse,
    const float* scale, const T* wa, const int32_t* entry_of, double* ybuf, size_t stride,
    const int32_t* agg_start,
    const int32_t* agg_rows, const T* wr, const double* wfac, double* rc, const int32_t* skip) {
  if (launch_skipped(skip)) return;
  if (blockIdx.x < product_groups) {  // uniform per workgroup
    schwarz_block_product<T, V>(blockIdx.x, members, sizes, inverse, scale, wa, entry_of, ybuf,
                                stride);
  } else {
    restrict_block<T>(static_cast<int>(blockIdx.x - product_groups), agg_start, agg_rows, wr,
                      wfac, rc);
  }
}
static_assert(kBlock == kMaxBlock, "one workgroup shape for the fused products and restriction");

// yc = Ainv rc, one block per row.
__global__ void __launch_bounds__(kBlock) coarse_product_kernel(int count, const double* ainv,
                                                                const double* rc,
                                                                const int32_t* status,
                                                                double* yc,
                                                                const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  __shared__ double red[kWaves];
  const int I = blockIdx.x;
  if (status[0]) {  // zero pivot in the inverse: no correction
    if (threadIdx.x == 0) yc[I] = 0.0;
    return;
  }
  const double* row = ainv + static_cast<size_t>(I) * count;
  double s = 0.0;
  for (int J = threadIdx.x; J < count; J += kBlock) s += row[J] * rc[J];
  s = block_sum(s, red);
  if (threadIdx.x == 0) yc[I] = s;
}

template <typename W>
__global__ void __launch_bounds__(kBlock) prolong_add_kernel(size_t n, const int32_t* agg_of_row,
                                                             const double* yc, W* v,
                                                             const int32_t* skip = nullptr) {
  if (launch_skipped(skip)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t a = agg_of_row[i];
  if (a >= 0) v[i] = static_cast<W>(static_cast<double>(v[i]) + yc[a]);
}

// gather_kernel (the Schwarz completion) and prolong_add_kernel fused: the
// same values and roundings, one pass over the rows instead of two launches.
template <typename T>
__global__ void __launch_bounds__(kBlock) complete_prolong_kernel(
    size_t n, const int32_t* cover_start, const double* ybuf, size_t ybuf_stride,
    const int32_t* agg_of_row, const double* yc, T* v, const int32_t* skip) {
  if (launch_skipped(skip)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool covered = cover_start[i + 1] > cover_start[i];
  const int32_t a = agg_of_row[i];
  if (!covered && a < 0) return;
  T vi = covered ? static_cast<T>(schwarz_row_mean(cover_start, ybuf, ybuf_stride,
                                                   static_cast<int32_t>(i)))
                 : v[i];
  if (a >= 0) vi = static_cast<T>(static_cast<double>(vi) + yc[a]);
  v[i] = vi;
}

unsigned wave_blocks(size_t items) { return static_cast<unsigned>((items + kWaves - 1) / kWaves); }

// The blocked Gauss-Jordan elimination of the count x count row-major matrix a
// in place (panel_rows: kMaxPanel x count doubles; perm: count; status[0] set
// on a zero pivot, after which the remaining launches return at once).
void gauss_jordan(double* a, int count, double* panel_rows, int32_t* perm, int32_t* status,
                  hipStream_t st) {
  const int b = panel_width(count);
  const int tiles = (count + 15) / 16;
  const unsigned rank_blocks =
      static_cast<unsigned>((static_cast<size_t>(tiles) * tiles + kWaves - 1) / kWaves);
  for (int k0 = 0; k0 < count; k0 += b) {
    const int w = std::min(b, count - k0);
    // (four waves with four rows a thread measured no faster at 729 rows)
    if (count <= kPanelThreads) {
      gj_panel_kernel<1, 16, kPanelThreads><<<1, kPanelThreads, 0, st>>>(a, count, k0, w, perm,
                                                                          status);
    } else if (count <= 2 * kPanelThreads) {
      gj_panel_kernel<2, 16, kPanelThreads><<<1, kPanelThreads, 0, st>>>(a, count, k0, w, perm,
                                                                          status);
    } else {
      gj_panel_kernel<4, 8, kPanelThreads><<<1, kPanelThreads, 0, st>>>(a, count, k0, w, perm,
                                                                         status);
    }
    if (count > w) {
      gj_swap_kernel<<<blocks_for(count), kBlock, 0, st>>>(a, count, k0, w, perm, status,
                                                          panel_rows);
      gj_rank_kernel<<<rank_blocks, kBlock, 0, st>>>(a, count, k0, w, panel_rows, status);
    }
  }
  LSMPS_LAUNCH_CHECK();
}

}  // namespace

void coarse_aggregate(Context& ctx, CoarseSpace& C, size_t n, const int32_t* labels,
                      const int32_t* cell_sorted, const int32_t* cell_begin,
                      const int32_t* cell_end, const int32_t dims[3], int factor) {
  C.count = 0;
  C.factor = 0;
  C.n = n;
  C.ready = false;
  C.tanks = 0;
  C.block_entries = 0;
  if (n == 0 || factor < 1 || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return;
  C.agg_of_row.reserve(n);
  const int32_t widest = std::max(dims[0], std::max(dims[1], dims[2]));
  for (int f = factor; f <= widest; ++f) {
    CoarseGrid g{dims[0], dims[1], dims[2], (dims[0] + f - 1) / f, (dims[1] + f - 1) / f,
                 (dims[2] + f - 1) / f, f};
    const size_t ccells = static_cast<size_t>(g.cx) * g.cy * g.cz;
    C.cell_count.reserve(ccells);
    C.cell_start.reserve(ccells);
    C.cell_flag.reserve(ccells);
    C.cell_id.reserve(ccells);
    count_members_kernel<<<wave_blocks(ccells), kBlock, 0, ctx.stream>>>(
        g, ccells, labels, cell_sorted, cell_begin, cell_end, C.cell_count.data(),
        C.cell_flag.data());
    LSMPS_LAUNCH_CHECK();
    exclusive_scan_i32(ctx, C.cell_count.data(), C.cell_start.data(), ccells);
    exclusive_scan_i32(ctx, C.cell_flag.data(), C.cell_id.data(), ccells);
    int32_t* h = ctx.host_i32;
    LSMPS_HIP(hipMemcpyAsync(h, C.cell_id.data() + ccells - 1, sizeof(int32_t),
                             hipMemcpyDeviceToHost, ctx.stream));
    LSMPS_HIP(hipMemcpyAsync(h + 1, C.cell_flag.data() + ccells - 1, sizeof(int32_t),
                             hipMemcpyDeviceToHost, ctx.stream));
    LSMPS_HIP(hipMemcpyAsync(h + 2, C.cell_start.data() + ccells - 1, sizeof(int32_t),
                             hipMemcpyDeviceToHost, ctx.stream));
    LSMPS_HIP(hipMemcpyAsync(h + 3, C.cell_count.data() + ccells - 1, sizeof(int32_t),
                             hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    const int count = h[0] + h[1];
    const int members = h[2] + h[3];
    if (count <= 1 || members <= 0) return;  // no useful coarse space
    if (count > kMaxCoarse) continue;        // too many aggregates: wider ones
    C.agg_start.reserve(static_cast<size_t>(count) + 1);
    C.agg_rows.reserve(static_cast<size_t>(members));
    fill_i32<<<blocks_for(n), kBlock, 0, ctx.stream>>>(C.agg_of_row.data(), n, -1);
    fill_members_kernel<<<wave_blocks(ccells), kBlock, 0, ctx.stream>>>(
        g, ccells, labels, cell_sorted, cell_begin, cell_end, C.cell_count.data(),
        C.cell_start.data(), C.cell_id.data(), C.agg_start.data(), C.agg_rows.data(),
        C.agg_of_row.data());
    set_last_kernel<<<1, 1, 0, ctx.stream>>>(C.agg_start.data(), count, C.cell_count.data(),
                                             C.cell_start.data(), ccells);
    LSMPS_LAUNCH_CHECK();
    C.count = count;
    C.factor = f;
    return;
  }
}

void coarse_aggregate_tanks(Context& ctx, CoarseSpace& C, size_t n, const int32_t* labels,
                            const int32_t* cell_sorted, const int32_t* cell_begin,
                            const int32_t* cell_end, const std::vector<CoarseTankGrid>& grids,
                            int factor) {
  C.count = 0;
  C.factor = 0;
  C.n = n;
  C.ready = false;
  C.tanks = 0;
  C.block_entries = 0;
  const int tanks = static_cast<int>(grids.size());
  if (n == 0 || factor < 1 || tanks == 0) return;
  int32_t widest = 1;
  for (const CoarseTankGrid& g : grids) {
    require(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.base >= 0, "Invalid cell grid configuration");
    widest = std::max(widest, std::max(g.nx, std::max(g.ny, g.nz)));
  }
  C.agg_of_row.reserve(n);
  C.tank_grids.reserve(static_cast<size_t>(tanks));
  C.tank_cc.reserve(static_cast<size_t>(tanks) + 1);
  C.tank_agg.reserve(static_cast<size_t>(tanks) + 2);
  C.block_offset.reserve(static_cast<size_t>(tanks) + 1);
  C.grids_host.resize(tanks);
  C.tank_cc_host.resize(static_cast<size_t>(tanks) + 1);
  C.tank_agg_host.resize(static_cast<size_t>(tanks) + 2);
  for (int f = factor; f <= widest; ++f) {
    int64_t ccells = 0;
    for (int k = 0; k < tanks; ++k) {
      const CoarseTankGrid& g = grids[k];
      TankCoarseGrid& t = C.grids_host[k];
      t = TankCoarseGrid{g.nx, g.ny, g.nz, (g.nx + f - 1) / f, (g.ny + f - 1) / f,
                         (g.nz + f - 1) / f, f, g.base, static_cast<int32_t>(ccells)};
      C.tank_cc_host[k] = static_cast<int32_t>(ccells);
      ccells += static_cast<int64_t>(t.cx) * t.cy * t.cz;
      require(ccells < INT32_MAX, "Cell count exceeds 32-bit index capacity");
    }
    C.tank_cc_host[tanks] = static_cast<int32_t>(ccells);
    const size_t nc = static_cast<size_t>(ccells);
    C.cell_count.reserve(nc);
    C.cell_start.reserve(nc);
    C.cell_flag.reserve(nc);
    C.cell_id.reserve(nc);
    LSMPS_HIP(hipMemcpyAsync(C.tank_grids.data(), C.grids_host.data(),
                             tanks * sizeof(TankCoarseGrid), hipMemcpyHostToDevice, ctx.stream));
    LSMPS_HIP(hipMemcpyAsync(C.tank_cc.data(), C.tank_cc_host.data(), (tanks + 1) * sizeof(int32_t),
                             hipMemcpyHostToDevice, ctx.stream));
    count_tank_members_kernel<<<wave_blocks(nc), kBlock, 0, ctx.stream>>>(
        C.tank_grids.data(), C.tank_cc.data(), tanks, nc, labels, cell_sorted, cell_begin,
        cell_end, C.cell_count.data(), C.cell_flag.data());
    LSMPS_LAUNCH_CHECK();
    exclusive_scan_i32(ctx, C.cell_count.data(), C.cell_start.data(), nc);
    exclusive_scan_i32(ctx, C.cell_flag.data(), C.cell_id.data(), nc);
    tank_ranges_kernel<<<blocks_for(static_cast<size_t>(tanks) + 2), kBlock, 0, ctx.stream>>>(
        C.tank_cc.data(), tanks, nc, C.cell_count.data(), C.cell_flag.data(),
        C.cell_start.data(), C.cell_id.data(), C.tank_agg.data());
    LSMPS_LAUNCH_CHECK();
    LSMPS_HIP(hipMemcpyAsync(C.tank_agg_host.data(), C.tank_agg.data(),
                             (tanks + 2) * sizeof(int32_t), hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    const int count = C.tank_agg_host[tanks];
    const int members = C.tank_agg_host[tanks + 1];
    if (count <= 0 || members <= 0) return;  // no aggregated rows
    int widest_block = 0;
    for (int k = 0; k < tanks; ++k)
      widest_block = std::max(widest_block, C.tank_agg_host[k + 1] - C.tank_agg_host[k]);
    if (widest_block > kMaxTankCoarse) continue;  // wider aggregates
    C.agg_start.reserve(static_cast<size_t>(count) + 1);
    C.agg_rows.reserve(static_cast<size_t>(members));
    C.agg_tank.reserve(static_cast<size_t>(count));
    fill_i32<<<blocks_for(n), kBlock, 0, ctx.stream>>>(C.agg_of_row.data(), n, -1);
    fill_tank_members_kernel<<<wave_blocks(nc), kBlock, 0, ctx.stream>>>(
        C.tank_grids.data(), C.tank_cc.data(), tanks, nc, labels, cell_sorted, cell_begin,
        cell_end, C.cell_count.data(), C.cell_start.data(), C.cell_id.data(), C.agg_start.data(),
        C.agg_rows.data(), C.agg_of_row.data());
    set_last_kernel<<<1, 1, 0, ctx.stream>>>(C.agg_start.data(), count, C.cell_count.data(),
                                             C.cell_start.data(), nc);
    agg_tank_kernel<<<static_cast<unsigned>(tanks), kBlock, 0, ctx.stream>>>(C.tank_agg.data(),
                                                                            C.agg_tank.data());
    LSMPS_LAUNCH_CHECK();
    C.offsets_host.resize(static_cast<size_t>(tanks) + 1);
    int64_t entries = 0;
    for (int k = 0; k < tanks; ++k) {
      C.offsets_host[k] = entries;
      const int64_t c = C.tank_agg_host[k + 1] - C.tank_agg_host[k];
      entries += c * c;
    }
    C.offsets_host[tanks] = entries;
    LSMPS_HIP(hipMemcpyAsync(C.block_offset.data(), C.offsets_host.data(),
                             (tanks + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx.stream));
    ctx.sync();  // the host sources of the uploads stay valid until here anyway
    C.count = count;
    C.factor = f;
    C.tanks = tanks;
    C.block_entries = static_cast<size_t>(entries);
    return;
  }
}

void coarse_reserve(CoarseSpace& C) {
  const int count = C.count;
  if (C.tanks > 0) {
    if (count < 1) return;
    C.work.reserve(std::max<size_t>(C.block_entries, 1));
    C.ainv.reserve(std::max<size_t>(C.block_entries, 1));
    C.status.reserve(static_cast<size_t>(C.tanks));
    C.rc.reserve(static_cast<size_t>(count));
    C.yc.reserve(static_cast<size_t>(count));
    C.rowabs.reserve(C.n);
    C.wfac.reserve(C.n);
    return;
  }
  if (count <= 1) return;
  const size_t nn = static_cast<size_t>(count) * count;
  C.work.reserve(nn);
  C.ainv.reserve(nn);
  C.colk.reserve(static_cast<size_t>(kMaxPanel) * count + kMaxPanel);
  C.perm.reserve(static_cast<size_t>(count));
  C.status.reserve(1);
  C.rc.reserve(static_cast<size_t>(count));
  C.yc.reserve(static_cast<size_t>(count));
  C.rowabs.reserve(C.n);
  C.wfac.reserve(C.n);
}

void coarse_setup(Context& ctx, CoarseSpace& C, const lsmps_csr& A, const double* diag,
                  const double* offabs, hipStream_t st) {
  (void)ctx;
  C.ready = false;
  const int count = C.count;
  if (C.tanks > 0) {
    if (count < 1 || C.n != A.rows) return;
    require(C.work.size() >= C.block_entries && C.ainv.size() >= C.block_entries &&
                C.rc.size() >= static_cast<size_t>(count) && C.rowabs.size() >= A.rows &&
                C.status.size() >= static_cast<size_t>(C.tanks),
            "coarse setup before coarse_reserve", LSMPS_ERR_INTERNAL);
    const unsigned tb = static_cast<unsigned>(C.tanks);
    fill_i32<<<blocks_for(C.tanks), kBlock, 0, st>>>(C.status.data(), C.tanks, 0);
    unsigned long long* acci = reinterpret_cast<unsigned long long*>(C.ainv.data());
    row_bound_kernel<<<blocks_for(A.rows), kBlock, 0, st>>>(
        A.rows, C.agg_of_row.data(), diag, offabs, C.rowabs.data());
    restrict_kernel<double><<<static_cast<unsigned>(count), kBlock, 0, st>>>(
        count, C.agg_start.data(), C.agg_rows.data(), C.rowabs.data(), nullptr, C.rc.data());
    fill_u64<<<blocks_for(C.block_entries), kBlock, 0, st>>>(acci, C.block_entries);
    tank_matrix_kernel<<<static_cast<unsigned>(count) * kMatrixParts, kBlock, 0, st>>>(
        C.agg_tank.data(), C.tank_agg.data(), C.block_offset.data(), C.agg_start.data(),
        C.agg_rows.data(), C.agg_of_row.data(), A.row_offsets, A.col_indices, A.values,
        C.rc.data(), acci);
    tank_convert_kernel<<<static_cast<unsigned>(count), kBlock, 0, st>>>(
        C.agg_tank.data(), C.tank_agg.data(), C.block_offset.data(), C.rc.data(), acci,
        C.work.data());
    tank_gj_kernel<<<tb, kBlock, 0, st>>>(C.tank_agg.data(), C.block_offset.data(),
                                          C.work.data(), C.ainv.data(), C.status.data());
    LSMPS_LAUNCH_CHECK();
    C.ready = true;  // singular blocks are flagged per tank in C.status
    return;
  }
  if (count <= 1 || C.n != A.rows) return;
  const size_t nn = static_cast<size_t>(count) * count;
  require(C.work.size() >= nn && C.ainv.size() >= nn && C.rc.size() >= static_cast<size_t>(count) &&
              C.rowabs.size() >= A.rows,
          "coarse setup before coarse_reserve", LSMPS_ERR_INTERNAL);
  fill_i32<<<1, 1, 0, st>>>(C.status.data(), 1, 0);
  // integer matrix in ainv's storage (overwritten by the inverse at the end)
  unsigned long long* acci = reinterpret_cast<unsigned long long*>(C.ainv.data());
  row_bound_kernel<<<blocks_for(A.rows), kBlock, 0, st>>>(
      A.rows, C.agg_of_row.data(), diag, offabs, C.rowabs.data());
  restrict_kernel<double><<<static_cast<unsigned>(count), kBlock, 0, st>>>(
      count, C.agg_start.data(), C.agg_rows.data(), C.rowabs.data(), nullptr, C.rc.data());
  fill_u64<<<blocks_for(nn), kBlock, 0, st>>>(acci, nn);
  coarse_matrix_kernel<<<static_cast<unsigned>(count) * kMatrixParts, kBlock, 0, st>>>(
      count, C.agg_start.data(), C.agg_rows.data(), C.agg_of_row.data(), A.row_offsets,
      A.col_indices, A.values, C.rc.data(), acci);
  coarse_convert_kernel<<<blocks_for(nn), kBlock, 0, st>>>(count, C.rc.data(), acci,
                                                                   C.work.data());
  LSMPS_LAUNCH_CHECK();
  const unsigned eb = blocks_for(nn);
  gauss_jordan(C.work.data(), count, C.colk.data(), C.perm.data(), C.status.data(), st);
  LSMPS_LAUNCH_CHECK();
  // idx reuses colk's storage as int32 (count entries fit in count doubles)
  int32_t* idx = reinterpret_cast<int32_t*>(C.colk.data());
  gj_order_kernel<<<1, kBlock, 0, st>>>(count, C.perm.data(), C.status.data(), idx);
  gj_permute_kernel<<<eb, kBlock, 0, st>>>(C.work.data(), count, idx, C.ainv.data());
  LSMPS_LAUNCH_CHECK();
  C.ready = true;  // a zero pivot is reported in C.status (the product then yields 0)
}

void dense_inverse(Context& ctx, int n, const double* matrix, double* inverse) {
  require(n >= 1 && n <= kMaxCoarse, "dense inverse size must be in 1..4096");
  require(matrix && inverse, "dense inverse needs device matrix and output pointers");
  const size_t nn = static_cast<size_t>(n) * n;
  DeviceArray<double> work, panel_rows;
  DeviceArray<int32_t> perm, status;
  work.reserve(nn);
  panel_rows.reserve(static_cast<size_t>(kMaxPanel) * n + kMaxPanel);
  perm.reserve(static_cast<size_t>(n));
  status.reserve(1);
  LSMPS_HIP(hipMemcpyAsync(work.data(), matrix, nn * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx.stream));
  fill_i32<<<1, 1, 0, ctx.stream>>>(status.data(), 1, 0);
  gauss_jordan(work.data(), n, panel_rows.data(), perm.data(), status.data(), ctx.stream);
  int32_t* idx = reinterpret_cast<int32_t*>(panel_rows.data());
  gj_order_kernel<<<1, kBlock, 0, ctx.stream>>>(n, perm.data(), status.data(), idx);
  gj_permute_kernel<<<blocks_for(nn), kBlock, 0, ctx.stream>>>(work.data(), n, idx, inverse);
  LSMPS_LAUNCH_CHECK();
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, status.data(), sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx.stream));
  ctx.sync();
  if (ctx.host_i32[0]) raise(LSMPS_ERR_NUMERICAL, "dense inverse: zero pivot");
}

void coarse_set_weights(Context& ctx, CoarseSpace& C, const double* pdiag) {
  require(C.wfac.size() >= C.n, "coarse weights before coarse_reserve", LSMPS_ERR_INTERNAL);
  inverse_kernel<<<blocks_for(C.n), kBlock, 0, ctx.stream>>>(C.n, pdiag, C.wfac.data());
  LSMPS_LAUNCH_CHECK();
}

bool coarse_usable(Context& ctx, CoarseSpace& C) {
  if (!C.ready) return false;
  if (C.tanks > 0) {
    C.status_host.resize(static_cast<size_t>(C.tanks));
    LSMPS_HIP(hipMemcpyAsync(C.status_host.data(), C.status.data(), C.tanks * sizeof(int32_t),
                             hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    for (int k = 0; k < C.tanks; ++k)
      if (!C.status_host[k] && C.tank_agg_host[k + 1] > C.tank_agg_host[k]) return true;
    return false;
  }
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, C.status.data(), sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx.stream));
  ctx.sync();
  return ctx.host_i32[0] == 0;
}

template <typename W>
void coarse_solve(Context& ctx, CoarseSpace& C, const W* w, bool scale_by_wfac,
                  const int32_t* skip) {
  restrict_kernel<W><<<static_cast<unsigned>(C.count), kBlock, 0, ctx.stream>>>(
      C.count, C.agg_start.data(), C.agg_rows.data(), w, scale_by_wfac ? C.wfac.data() : nullptr,
      C.rc.data(), skip);
  if (C.tanks > 0) {
    tank_product_kernel<<<wave_blocks(static_cast<size_t>(C.count)), kBlock, 0, ctx.stream>>>(
        C.count, C.agg_tank.data(), C.tank_agg.data(), C.block_offset.data(), C.ainv.data(),
        C.rc.data(), C.status.data(), C.yc.data(), skip);
    LSMPS_LAUNCH_CHECK();
    return;
  }
  coarse_product_kernel<<<static_cast<unsigned>(C.count), kBlock, 0, ctx.stream>>>(
      C.count, C.ainv.data(), C.rc.data(), C.status.data(), C.yc.data(), skip);
  LSMPS_LAUNCH_CHECK();
}

void coarse_product(Context& ctx, CoarseSpace& C, const int32_t* skip) {
  if (C.tanks > 0) {
    tank_product_kernel<<<wave_blocks(static_cast<size_t>(C.count)), kBlock, 0, ctx.stream>>>(
        C.count, C.agg_tank.data(), C.tank_agg.data(), C.block_offset.data(), C.ainv.data(),
        C.rc.data(), C.status.data(), C.yc.data(), skip);
  } else {
    coarse_product_kernel<<<static_cast<unsigned>(C.count), kBlock, 0, ctx.stream>>>(
        C.count, C.ainv.data(), C.rc.data(), C.status.data(), C.yc.data(), skip);
  }
  LSMPS_LAUNCH_CHECK();
}

template <typename T>
void products_and_restrict(Context& ctx, const SchwarzPrec& P, CoarseSpace& C, const T* wa,
                           const T* wr, bool scale_by_wfac, const int32_t* skip) {
  const unsigned groups = static_cast<unsigned>(P.blocks * kApplyParts);
  const unsigned grid = groups + static_cast<unsigned>(C.count);
  const double* wfac = scale_by_wfac ? C.wfac.data() : nullptr;
  if (P.half) {
    products_restrict_kernel<T, _Float16><<<grid, kBlock, 0, ctx.stream>>>(
        groups, P.members.data(), P.sizes.data(),
        reinterpret_cast<const _Float16*>(P.inv16.data()), P.col_scale.data(), wa,
        P.entry_of.data(), P.ybuf.data(), P.ybuf_stride,
        C.agg_start.data(), C.agg_rows.data(), wr, wfac, C.rc.data(), skip);
  } else if (P.fp32) {
    products_restrict_kernel<T, float><<<grid, kBlock, 0, ctx.stream>>>(
        groups, P.members.data(), P.sizes.data(), P.inv32.data(), nullptr, wa, P.entry_of.data(),
        P.ybuf.data(), P.ybuf_stride,
        C.agg_start.data(), C.agg_rows.data(), wr, wfac, C.rc.data(), skip);
  } else {
    products_restrict_kernel<T, double><<<grid, kBlock, 0, ctx.stream>>>(
        groups, P.members.data(), P.sizes.data(), P.inv64.data(), nullptr, wa, P.entry_of.data(),
        P.ybuf.data(), P.ybuf_stride,
        C.agg_start.data(), C.agg_rows.data(), wr, wfac, C.rc.data(), skip);
  }
  LSMPS_LAUNCH_CHECK();
}

template <typename W>
void coarse_prolong_add(Context& ctx, const CoarseSpace& C, W* v, const int32_t* skip) {
  prolong_add_kernel<W><<<blocks_for(C.n), kBlock, 0, ctx.stream>>>(C.n, C.agg_of_row.data(),
                                                                    C.yc.data(), v, skip);
  LSMPS_LAUNCH_CHECK();
}

template <typename T>
void schwarz_complete_prolong_add(Context& ctx, const SchwarzPrec& P, const CoarseSpace& C, T* v,
                                  const int32_t* skip) {
  require(P.n == C.n, "Schwarz blocks and coarse space cover different systems");
  if (C.n == 0) return;
  complete_prolong_kernel<T><<<blocks_for(C.n), kBlock, 0, ctx.stream>>>(
      C.n, P.cover_start.data(), P.ybuf.data(), P.ybuf_stride, C.agg_of_row.data(), C.yc.data(),
      v, skip);
  LSMPS_LAUNCH_CHECK();
}

template void schwarz_complete_prolong_add<double>(Context&, const SchwarzPrec&,
                                                   const CoarseSpace&, double*, const int32_t*);
template void products_and_restrict<double>(Context&, const SchwarzPrec&, CoarseSpace&,
                                            const double*, const double*, bool, const int32_t*);
template void coarse_solve<double>(Context&, CoarseSpace&, const double*, bool, const int32_t*);
template void coarse_solve<float>(Context&, CoarseSpace&, const float*, bool, const int32_t*);
template void coarse_prolong_add<double>(Context&, const CoarseSpace&, double*, const int32_t*);
template void coarse_prolong_add<float>(Context&, const CoarseSpace&, float*, const int32_t*);

}  // namespace lsmps
