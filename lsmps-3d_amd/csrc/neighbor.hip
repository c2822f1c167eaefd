// Neighbor search on gfx950: cell list + fluid-fluid / fluid-wall CSR tables.
//
// Replaces build_cell_list / build_neighbor_lists (src/neighbor/neighbor_search.cu:255-367).
// Output is bit-identical to a stable sort of the cell ids and a 27-cell scan in
// (dz, dy, dx) order, within a cell in ascending particle index:
//   1. cell id per particle (FP64, same operations as clamp_cell_index/flatten_cell),
//      per-cell counts with atomics;
//   2. device-wide exclusive scan -> cell starts;
//   3. scatter with per-cell atomic cursors, then each cell's slice is sorted
//      ascending (the slot order from atomics is arbitrary, the sorted slice is
//      what a stable sort produces because input order = index order);
//   4. sorted source points packed as 32-byte {x, y, z, index} records so a
//      candidate costs one contiguous load;
//   5. count pass, scan, capacity check, write pass.  Rows are visited in the
//      fluid cell order so a wavefront's 64 rows share the same neighbor cells.
// Compiled with -ffp-contract=off: the distance test must round exactly like
// the oracle's.
#include <type_traits>

#include "common.hpp"
#include "stages.hpp"
#include "tune.hpp"

namespace lsmps {
namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048

struct alignas(32) SortedPoint {
  double x, y, z;
  int32_t index;
  int32_t pad;
};
static_assert(sizeof(SortedPoint) == kSortedPointBytes, "sorted point record size");

struct GridDev {
  double ox, oy, oz, h;
  int32_t nx, ny, nz;
};

__global__ void fill_i32_kernel(int32_t* p, size_t n, int32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_f64_kernel(double* p, size_t n, double v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_f32_kernel(float* p, size_t n, float v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void copy_f64_kernel(double* dst, const double* src, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---- block scan helpers (64-lane wavefronts, 256-thread blocks) ----------
__device__ __forceinline__ int32_t wave_inclusive_scan(int32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan over the block; returns the prefix of this thread, total in `total`.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t& total) {
  __shared__ int32_t wave_sums[kBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t inc = wave_inclusive_scan(v);
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  int32_t offset = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const int32_t s = wave_sums[w];
    offset += (w < wave) ? s : 0;
    sum += s;
  }
  __syncthreads();
  total = sum;
  return inc - v + offset;
}

__global__ void __launch_bounds__(kBlock) scan_reduce_kernel(const int32_t* in, size_t n,
                                                             int32_t* partials) {
  const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int32_t total;
  (void)block_exclusive_scan(s, total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Scans one tile per block, adding the block's carry from `carries` (or 0).
__global__ void __launch_bounds__(kBlock) scan_tile_kernel(const int32_t* in, int32_t* out,
                                                           size_t n, const int32_t* carries,
                                                           int write_total) {
  __shared__ int32_t tile[kScanTile];
  const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
    tile[k * kBlock + threadIdx.x] = idx < n ? in[idx] : 0;
  }
  __syncthreads();
  int32_t local[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    local[k] = tile[threadIdx.x * kScanItems + k];
    s += local[k];
  }
  int32_t total;
  int32_t prefix = block_exclusive_scan(s, total);
  const int32_t carry = carries ? carries[blockIdx.x] : 0;
  prefix += carry;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    tile[threadIdx.x * kScanItems + k] = prefix;
    prefix += local[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
    if (idx < n) out[idx] = tile[k * kBlock + threadIdx.x];
  }
  if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) {
    out[n] = carry + total;
  }
}

// ---- cell list ------------------------------------------------------------

// clamp_cell_index (neighbor_search.cu:36-46)
__device__ __forceinline__ int32_t clamp_cell(double coordinate, double origin, double h,
                                              int32_t dim) {
  const double scaled = (coordinate - origin) / h;
  const int32_t index = static_cast<int32_t>(floor(scaled));
  if (index < 0) return 0;
  if (index >= dim) return dim - 1;
  return index;
}

// flatten_cell (neighbor_search.cu:32-34)
__device__ __forceinline__ int32_t flatten(int32_t ix, int32_t iy, int32_t iz, const GridDev& g) {
  return (iz * g.ny + iy) * g.nx + ix;
}

__device__ __forceinline__ GridDev seg_grid(const SegmentGrid& s) {
  GridDev g;
  g.ox = s.ox;
  g.oy = s.oy;
  g.oz = s.oz;
  g.h = s.h;
  g.nx = s.nx;
  g.ny = s.ny;
  g.nz = s.nz;
  return g;
}

// assign_cells_kernel (neighbor_search.cu:61-73) + per-cell counting.  With
// segments, particle i uses the grid of its tank seg[i] (same arithmetic as a
// single-tank run, cell ids offset by the tank's base).
template <bool kSeg>
__global__ void assign_cells_kernel(size_t n, const double* __restrict__ x,
                                    const double* __restrict__ y, const double* __restrict__ z,
                                    GridDev g0, const SegmentGrid* __restrict__ grids,
                                    const int32_t* __restrict__ seg,
                                    int32_t* __restrict__ cell_of,
                                    int32_t* __restrict__ counts) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  GridDev g = g0;
  int32_t base = 0;
  if (kSeg) {
    const SegmentGrid sg = grids[seg[i]];
    g = seg_grid(sg);
    base = sg.base;
  }
  const int32_t ix = clamp_cell(x[i], g.ox, g.h, g.nx);
  const int32_t iy = clamp_cell(y[i], g.oy, g.h, g.ny);
  const int32_t iz = clamp_cell(z[i], g.oz, g.h, g.nz);
  const int32_t c = base + flatten(ix, iy, iz, g);
  cell_of[i] = c;
  atomicAdd(&counts[c], 1);
}

__global__ void scatter_cells_kernel(size_t n, const int32_t* __restrict__ cell_of,
                                     const int32_t* __restrict__ starts,
                                     int32_t* __restrict__ cursor,
                                     int32_t* __restrict__ sorted) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = cell_of[i];
  const int32_t slot = atomicAdd(&cursor[c], 1);
  sorted[starts[c] + slot] = static_cast<int32_t>(i);
}

// Sorts each cell's slice ascending and writes cell_ids / cell_begin / cell_end
// (build_cell_ranges_kernel, neighbor_search.cu:85-98).
__global__ void finalize_cells_kernel(size_t cells, const int32_t* __restrict__ starts,
                                      int32_t* __restrict__ sorted,
                                      int32_t* __restrict__ cell_ids,
                                      int32_t* __restrict__ begin, int32_t* __restrict__ end) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int32_t b = starts[c];
  const int32_t e = starts[c + 1];
  if (b == e) return;  // ranges were filled with -1
  for (int32_t p = b + 1; p < e; ++p) {
    const int32_t v = sorted[p];
    int32_t q = p - 1;
    while (q >= b && sorted[q] > v) {
      sorted[q + 1] = sorted[q];
      --q;
    }
    sorted[q + 1] = v;
  }
  for (int32_t p = b; p < e; ++p) cell_ids[p] = static_cast<int32_t>(c);
  begin[c] = b;
  end[c] = e;
}

__global__ void gather_points_kernel(size_t n, const int32_t* __restrict__ sorted,
                                     const double* __restrict__ x, const double* __restrict__ y,
                                     const double* __restrict__ z,
                                     SortedPoint* __restrict__ points) {
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t j = sorted[p];
  SortedPoint sp;
  sp.x = x[j];
  sp.y = y[j];
  sp.z = z[j];
  sp.index = j;
  sp.pad = 0;
  points[p] = sp;
}

// count_neighbors_kernel / write_neighbors_kernel (neighbor_search.cu:221-253) with
// count/write_source_neighbors_for_fluid (:100-219) inlined.  Row t of the launch
// handles fluid particle order[t] (cell order) so neighboring rows share cells.
// kScanLanes lanes per particle (a power of two dividing 64, picked by the
// autotuner; by default 8 for the fill pass and 1 for the count pass, which
// writes nothing per hit).  The table is the same for every choice.
template <bool kWrite, bool kSeg, int kScanLanes>
__global__ void __launch_bounds__(kBlock) neighbor_scan_kernel(
    size_t n_fluid, const int32_t* __restrict__ order, const double* __restrict__ fx,
    const double* __restrict__ fy, const double* __restrict__ fz, GridDev g0,
    const SegmentGrid* __restrict__ grids, const int32_t* __restrict__ fseg,
    const int32_t* __restrict__ src_begin, const int32_t* __restrict__ src_end,
    const SortedPoint* __restrict__ src, double radius_squared, int exclude_self,
    int32_t* __restrict__ offsets, int32_t* __restrict__ indices) {
  constexpr uint32_t kScanMask = (1u << kScanLanes) - 1u;
  const size_t gt = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t t = gt / kScanLanes;
  const int sub = static_cast<int>(gt % kScanLanes);
  if (t >= n_fluid) return;  // whole groups
  const int group_lane = static_cast<int>(threadIdx.x & 63) - sub;  // first lane of the group
  const int32_t i = order ? order[t] : static_cast<int32_t>(t);
  GridDev g = g0;
  int32_t base = 0;
  if (kSeg) {  // the tank's own grid: sources of other tanks are never visited
    const SegmentGrid sg = grids[fseg[i]];
    g = seg_grid(sg);
    base = sg.base;
  }
  const double px = fx[i];
  const double py = fy[i];
  const double pz = fz[i];
  const int32_t cx = clamp_cell(px, g.ox, g.h, g.nx);
  const int32_t cy = clamp_cell(py, g.oy, g.h, g.ny);
  const int32_t cz = clamp_cell(pz, g.oz, g.h, g.nz);
  int32_t count = 0;
  int32_t out = kWrite ? offsets[i] : 0;
  for (int32_t dz = -1; dz <= 1; ++dz) {
    const int32_t nz = cz + dz;
    if (nz < 0 || nz >= g.nz) continue;
    for (int32_t dy = -1; dy <= 1; ++dy) {
      const int32_t ny = cy + dy;
      if (ny < 0 || ny >= g.ny) continue;
      for (int32_t dx = -1; dx <= 1; ++dx) {
        const int32_t nx = cx + dx;
        if (nx < 0 || nx >= g.nx) continue;
        const int32_t cell = base + flatten(nx, ny, nz, g);
        const int32_t b = src_begin[cell];
        const int32_t e = src_end[cell];
        if (b < 0 || e < 0) continue;
        // the group's lanes test points p0 + sub; hits are compacted in point
        // order, so the list is the one a sequential scan writes
        for (int32_t p0 = b; p0 < e; p0 += kScanLanes) {
          const int32_t p = p0 + sub;
          bool hit = false;
          int32_t qi = 0;
          if (p < e) {
            const SortedPoint q = src[p];
            qi = q.index;
            // within_radius (neighbor_search.cu:48-59)
            const double ddx = px - q.x;
            const double ddy = py - q.y;
            const double ddz = pz - q.z;
            hit = !(exclude_self && q.index == i) &&
                  ddx * ddx + ddy * ddy + ddz * ddz <= radius_squared;
          }
          const unsigned long long bits = __ballot(hit);
          const uint32_t mine = static_cast<uint32_t>(bits >> group_lane) & kScanMask;
          if (kWrite && hit) indices[out + __popc(mine & ((1u << sub) - 1u))] = qi;
          const int32_t found = __popc(mine);
          out += found;
          count += found;
        }
      }
    }
  }
  if (!kWrite && sub == 0) offsets[i] = count;
}

GridDev to_dev(const lsmps_cell_grid& g) {
  GridDev d;
  d.ox = g.origin[0];
  d.oy = g.origin[1];
  d.oz = g.origin[2];
  d.h = g.cell_size;
  d.nx = g.dims[0];
  d.ny = g.dims[1];
  d.nz = g.dims[2];
  return d;
}

}  // namespace

void fill_i32(hipStream_t s, int32_t* p, size_t n, int32_t v) {
  if (n == 0) return;
  fill_i32_kernel<<<blocks_for(n), kBlock, 0, s>>>(p, n, v);
  LSMPS_LAUNCH_CHECK();
}
void fill_f64(hipStream_t s, double* p, size_t n, double v) {
  if (n == 0) return;
  fill_f64_kernel<<<blocks_for(n), kBlock, 0, s>>>(p, n, v);
  LSMPS_LAUNCH_CHECK();
}
void fill_f32(hipStream_t s, float* p, size_t n, float v) {
  if (n == 0) return;
  fill_f32_kernel<<<blocks_for(n), kBlock, 0, s>>>(p, n, v);
  LSMPS_LAUNCH_CHECK();
}
void copy_f64(hipStream_t s, double* dst, const double* src, size_t n) {
  if (n == 0) return;
  copy_f64_kernel<<<blocks_for(n), kBlock, 0, s>>>(dst, src, n);
  LSMPS_LAUNCH_CHECK();
}

namespace {
// Recursive three-phase scan: tile sums, scan of the sums, tile scans with carries.
void scan_level(Context& ctx, const int32_t* in, int32_t* out, size_t n, int32_t* scratch,
                bool write_total) {
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    scan_tile_kernel<<<1, kBlock, 0, ctx.stream>>>(in, out, n, nullptr, write_total ? 1 : 0);
    LSMPS_LAUNCH_CHECK();
    return;
  }
  int32_t* partials = scratch;
  int32_t* next = scratch + tiles + 1;
  scan_reduce_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, ctx.stream>>>(in, n, partials);
  LSMPS_LAUNCH_CHECK();
  scan_level(ctx, partials, partials, tiles, next, false);
  scan_tile_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, ctx.stream>>>(
      in, out, n, partials, write_total ? 1 : 0);
  LSMPS_LAUNCH_CHECK();
}
}  // namespace

size_t scan_partials_needed(size_t n) {
  size_t need = 16;
  for (size_t m = n; m > 1;) {
    const size_t tiles = (m + kScanTile - 1) / kScanTile;
    need += tiles + 1;
    if (tiles <= 1) break;
    m = tiles;
  }
  return need;
}

void exclusive_scan_i32_with(Context& ctx, const int32_t* in, int32_t* out, size_t n,
                             int32_t* partials, size_t partials_capacity) {
  if (n == 0) {
    fill_i32(ctx.stream, out, 1, 0);
    return;
  }
  require(scan_partials_needed(n) <= partials_capacity, "scan partials buffer too small",
          LSMPS_ERR_INTERNAL);
  scan_level(ctx, in, out, n, partials, true);
}

void exclusive_scan_i32(Context& ctx, const int32_t* in, int32_t* out, size_t n) {
  // one allocation of 2^20 partials covers every n below 2^31 (need ~ n / 2047),
  // so the buffer never grows between scans already enqueued
  ctx.scan_partials.reserve(std::max<size_t>(scan_partials_needed(n), size_t(1) << 20));
  exclusive_scan_i32_with(ctx, in, out, n, ctx.scan_partials.data(), ctx.scan_partials.size());
}

GeometryScratch context_geometry_scratch(Context& ctx, size_t cells, size_t fluid_count,
                                         size_t wall_count) {
  const size_t most = std::max<size_t>(std::max(fluid_count, wall_count), 1);
  ctx.tmp_counts.reserve(cells + 1);
  ctx.tmp_cursor.reserve(cells + 1);
  ctx.tmp_labels.reserve(most);
  ctx.tmp_points.reserve(most * kSortedPointBytes);
  ctx.scan_partials.reserve(std::max<size_t>(scan_partials_needed(std::max(most, cells)),
                                             size_t(1) << 20));
  GeometryScratch g;
  g.counts = ctx.tmp_counts.data();
  g.cursor = ctx.tmp_cursor.data();
  g.cells_capacity = ctx.tmp_counts.size();
  g.labels = ctx.tmp_labels.data();
  g.particles_capacity = std::min(ctx.tmp_labels.size(), ctx.tmp_points.size() / kSortedPointBytes);
  g.points = ctx.tmp_points.data();
  g.partials = ctx.scan_partials.data();
  g.partials_capacity = ctx.scan_partials.size();
  return g;
}

namespace {
// Counting sort of `count` particles into `cc` cells (one grid, or per-particle
// segment grids when `grids` is set).
void cell_list_impl(Context& ctx, const GeometryScratch& sc, size_t count, const double* x,
                    const double* y, const double* z, const GridDev& g,
                    const SegmentGrid* grids, const int32_t* seg, size_t cc,
                    const lsmps_cell_list& cells) {
  fill_i32(ctx.stream, cells.cell_begin, cc, -1);
  fill_i32(ctx.stream, cells.cell_end, cc, -1);
  if (count == 0) return;
  require(x && y && z && cells.cell_ids && cells.sorted_indices && cells.cell_begin && cells.cell_end,
          "Cell list requires position and cell buffers");
  require(sc.cells_capacity >= cc + 1 && sc.particles_capacity >= count,
          "cell list scratch too small", LSMPS_ERR_INTERNAL);
  int32_t* counts = sc.counts;
  int32_t* starts = sc.cursor;
  int32_t* cell_of = sc.labels;
  fill_i32(ctx.stream, counts, cc, 0);
  if (grids) {
    assign_cells_kernel<true><<<blocks_for(count), kBlock, 0, ctx.stream>>>(
        count, x, y, z, g, grids, seg, cell_of, counts);
  } else {
    assign_cells_kernel<false><<<blocks_for(count), kBlock, 0, ctx.stream>>>(
        count, x, y, z, g, nullptr, nullptr, cell_of, counts);
  }
  LSMPS_LAUNCH_CHECK();
  exclusive_scan_i32_with(ctx, counts, starts, cc, sc.partials, sc.partials_capacity);
  fill_i32(ctx.stream, counts, cc, 0);
  scatter_cells_kernel<<<blocks_for(count), kBlock, 0, ctx.stream>>>(count, cell_of, starts,
                                                                     counts, cells.sorted_indices);
  LSMPS_LAUNCH_CHECK();
  finalize_cells_kernel<<<blocks_for(cc), kBlock, 0, ctx.stream>>>(
      cc, starts, cells.sorted_indices, cells.cell_ids, cells.cell_begin, cells.cell_end);
  LSMPS_LAUNCH_CHECK();
}
}  // namespace

namespace {
void check_cell_list(size_t count, const lsmps_cell_grid& grid, const lsmps_cell_list& cells) {
  // validation messages follow build_cell_list_impl (neighbor_search.cu:257-273)
  require(grid.cell_size > 0.0 && grid.dims[0] > 0 && grid.dims[1] > 0 && grid.dims[2] > 0,
          "Invalid cell grid configuration");
  const int64_t cell_count = static_cast<int64_t>(grid.dims[0]) * grid.dims[1] * grid.dims[2];
  require(cell_count == static_cast<int64_t>(cells.cell_count),
          "Cell list capacity does not match grid cell count");
  require(count <= cells.particle_count,
          "Cell list particle capacity is smaller than particle count");
  require(count <= static_cast<size_t>(INT32_MAX), "Particle count exceeds 32-bit index capacity");
  require(cell_count < static_cast<int64_t>(INT32_MAX), "Cell count exceeds 32-bit index capacity");
}

}  // namespace

void build_cell_list(Context& ctx, size_t count, const double* x, const double* y,
                     const double* z, const lsmps_cell_grid& grid, const lsmps_cell_list& cells) {
  check_cell_list(count, grid, cells);
  const size_t cc = static_cast<size_t>(grid.dims[0]) * grid.dims[1] * grid.dims[2];
  const GeometryScratch sc = context_geometry_scratch(ctx, cc, count, 0);
  cell_list_impl(ctx, sc, count, x, y, z, to_dev(grid), nullptr, nullptr, cc, cells);
}

namespace {
// build_one_neighbor_list (neighbor_search.cu:294-333)
void one_neighbor_list(Context& ctx, const GeometryScratch& sc, const lsmps_fluid_soa& fluid,
                       size_t source_count,
                       const double* sx, const double* sy, const double* sz,
                       const lsmps_cell_list& source_cells, const lsmps_cell_list& fluid_cells,
                       const lsmps_cell_grid& grid, const SegmentGrid* grids,
                       const int32_t* fluid_seg, double radius, bool exclude_self,
                       const lsmps_neighbor_list& nbrs, int32_t* device_total) {
  require(fluid.count <= nbrs.particle_count,
          "Neighbor list particle capacity is smaller than fluid count");
  require(nbrs.offsets != nullptr, "Neighbor list requires an offsets buffer");
  const double radius_squared = radius * radius;
  if (fluid.count == 0) {
    fill_i32(ctx.stream, nbrs.offsets, 1, 0);
    return;
  }
  const GridDev g = to_dev(grid);
  require(sc.particles_capacity >= source_count, "neighbor scan scratch too small",
          LSMPS_ERR_INTERNAL);
  SortedPoint* points = reinterpret_cast<SortedPoint*>(sc.points);
  if (source_count > 0) {
    gather_points_kernel<<<blocks_for(source_count), kBlock, 0, ctx.stream>>>(
        source_count, source_cells.sorted_indices, sx, sy, sz, points);
    LSMPS_LAUNCH_CHECK();
  }
  const int32_t* order = fluid_cells.sorted_indices;
  const size_t nf = fluid.count;
  const int self = exclude_self ? 1 : 0;
  // one pass of the scan with L lanes per particle, over one grid or the segments' grids
  auto scan = [&](auto write_tag, auto lanes_tag) {
    constexpr bool W = decltype(write_tag)::value;
    constexpr int L = decltype(lanes_tag)::value;
    int32_t* idx = W ? nbrs.indices : nullptr;
    if (grids) {
      neighbor_scan_kernel<W, true, L><<<blocks_for(nf * L), kBlock, 0, ctx.stream>>>(
          nf, order, fluid.x, fluid.y, fluid.z, g, grids, fluid_seg, source_cells.cell_begin,
          source_cells.cell_end, points, radius_squared, self, nbrs.offsets, idx);
    } else {
      neighbor_scan_kernel<W, false, L><<<blocks_for(nf * L), kBlock, 0, ctx.stream>>>(
          nf, order, fluid.x, fluid.y, fluid.z, g, nullptr, nullptr, source_cells.cell_begin,
          source_cells.cell_end, points, radius_squared, self, nbrs.offsets, idx);
    }
  };
  using Count = std::false_type;
  using Write = std::true_type;
  launch_tuned(ctx, kTuneScanCount, {1, 2, 4}, 1, nf, nullptr, [&](int lanes) {
    switch (lanes) {
      case 1: scan(Count{}, std::integral_constant<int, 1>{}); break;
      case 2: scan(Count{}, std::integral_constant<int, 2>{}); break;
      default: scan(Count{}, std::integral_constant<int, 4>{}); break;
    }
  });
  exclusive_scan_i32_with(ctx, nbrs.offsets, nbrs.offsets, fluid.count, sc.partials,
                          sc.partials_capacity);
  // Capacity check (require_neighbor_capacity, neighbor_search.cu:22-30,323-328).
  LSMPS_HIP(hipMemcpyAsync(ctx.host_i32, nbrs.offsets + fluid.count, sizeof(int32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  const int64_t required = ctx.host_i32[0];
  if (device_total) *device_total = static_cast<int32_t>(required);
  if (required < 0 || static_cast<size_t>(required) > nbrs.neighbor_count)
    raise(LSMPS_ERR_CAPACITY, "Neighbor list capacity exceeded: required " +
                                  std::to_string(required) + ", capacity " +
                                  std::to_string(nbrs.neighbor_count));
  require(required == 0 || nbrs.indices != nullptr, "Neighbor list requires an indices buffer");
  launch_tuned(ctx, kTuneScanWrite, {4, 8, 16}, 8, nf, nbrs.offsets, [&](int lanes) {
    switch (lanes) {
      case 4: scan(Write{}, std::integral_constant<int, 4>{}); break;
      case 8: scan(Write{}, std::integral_constant<int, 8>{}); break;
      default: scan(Write{}, std::integral_constant<int, 16>{}); break;
    }
  });
  LSMPS_LAUNCH_CHECK();
}
}  // namespace

lsmps_cell_grid neighbor_grid(const lsmps_config& config) {
  lsmps_cell_grid grid;
  for (int k = 0; k < 3; ++k) {
    grid.origin[k] = config.cell_origin[k];
    grid.dims[k] = config.cell_dims[k];
  }
  grid.cell_size = config.cell_size;
  return grid;
}

void build_neighbor_lists(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                          const lsmps_config& config, const lsmps_cell_list& fluid_cells,
                          const lsmps_cell_list& wall_cells,
                          const lsmps_neighbor_list& fluid_neighbors,
                          const lsmps_neighbor_list& wall_neighbors, int32_t* fluid_total,
                          int32_t* wall_total) {
  const lsmps_cell_grid grid = neighbor_grid(config);
  const size_t cc = static_cast<size_t>(std::max(grid.dims[0], 1)) * std::max(grid.dims[1], 1) *
                    std::max(grid.dims[2], 1);
  const GeometryScratch sc = context_geometry_scratch(ctx, cc, fluid.count, walls.count);
  build_neighbor_lists_with(ctx, sc, fluid, walls, config, fluid_cells, wall_cells,
                            fluid_neighbors, wall_neighbors, fluid_total, wall_total);
}

void build_neighbor_lists_with(Context& ctx, const GeometryScratch& sc,
                               const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                               const lsmps_config& config, const lsmps_cell_list& fluid_cells,
                               const lsmps_cell_list& wall_cells,
                               const lsmps_neighbor_list& fluid_neighbors,
                               const lsmps_neighbor_list& wall_neighbors, int32_t* fluid_total,
                               int32_t* wall_total) {
  const lsmps_cell_grid grid = neighbor_grid(config);
  const double radius = config.support_radius;
  require(radius > 0.0, "Neighbor search radius must be positive");
  require(radius <= grid.cell_size, "Neighbor search currently requires radius <= cell_size");
  check_cell_list(fluid.count, grid, fluid_cells);
  check_cell_list(walls.count, grid, wall_cells);
  const size_t cc = static_cast<size_t>(grid.dims[0]) * grid.dims[1] * grid.dims[2];
  cell_list_impl(ctx, sc, fluid.count, fluid.x, fluid.y, fluid.z, to_dev(grid), nullptr, nullptr,
                 cc, fluid_cells);
  cell_list_impl(ctx, sc, walls.count, walls.x, walls.y, walls.z, to_dev(grid), nullptr, nullptr,
                 cc, wall_cells);
  one_neighbor_list(ctx, sc, fluid, fluid.count, fluid.x, fluid.y, fluid.z, fluid_cells,
                    fluid_cells, grid, nullptr, nullptr, radius, true, fluid_neighbors,
                    fluid_total);
  one_neighbor_list(ctx, sc, fluid, walls.count, walls.x, walls.y, walls.z, wall_cells,
                    fluid_cells, grid, nullptr, nullptr, radius, false, wall_neighbors,
                    wall_total);
}

void build_neighbor_lists_segmented(Context& ctx, const lsmps_fluid_soa& fluid,
                                    const lsmps_wall_soa& walls, const lsmps_config& config,
                                    const Segments& segments, const lsmps_cell_list& fluid_cells,
                                    const lsmps_cell_list& wall_cells,
                                    const lsmps_neighbor_list& fluid_neighbors,
                                    const lsmps_neighbor_list& wall_neighbors) {
  const double radius = config.support_radius;
  require(radius > 0.0, "Neighbor search radius must be positive");
  require(segments.count > 0 && segments.grids && segments.fluid_seg &&
              (walls.count == 0 || segments.wall_seg),
          "Batched neighbor search requires segment grids and particle segment ids");
  require(segments.cells == fluid_cells.cell_count && segments.cells == wall_cells.cell_count,
          "Cell list capacity does not match the segments' cell count");
  require(segments.cells < static_cast<size_t>(INT32_MAX), "Cell count exceeds 32-bit index capacity");
  require(fluid.count <= fluid_cells.particle_count && walls.count <= wall_cells.particle_count,
          "Cell list particle capacity is smaller than particle count");
  const GridDev none{};
  const GeometryScratch sc = context_geometry_scratch(ctx, segments.cells, fluid.count,
                                                      walls.count);
  cell_list_impl(ctx, sc, fluid.count, fluid.x, fluid.y, fluid.z, none, segments.grids,
                 segments.fluid_seg, segments.cells, fluid_cells);
  cell_list_impl(ctx, sc, walls.count, walls.x, walls.y, walls.z, none, segments.grids,
                 segments.wall_seg, segments.cells, wall_cells);
  lsmps_cell_grid unused{};
  unused.cell_size = 1.0;
  unused.dims[0] = unused.dims[1] = unused.dims[2] = 1;
  one_neighbor_list(ctx, sc, fluid, fluid.count, fluid.x, fluid.y, fluid.z, fluid_cells,
                    fluid_cells, unused, segments.grids, segments.fluid_seg, radius, true,
                    fluid_neighbors, nullptr);
  one_neighbor_list(ctx, sc, fluid, walls.count, walls.x, walls.y, walls.z, wall_cells,
                    fluid_cells, unused, segments.grids, segments.fluid_seg, radius, false,
                    wall_neighbors, nullptr);
}

}  // namespace lsmps
