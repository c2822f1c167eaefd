// lsmps_oracle.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// An FP64 C++17/OpenMP restatement of lsmps-3D's per-step particle pipeline.
// It is the parity checker of the HIP library and the CPU baseline timed by
// bench.py; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  Every function cites the reference file:line it follows
// (paths relative to the lsmps-3D source tree).  All arithmetic is IEEE double
// compiled with -ffp-contract=off, so the neighbor tables and surface labels
// are reproducible bit for bit by any implementation that performs the same
// correctly rounded operations in the same order.
//
// Deliberate differences from the CUDA reference (each one is a documented
// restatement choice):
//  * rsqrt (surface_detection.cu:21-24) is replaced by 1/sqrt, correctly rounded.
//  * Thrust's sort_by_key (neighbor_search.cu:287) is restated as a stable
//    counting sort: within a cell, ascending particle index.
//  * cuSOLVER potrf/potrs (moment_matrix.cu:430-471) are restated as a textbook
//    Cholesky-Banachiewicz factorisation and per-column triangular solves.
//  * AMGX GMRES (ppe_matrix.cu:586-679) is replaced by IDR(s) (default; with
//    Jacobi or restricted-additive-Schwarz preconditioning), BiCGStab or PCG to
//    a relative true-residual tolerance, from p = 0, followed by the same clamp
//    p = max(p, 0) (ppe_matrix.cu:662).
//
// tests/test_oracle_*.py pin each function against the reference's own
// unit-test known answers, closed forms and brute force.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include <omp.h>

#include "lsmps3d_amd.h"

namespace {

constexpr int kTypeA = 9;
constexpr int kTypeB = 10;

thread_local std::string g_error;

int fail(int code, const std::string& message) {
  g_error = message;
  return code;
}

// ---------------------------------------------------------------------------
// weights
// particle_weight (surface_detection.cu:26-31), moment_weight (moment_matrix.cu:145-150),
// provision_weight (explicit_update.cu:121-126), ppe moment_weight (ppe_matrix.cu:148-153).
inline double weight_fn(double distance, double support_radius, int kernel) {
  if (distance >= support_radius) return 0.0;
  const double t = 1.0 - distance / support_radius;
  return kernel == LSMPS_WEIGHT_QUADRATIC ? t * t : t;
}
// correction_weight (pressure_correction.cu:127-132): also zero at distance <= 0.
inline double correction_weight_fn(double distance, double support_radius, int kernel) {
  if (distance >= support_radius || distance <= 0.0) return 0.0;
  const double t = 1.0 - distance / support_radius;
  return kernel == LSMPS_WEIGHT_QUADRATIC ? t * t : t;
}

// ---------------------------------------------------------------------------
// neighbor search (src/neighbor/neighbor_search.cu)

// flatten_cell (neighbor_search.cu:32-34)
inline int32_t flatten_cell(int32_t ix, int32_t iy, int32_t iz, const int32_t dims[3]) {
  return (iz * dims[1] + iy) * dims[0] + ix;
}
// clamp_cell_index (neighbor_search.cu:36-46)
inline int32_t clamp_cell_index(double coordinate, double origin, double cell_size, int32_t dim) {
  const double scaled = (coordinate - origin) / cell_size;
  const int32_t index = static_cast<int32_t>(std::floor(scaled));
  if (index < 0) return 0;
  if (index >= dim) return dim - 1;
  return index;
}
// within_radius (neighbor_search.cu:48-59)
inline bool within_radius(double ax, double ay, double az, double bx, double by, double bz,
                          double radius_squared) {
  const double dx = ax - bx;
  const double dy = ay - by;
  const double dz = az - bz;
  return dx * dx + dy * dy + dz * dz <= radius_squared;
}

// build_cell_list_impl (neighbor_search.cu:255-292): reset ranges, assign cells,
// stable sort by cell id, cell ranges.
int build_cell_list(size_t count, const double* x, const double* y, const double* z,
                    const lsmps_cell_grid& grid, const lsmps_cell_list& cells) {
  if (grid.cell_size <= 0.0 || grid.dims[0] <= 0 || grid.dims[1] <= 0 || grid.dims[2] <= 0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Invalid cell grid configuration");
  const int64_t cell_count = static_cast<int64_t>(grid.dims[0]) * grid.dims[1] * grid.dims[2];
  if (cell_count != static_cast<int64_t>(cells.cell_count))
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Cell list capacity does not match grid cell count");
  if (count > cells.particle_count)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Cell list particle capacity is smaller than particle count");
  for (int64_t c = 0; c < cell_count; ++c) {
    cells.cell_begin[c] = -1;
    cells.cell_end[c] = -1;
  }
  if (count == 0) return LSMPS_OK;
  std::vector<int32_t> ids(count);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(count); ++i) {
    const int32_t ix = clamp_cell_index(x[i], grid.origin[0], grid.cell_size, grid.dims[0]);
    const int32_t iy = clamp_cell_index(y[i], grid.origin[1], grid.cell_size, grid.dims[1]);
    const int32_t iz = clamp_cell_index(z[i], grid.origin[2], grid.cell_size, grid.dims[2]);
    ids[i] = flatten_cell(ix, iy, iz, grid.dims);
  }
  // Stable counting sort (restates thrust::sort_by_key, neighbor_search.cu:285-287).
  std::vector<int64_t> start(static_cast<size_t>(cell_count) + 1, 0);
  for (size_t i = 0; i < count; ++i) ++start[static_cast<size_t>(ids[i]) + 1];
  for (int64_t c = 0; c < cell_count; ++c) start[c + 1] += start[c];
  std::vector<int64_t> cursor(start.begin(), start.end() - 1);
  for (size_t i = 0; i < count; ++i) {
    const int64_t pos = cursor[ids[i]]++;
    cells.sorted_indices[pos] = static_cast<int32_t>(i);
    cells.cell_ids[pos] = ids[i];
  }
  // build_cell_ranges_kernel (neighbor_search.cu:85-98)
  for (size_t pos = 0; pos < count; ++pos) {
    const int32_t cell = cells.cell_ids[pos];
    if (pos == 0 || cells.cell_ids[pos - 1] != cell) cells.cell_begin[cell] = static_cast<int32_t>(pos);
    if (pos + 1 == count || cells.cell_ids[pos + 1] != cell)
      cells.cell_end[cell] = static_cast<int32_t>(pos + 1);
  }
  return LSMPS_OK;
}

// count/write_source_neighbors_for_fluid (neighbor_search.cu:100-219): 27 cells in
// (dz, dy, dx) order, each -1..1, within a cell in sorted order.
template <bool kWrite>
int32_t scan_neighbors(size_t i, const lsmps_fluid_soa& fluid, const double* sx, const double* sy,
                       const double* sz, const lsmps_cell_list& cells, const lsmps_cell_grid& grid,
                       double radius_squared, bool exclude_self, int32_t* out) {
  int32_t count = 0;
  const double px = fluid.x[i];
  const double py = fluid.y[i];
  const double pz = fluid.z[i];
  const int32_t cx = clamp_cell_index(px, grid.origin[0], grid.cell_size, grid.dims[0]);
  const int32_t cy = clamp_cell_index(py, grid.origin[1], grid.cell_size, grid.dims[1]);
  const int32_t cz = clamp_cell_index(pz, grid.origin[2], grid.cell_size, grid.dims[2]);
  for (int32_t dz = -1; dz <= 1; ++dz) {
    const int32_t nz = cz + dz;
    if (nz < 0 || nz >= grid.dims[2]) continue;
    for (int32_t dy = -1; dy <= 1; ++dy) {
      const int32_t ny = cy + dy;
      if (ny < 0 || ny >= grid.dims[1]) continue;
      for (int32_t dx = -1; dx <= 1; ++dx) {
        const int32_t nx = cx + dx;
        if (nx < 0 || nx >= grid.dims[0]) continue;
        const int32_t cell = flatten_cell(nx, ny, nz, grid.dims);
        const int32_t begin = cells.cell_begin[cell];
        const int32_t end = cells.cell_end[cell];
        if (begin < 0 || end < 0) continue;
        for (int32_t pos = begin; pos < end; ++pos) {
          const int32_t j = cells.sorted_indices[pos];
          if (exclude_self && static_cast<size_t>(j) == i) continue;
          if (within_radius(px, py, pz, sx[j], sy[j], sz[j], radius_squared)) {
            if (kWrite) out[count] = j;
            ++count;
          }
        }
      }
    }
  }
  return count;
}

// build_one_neighbor_list (neighbor_search.cu:294-333)
int build_one_neighbor_list(const lsmps_fluid_soa& fluid, size_t source_count, const double* sx,
                            const double* sy, const double* sz, const lsmps_cell_list& cells,
                            const lsmps_cell_grid& grid, double radius, bool exclude_self,
                            const lsmps_neighbor_list& nbrs) {
  if (fluid.count > nbrs.particle_count)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Neighbor list particle capacity is smaller than fluid count");
  (void)source_count;
  const double radius_squared = radius * radius;
  if (fluid.count == 0) {
    nbrs.offsets[0] = 0;
    return LSMPS_OK;
  }
  const int64_t n = static_cast<int64_t>(fluid.count);
  std::vector<int32_t> counts(fluid.count);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i)
    counts[i] = scan_neighbors<false>(static_cast<size_t>(i), fluid, sx, sy, sz, cells, grid,
                                      radius_squared, exclude_self, nullptr);
  int64_t total = 0;
  for (int64_t i = 0; i < n; ++i) {
    nbrs.offsets[i] = static_cast<int32_t>(total);
    total += counts[i];
  }
  nbrs.offsets[n] = static_cast<int32_t>(total);
  if (static_cast<size_t>(total) > nbrs.neighbor_count)
    return fail(LSMPS_ERR_CAPACITY, "Neighbor list capacity exceeded: required " +
                                        std::to_string(total) + ", capacity " +
                                        std::to_string(nbrs.neighbor_count));
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i)
    scan_neighbors<true>(static_cast<size_t>(i), fluid, sx, sy, sz, cells, grid, radius_squared,
                         exclude_self, nbrs.indices + nbrs.offsets[i]);
  return LSMPS_OK;
}

// ---------------------------------------------------------------------------
// surface detection (src/surface/surface_detection.cu)

// safe_rsqrt (surface_detection.cu:21-24) restated with a correctly rounded 1/sqrt.
inline double safe_inv_sqrt(double value) {
  const double kEpsilon = 1.0e-20;
  return 1.0 / std::sqrt(value > kEpsilon ? value : kEpsilon);
}

// accumulate_neighbor_moments (surface_detection.cu:33-59)
inline void accumulate_neighbor_moments(double dx, double dy, double dz, double re, int kernel,
                                        double& number_density, double& sx, double& sy,
                                        double& sz, int& weighted_count) {
  const double distance_squared = dx * dx + dy * dy + dz * dz;
  if (distance_squared <= 0.0) return;
  const double distance = std::sqrt(distance_squared);
  const double weight = weight_fn(distance, re, kernel);
  if (weight <= 0.0) return;
  const double inv_distance = 1.0 / distance;
  number_density += weight;
  sx += weight * dx * inv_distance;
  sy += weight * dy * inv_distance;
  sz += weight * dz * inv_distance;
  ++weighted_count;
}

struct WallBasis {
  int count = 0;
  double x[3] = {0, 0, 0};
  double y[3] = {0, 0, 0};
  double z[3] = {0, 0, 0};
};

// add_wall_basis_direction (surface_detection.cu:123-149)
inline void add_wall_basis_direction(WallBasis& basis, double x, double y, double z,
                                     double threshold) {
  double rx = x, ry = y, rz = z;
  for (int k = 0; k < basis.count; ++k) {
    const double projection = rx * basis.x[k] + ry * basis.y[k] + rz * basis.z[k];
    rx -= projection * basis.x[k];
    ry -= projection * basis.y[k];
    rz -= projection * basis.z[k];
  }
  const double length_squared = rx * rx + ry * ry + rz * rz;
  if (length_squared <= threshold * threshold || basis.count >= 3) return;
  const double inv_length = safe_inv_sqrt(length_squared);
  basis.x[basis.count] = rx * inv_length;
  basis.y[basis.count] = ry * inv_length;
  basis.z[basis.count] = rz * inv_length;
  ++basis.count;
}

// build_wall_basis (surface_detection.cu:151-197)
WallBasis build_wall_basis(size_t i, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                           const lsmps_neighbor_list& wn, const lsmps_config& cfg) {
  WallBasis basis;
  if (!cfg.include_wall_neighbors) return basis;
  const double px = fluid.x[i], py = fluid.y[i], pz = fluid.z[i];
  for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
    const int32_t j = wn.indices[c];
    const double dx = walls.x[j] - px;
    const double dy = walls.y[j] - py;
    const double dz = walls.z[j] - pz;
    const double distance_squared = dx * dx + dy * dy + dz * dz;
    if (distance_squared <= 0.0) continue;
    const double distance = std::sqrt(distance_squared);
    if (weight_fn(distance, cfg.support_radius, cfg.weight_kernel) <= 0.0) continue;
    const double nx = walls.normal_x[j], ny = walls.normal_y[j], nz = walls.normal_z[j];
    const double normal_length_squared = nx * nx + ny * ny + nz * nz;
    if (normal_length_squared <= 0.0) continue;
    const double inv_normal_length = safe_inv_sqrt(normal_length_squared);
    add_wall_basis_direction(basis, -nx * inv_normal_length, -ny * inv_normal_length,
                             -nz * inv_normal_length, cfg.wall_normal_independence_threshold);
  }
  return basis;
}

// compute_uniform_reference_number_density (surface_detection.cu:494-520)
double reference_number_density(double l0, double re, int kernel) {
  const int32_t extent = static_cast<int32_t>(std::ceil(re / l0));
  double density = 0.0;
  for (int32_t iz = -extent; iz <= extent; ++iz)
    for (int32_t iy = -extent; iy <= extent; ++iy)
      for (int32_t ix = -extent; ix <= extent; ++ix) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        const double dx = static_cast<double>(ix) * l0;
        const double dy = static_cast<double>(iy) * l0;
        const double dz = static_cast<double>(iz) * l0;
        const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
        density += weight_fn(distance, re, kernel);
      }
  return density;
}

// classify_primary_surface_kernel body (surface_detection.cu:199-320)
void classify_one(size_t i, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                  const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                  const lsmps_config& cfg, double n0, const lsmps_surface_diagnostics* diag) {
  const int32_t fluid_count = fn.offsets[i + 1] - fn.offsets[i];
  const int32_t wall_count = wn.offsets[i + 1] - wn.offsets[i];
  const int32_t classification_count =
      cfg.include_wall_neighbors ? fluid_count + wall_count : fluid_count;
  double number_density = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  int weighted_count = 0;
  const double px = fluid.x[i], py = fluid.y[i], pz = fluid.z[i];
  for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
    const int32_t j = fn.indices[c];
    accumulate_neighbor_moments(fluid.x[j] - px, fluid.y[j] - py, fluid.z[j] - pz,
                                cfg.support_radius, cfg.weight_kernel, number_density, sx, sy, sz,
                                weighted_count);
  }
  if (cfg.include_wall_neighbors) {
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      accumulate_neighbor_moments(walls.x[j] - px, walls.y[j] - py, walls.z[j] - pz,
                                  cfg.support_radius, cfg.weight_kernel, number_density, sx, sy,
                                  sz, weighted_count);
    }
  }
  double anisotropy = 0.0;
  if (weighted_count > 0)
    anisotropy = std::sqrt(sx * sx + sy * sy + sz * sz) / static_cast<double>(weighted_count);
  const double ratio = number_density / n0;
  const double mx = -sx, my = -sy, mz = -sz;
  const double missing_length_squared = mx * mx + my * my + mz * mz;
  double ax = mx, ay = my, az = mz;
  if (missing_length_squared > 0.0) {
    const WallBasis basis = build_wall_basis(i, fluid, walls, wn, cfg);
    for (int k = 0; k < basis.count; ++k) {
      const double projection = ax * basis.x[k] + ay * basis.y[k] + az * basis.z[k];
      if (projection > 0.0) {
        ax -= projection * basis.x[k];
        ay -= projection * basis.y[k];
        az -= projection * basis.z[k];
      }
    }
  }
  const double air_length_squared = ax * ax + ay * ay + az * az;
  const double air_length = std::sqrt(air_length_squared);
  const double missing_length = std::sqrt(missing_length_squared);
  const double open_ratio = missing_length > 0.0 ? air_length / missing_length : 0.0;
  const double air_anisotropy =
      weighted_count > 0 ? air_length / static_cast<double>(weighted_count) : 0.0;
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if (air_length_squared > 0.0) {
    const double inv = safe_inv_sqrt(air_length_squared);
    nx = ax * inv;
    ny = ay * inv;
    nz = az * inv;
  }
  int32_t type = LSMPS_SURFACE_INNER;
  if (classification_count < cfg.splash_neighbor_threshold) {
    type = LSMPS_SURFACE_SPLASH;
  } else if (ratio < cfg.number_density_ratio_threshold &&
             open_ratio > cfg.air_open_ratio_threshold &&
             air_anisotropy > cfg.air_anisotropy_threshold) {
    type = LSMPS_SURFACE_SURFACE;
  }
  fluid.surface_type[i] = type;
  if (diag) {
    if (diag->fluid_neighbor_count) diag->fluid_neighbor_count[i] = fluid_count;
    if (diag->wall_neighbor_count) diag->wall_neighbor_count[i] = wall_count;
    if (diag->number_density) diag->number_density[i] = number_density;
    if (diag->number_density_ratio) diag->number_density_ratio[i] = ratio;
    if (diag->anisotropy) diag->anisotropy[i] = anisotropy;
    if (diag->air_open_ratio) diag->air_open_ratio[i] = open_ratio;
    if (diag->air_anisotropy) diag->air_anisotropy[i] = air_anisotropy;
    if (diag->surface_normal_x) diag->surface_normal_x[i] = nx;
    if (diag->surface_normal_y) diag->surface_normal_y[i] = ny;
    if (diag->surface_normal_z) diag->surface_normal_z[i] = nz;
  }
}

// classify_surface_particles (surface_detection.cu:522-553).  The near-surface
// expansion (expand_near_surface_kernel, :322-348) only promotes Inner -> NearSurface
// and only tests `== Surface`, which it never writes, so evaluating it against a
// snapshot of the primary labels gives the reference's deterministic result.
int classify_surface(const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                     const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                     const lsmps_config& cfg, const lsmps_surface_diagnostics* diag) {
  if (cfg.support_radius <= 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Surface detection support radius must be positive");
  if (fluid.count > fn.particle_count || fluid.count > wn.particle_count)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Surface detection neighbor lists are smaller than fluid particle count");
  if (cfg.near_surface_radius < 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Near-surface radius must be non-negative");
  if (cfg.particle_spacing <= 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Reference number density requires positive spacing and support radius");
  const double n0 = reference_number_density(cfg.particle_spacing, cfg.support_radius,
                                             cfg.weight_kernel);
  if (n0 <= 0.0) return fail(LSMPS_ERR_INVALID_ARGUMENT, "Reference number density must be positive");
  const int64_t n = static_cast<int64_t>(fluid.count);
  if (n == 0) return LSMPS_OK;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) classify_one(static_cast<size_t>(i), fluid, walls, fn, wn, cfg, n0, diag);
  if (cfg.near_surface_radius > 0.0) {
    const double r2 = cfg.near_surface_radius * cfg.near_surface_radius;
    std::vector<int32_t> primary(fluid.surface_type, fluid.surface_type + n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      if (primary[i] != LSMPS_SURFACE_INNER) continue;
      const double px = fluid.x[i], py = fluid.y[i], pz = fluid.z[i];
      for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
        const int32_t j = fn.indices[c];
        if (primary[j] != LSMPS_SURFACE_SURFACE) continue;
        const double dx = fluid.x[j] - px;
        const double dy = fluid.y[j] - py;
        const double dz = fluid.z[j] - pz;
        if (dx * dx + dy * dy + dz * dz <= r2) {
          fluid.surface_type[i] = LSMPS_SURFACE_NEAR_SURFACE;
          break;
        }
      }
    }
  }
  return LSMPS_OK;
}

// virtual_light_direction (surface_detection.cu:350-383).
void light_direction(int k, double& ux, double& uy, double& uz) {
  const double h = 0.7071067811865475;
  static const double table[14][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
                                      {0, 0, -1}, {h, h, 0},  {h, -h, 0}, {-h, h, 0}, {-h, -h, 0},
                                      {h, 0, h},  {h, 0, -h}, {0, h, h},  {0, h, -h}};
  ux = table[k][0];
  uy = table[k][1];
  uz = table[k][2];
}

// fluid_neighbor_blocks_light / wall_neighbor_blocks_light (surface_detection.cu:385-447):
// does some neighbor of row i lie within the cone around direction k?
bool row_blocks_light(size_t i, int k, const lsmps_fluid_soa& fluid, const double* nx,
                      const double* ny, const double* nz, const lsmps_neighbor_list& nl,
                      const lsmps_config& cfg) {
  double ux, uy, uz;
  light_direction(k, ux, uy, uz);
  const double px = fluid.x[i], py = fluid.y[i], pz = fluid.z[i];
  for (int32_t c = nl.offsets[i]; c < nl.offsets[i + 1]; ++c) {
    const int32_t j = nl.indices[c];
    const double dx = nx[j] - px, dy = ny[j] - py, dz = nz[j] - pz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= 0.0 || d2 > cfg.support_radius * cfg.support_radius) continue;
    const double proj = (dx * ux + dy * uy + dz * uz) * (1.0 / std::sqrt(d2 > 1.0e-20 ? d2 : 1.0e-20));
    if (proj >= cfg.virtual_light_cone_cosine) return true;
  }
  return false;
}

// compute_virtual_light_diagnostics + virtual_light_kernel
// (surface_detection.cu:449-475,555-578): direction by direction, walls only for a
// direction the fluid left open.
int virtual_light(const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                  const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                  const lsmps_config& cfg, int32_t* open_count, double* open_fraction) {
  if (cfg.support_radius <= 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Surface detection support radius must be positive");
  if (fluid.count > fn.particle_count || fluid.count > wn.particle_count)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Surface detection neighbor lists are smaller than fluid particle count");
  if (cfg.virtual_light_cone_cosine < -1.0 || cfg.virtual_light_cone_cosine > 1.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Virtual light cone cosine must be in [-1, 1]");
  if (open_count == nullptr && open_fraction == nullptr)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "Virtual light diagnostics require at least one output buffer");
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t s = 0; s < n; ++s) {
    const size_t i = static_cast<size_t>(s);
    int open = 0;
    for (int k = 0; k < 14; ++k) {
      bool blocked = row_blocks_light(i, k, fluid, fluid.x, fluid.y, fluid.z, fn, cfg);
      if (!blocked && cfg.include_wall_neighbors)
        blocked = row_blocks_light(i, k, fluid, walls.x, walls.y, walls.z, wn, cfg);
      if (!blocked) ++open;
    }
    if (open_count) open_count[i] = open;
    if (open_fraction) open_fraction[i] = static_cast<double>(open) / 14.0;
  }
  return LSMPS_OK;
}

// ---------------------------------------------------------------------------
// moment matrices (src/moment_matrix/moment_matrix.cu)

struct MomentRequest {
  int kind;
  int boundary;  // 0 none, 1 wall Dirichlet samples, 2 wall pressure Neumann
  int basis;     // LSMPS_BASIS_*
  int n;
};

// make_matrix_request (moment_matrix.cu:494-538)
MomentRequest make_request(int kind) {
  switch (kind) {
    case LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A: return {kind, 1, LSMPS_BASIS_TYPE_A, kTypeA};
    case LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A: return {kind, 2, LSMPS_BASIS_TYPE_A, kTypeA};
    case LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_B: return {kind, 2, LSMPS_BASIS_TYPE_B, kTypeB};
    default: return {LSMPS_MOMENT_FLUID_ONLY_A, 0, LSMPS_BASIS_TYPE_A, kTypeA};
  }
}

// basis_vector (moment_matrix.cu:152-175): note sx = dx / re (division).
inline int basis_vector(double dx, double dy, double dz, int basis_kind, double re, double* b) {
  const double sx = dx / re, sy = dy / re, sz = dz / re;
  int o = 0;
  if (basis_kind == LSMPS_BASIS_TYPE_B) b[o++] = 1.0;
  b[o++] = sx;
  b[o++] = sy;
  b[o++] = sz;
  b[o++] = sx * sx;
  b[o++] = sy * sy;
  b[o++] = sz * sz;
  b[o++] = sx * sy;
  b[o++] = sy * sz;
  b[o++] = sz * sx;
  return o;
}

// wall_neumann_vector (moment_matrix.cu:177-201); also ppe_wall_neumann_vector
// (ppe_matrix.cu:175-193) and correction_wall_neumann_vector (pressure_correction.cu:158-176).
inline int wall_neumann_vector(double dx, double dy, double dz, double nx, double ny, double nz,
                               int basis_kind, double re, double* q) {
  const double inv = 1.0 / re;
  int o = 0;
  if (basis_kind == LSMPS_BASIS_TYPE_B) q[o++] = 0.0;
  q[o++] = nx;
  q[o++] = ny;
  q[o++] = nz;
  q[o++] = 2.0 * dx * nx * inv;
  q[o++] = 2.0 * dy * ny * inv;
  q[o++] = 2.0 * dz * nz * inv;
  q[o++] = (dy * nx + dx * ny) * inv;
  q[o++] = (dz * ny + dy * nz) * inv;
  q[o++] = (dx * nz + dz * nx) * inv;
  return o;
}

// add_outer_product (moment_matrix.cu:203-209)
inline void add_outer(const double* v, int n, double w, double* m) {
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) m[r + c * n] += w * v[r] * v[c];
}

// Cholesky-Banachiewicz (lower) + per-column solves: restates
// cusolverDnDpotrfBatched + n x cusolverDnDpotrsBatched (moment_matrix.cu:430-471).
// Returns info (0 = success, k = order of the failing leading minor).
int cholesky_inverse(const double* m, int n, double* inv) {
  double l[kTypeB * kTypeB];
  for (int j = 0; j < n; ++j) {
    double s = m[j + j * n];
    for (int k = 0; k < j; ++k) s -= l[j + k * n] * l[j + k * n];
    if (!(s > 0.0)) {
      for (int e = 0; e < n * n; ++e) inv[e] = std::numeric_limits<double>::quiet_NaN();
      return j + 1;
    }
    const double d = std::sqrt(s);
    l[j + j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = m[i + j * n];
      for (int k = 0; k < j; ++k) t -= l[i + k * n] * l[j + k * n];
      l[i + j * n] = t / d;
    }
  }
  for (int col = 0; col < n; ++col) {
    double y[kTypeB];
    for (int i = 0; i < n; ++i) {
      double t = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) t -= l[i + k * n] * y[k];
      y[i] = t / l[i + i * n];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < n; ++k) t -= l[k + i * n] * inv[k + col * n];
      inv[i + col * n] = t / l[i + i * n];
    }
  }
  return 0;
}

// build_matrix_cache_kernel (moment_matrix.cu:254-309) + ensure_matrix_internal (:398-476)
void build_moment_one(size_t i, const MomentRequest& rq, const lsmps_fluid_soa& fluid,
                      const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                      const lsmps_neighbor_list& wn, const lsmps_config& cfg, double* matrix,
                      double* inverse, double* trace_out, double* reg_out, int32_t* info) {
  const int n = rq.n;
  const double re = cfg.support_radius;
  for (int e = 0; e < n * n; ++e) matrix[e] = 0.0;
  const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
  double v[kTypeB];
  // add_sample_to_matrix (moment_matrix.cu:211-229)
  auto add_sample = [&](double dx, double dy, double dz, double scale) {
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= 0.0) return;
    const double distance = std::sqrt(d2);
    const double w = scale * weight_fn(distance, re, cfg.weight_kernel);
    if (w <= 0.0) return;
    const int nb = basis_vector(dx, dy, dz, rq.basis, re, v);
    add_outer(v, nb, w, matrix);
  };
  for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
    const int32_t j = fn.indices[c];
    add_sample(fluid.x[j] - xi, fluid.y[j] - yi, fluid.z[j] - zi, 1.0);
  }
  if (rq.boundary != 0) {
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double dx = walls.x[j] - xi, dy = walls.y[j] - yi, dz = walls.z[j] - zi;
      if (rq.boundary == 2) {
        // add_wall_neumann_to_matrix (moment_matrix.cu:231-252)
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= 0.0) continue;
        const double distance = std::sqrt(d2);
        const double w = cfg.lsmps_wall_weight_scale * weight_fn(distance, re, cfg.weight_kernel);
        if (w <= 0.0) continue;
        const int nq = wall_neumann_vector(dx, dy, dz, walls.normal_x[j], walls.normal_y[j],
                                           walls.normal_z[j], rq.basis, re, v);
        add_outer(v, nq, w, matrix);
      } else {
        add_sample(dx, dy, dz, cfg.lsmps_wall_weight_scale);
      }
    }
  }
  double trace = 0.0;
  for (int r = 0; r < n; ++r) trace += matrix[r + r * n];
  const double reg = cfg.lsmps_regularization * (trace > 0.0 ? trace / static_cast<double>(n) : 1.0);
  for (int r = 0; r < n; ++r) matrix[r + r * n] += reg;
  *trace_out = trace;
  *reg_out = reg;
  *info = cholesky_inverse(matrix, n, inverse);
}

int moment_prepare(int kind, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                   const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                   const lsmps_config& cfg, double* matrices, double* inverses, double* trace,
                   double* reg, int32_t* info) {
  if (cfg.support_radius <= 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Moment support radius must be positive");
  if (cfg.lsmps_regularization < 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Moment regularization must be non-negative");
  if (cfg.lsmps_wall_weight_scale < 0.0)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Moment wall weight scale must be non-negative");
  const MomentRequest rq = make_request(kind);
  const int nn = rq.n * rq.n;
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    double m_local[kTypeB * kTypeB];
    double* m = matrices ? matrices + i * nn : m_local;
    build_moment_one(static_cast<size_t>(i), rq, fluid, walls, fn, wn, cfg, m, inverses + i * nn,
                     trace + i, reg + i, info + i);
  }
  return LSMPS_OK;
}

// ---------------------------------------------------------------------------
// provision (src/provision/explicit_update.cu)

// provision_type_a_basis (explicit_update.cu:128-146); same as ppe_type_a_basis
// (ppe_matrix.cu:155-173) and correction_type_a_basis (pressure_correction.cu:138-156):
// sx = dx * (1/re).
inline void type_a_basis_mul(double dx, double dy, double dz, double re, double* b) {
  const double inv = 1.0 / re;
  const double sx = dx * inv, sy = dy * inv, sz = dz * inv;
  b[0] = sx;
  b[1] = sy;
  b[2] = sz;
  b[3] = sx * sx;
  b[4] = sy * sy;
  b[5] = sz * sz;
  b[6] = sx * sy;
  b[7] = sy * sz;
  b[8] = sz * sx;
}

// accumulate_velocity_rhs (explicit_update.cu:148-162; ppe_matrix.cu:217-231)
inline void accumulate_velocity_rhs(const double* b, double w, double ux, double uy, double uz,
                                    double* rx, double* ry, double* rz) {
  for (int r = 0; r < kTypeA; ++r) {
    const double wb = w * b[r];
    rx[r] += wb * ux;
    ry[r] += wb * uy;
    rz[r] += wb * uz;
  }
}

// laplacian_from_inverse_rhs (explicit_update.cu:164-176)
inline double laplacian_from_inverse_rhs(const double* inv, const double* rhs, double re) {
  double second = 0.0;
  for (int r = 3; r <= 5; ++r) {
    double coeff = 0.0;
    for (int c = 0; c < kTypeA; ++c) coeff += inv[r + c * kTypeA] * rhs[c];
    second += coeff;
  }
  return 2.0 * second / (re * re);
}

// compute_velocity_laplacian_kernel (explicit_update.cu:178-248)
void velocity_laplacian(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                        const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                        const lsmps_neighbor_list& wn, const double* vel_inverse, double* lx,
                        double* ly, double* lz) {
  const int64_t n = static_cast<int64_t>(fluid.count);
  const double re = cfg.support_radius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
    double rx[kTypeA] = {0}, ry[kTypeA] = {0}, rz[kTypeA] = {0}, b[kTypeA];
    for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
      const int32_t j = fn.indices[c];
      const double dx = fluid.x[j] - xi, dy = fluid.y[j] - yi, dz = fluid.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = weight_fn(distance, re, cfg.weight_kernel);
      if (w <= 0.0) continue;
      type_a_basis_mul(dx, dy, dz, re, b);
      accumulate_velocity_rhs(b, w, fluid.vx[j] - fluid.vx[i], fluid.vy[j] - fluid.vy[i],
                              fluid.vz[j] - fluid.vz[i], rx, ry, rz);
    }
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double dx = walls.x[j] - xi, dy = walls.y[j] - yi, dz = walls.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = cfg.lsmps_wall_weight_scale * weight_fn(distance, re, cfg.weight_kernel);
      if (w <= 0.0) continue;
      type_a_basis_mul(dx, dy, dz, re, b);
      accumulate_velocity_rhs(b, w, walls.vx[j] - fluid.vx[i], walls.vy[j] - fluid.vy[i],
                              walls.vz[j] - fluid.vz[i], rx, ry, rz);
    }
    const double* inv = vel_inverse + i * kTypeA * kTypeA;
    lx[i] = laplacian_from_inverse_rhs(inv, rx, re);
    ly[i] = laplacian_from_inverse_rhs(inv, ry, re);
    lz[i] = laplacian_from_inverse_rhs(inv, rz, re);
  }
}

// combine_explicit_velocity_kernel (explicit_update.cu:250-270)
void combine_velocity(const lsmps_config& cfg, const lsmps_fluid_soa& fluid, const double* lx,
                      const double* ly, const double* lz, const lsmps_fluid_soa& tmp) {
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const bool splash = fluid.surface_type && fluid.surface_type[i] == LSMPS_SURFACE_SPLASH;
    const double nu = splash ? 0.0 : cfg.kinematic_viscosity;
    const double dt = cfg.time_step;
    tmp.vx[i] = fluid.vx[i] + dt * (nu * lx[i] + cfg.gravity[0]);
    tmp.vy[i] = fluid.vy[i] + dt * (nu * ly[i] + cfg.gravity[1]);
    tmp.vz[i] = fluid.vz[i] + dt * (nu * lz[i] + cfg.gravity[2]);
  }
}

// combine_explicit_wall_velocity_kernel (explicit_update.cu:272-284)
void combine_wall_velocity(const lsmps_config& cfg, const lsmps_wall_soa& walls,
                           const lsmps_wall_soa& tmp) {
  const int64_t n = static_cast<int64_t>(walls.count);
  for (int64_t i = 0; i < n; ++i) {
    tmp.vx[i] = walls.vx[i] + cfg.time_step * cfg.gravity[0];
    tmp.vy[i] = walls.vy[i] + cfg.time_step * cfg.gravity[1];
    tmp.vz[i] = walls.vz[i] + cfg.time_step * cfg.gravity[2];
  }
}

// ---------------------------------------------------------------------------
// PPE (src/ppe/ppe_matrix.cu)

inline double dot9(const double* a, const double* b) {
  double v = 0.0;
  for (int k = 0; k < 9; ++k) v += a[k] * b[k];
  return v;
}

// build_ppe_row_offsets_kernel (ppe_matrix.cu:133-146) + assemble_ppe_system_kernel (:243-370)
int ppe_assemble(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                 const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                 const lsmps_neighbor_list& wn, const lsmps_fluid_soa& tmp,
                 const lsmps_wall_soa& tmpw, const double* vel_inverse, const double* pre_inverse,
                 const lsmps_ppe_workspace& ws) {
  const int64_t n = static_cast<int64_t>(fluid.count);
  if (ws.matrix.rows < fluid.count || ws.matrix.cols < fluid.count ||
      ws.matrix.nnz < fluid.count + fn.neighbor_count)
    return fail(LSMPS_ERR_INVALID_ARGUMENT,
                "PPE matrix workspace is smaller than the requested fluid system");
  if (n == 0) return LSMPS_OK;
  for (int64_t i = 0; i < n; ++i) ws.matrix.row_offsets[i] = static_cast<int32_t>(i) + fn.offsets[i];
  ws.matrix.row_offsets[n] = static_cast<int32_t>(n) + fn.offsets[n];
  const double re = cfg.support_radius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t row_begin = ws.matrix.row_offsets[i];
    const int32_t nb = fn.offsets[i], ne = fn.offsets[i + 1];
    const int32_t st = fluid.surface_type ? fluid.surface_type[i] : 0;
    if (st == LSMPS_SURFACE_SURFACE || st == LSMPS_SURFACE_SPLASH) {
      ws.matrix.col_indices[row_begin] = static_cast<int32_t>(i);
      ws.matrix.values[row_begin] = 1.0;
      for (int32_t c = nb; c < ne; ++c) {
        const int32_t out = row_begin + 1 + c - nb;
        ws.matrix.col_indices[out] = fn.indices[c];
        ws.matrix.values[out] = 0.0;
      }
      ws.rhs[i] = 0.0;
      ws.divergence[i] = 0.0;
      if (ws.pressure_laplacian) ws.pressure_laplacian[i] = 0.0;
      continue;
    }
    const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
    const double* pinv = pre_inverse + i * 81;
    double L[9];
    // compute_pressure_laplacian_row (ppe_matrix.cu:203-215)
    for (int c = 0; c < 9; ++c) L[c] = pinv[3 + c * 9] + pinv[4 + c * 9] + pinv[5 + c * 9];
    double rx[9] = {0}, ry[9] = {0}, rz[9] = {0}, b[9], q[9];
    double diagonal = 0.0;
    for (int32_t c = nb; c < ne; ++c) {
      const int32_t j = fn.indices[c];
      const double dx = fluid.x[j] - xi, dy = fluid.y[j] - yi, dz = fluid.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = weight_fn(distance, re, cfg.weight_kernel);
      double off = 0.0;
      if (w > 0.0) {
        type_a_basis_mul(dx, dy, dz, re, b);
        off = 2.0 * w * dot9(L, b) / (re * re * cfg.density);
        diagonal -= off;
        accumulate_velocity_rhs(b, w, tmp.vx[j] - tmp.vx[i], tmp.vy[j] - tmp.vy[i],
                                tmp.vz[j] - tmp.vz[i], rx, ry, rz);
      }
      const int32_t out = row_begin + 1 + c - nb;
      ws.matrix.col_indices[out] = j;
      ws.matrix.values[out] = off;
    }
    double wall_rhs = 0.0;
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double dx = walls.x[j] - xi, dy = walls.y[j] - yi, dz = walls.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double bw = weight_fn(distance, re, cfg.weight_kernel);
      if (bw <= 0.0) continue;
      type_a_basis_mul(dx, dy, dz, re, b);
      accumulate_velocity_rhs(b, cfg.lsmps_wall_weight_scale * bw, tmpw.vx[j] - tmp.vx[i],
                              tmpw.vy[j] - tmp.vy[i], tmpw.vz[j] - tmp.vz[i], rx, ry, rz);
      const double ng = cfg.gravity[0] * walls.normal_x[j] + cfg.gravity[1] * walls.normal_y[j] +
                        cfg.gravity[2] * walls.normal_z[j];
      wall_neumann_vector(dx, dy, dz, walls.normal_x[j], walls.normal_y[j], walls.normal_z[j],
                          LSMPS_BASIS_TYPE_A, re, q);
      wall_rhs += 2.0 * cfg.lsmps_wall_weight_scale * bw * ng * dot9(L, q) / re;
    }
    const double* vinv = vel_inverse + i * 81;
    // first_coefficient_from_inverse_rhs (ppe_matrix.cu:233-241)
    auto coeff = [&](const double* rhs, int row) {
      double s = 0.0;
      for (int c = 0; c < 9; ++c) s += vinv[row + c * 9] * rhs[c];
      return s;
    };
    const double divergence = (coeff(rx, 0) + coeff(ry, 1) + coeff(rz, 2)) / re;
    ws.matrix.col_indices[row_begin] = static_cast<int32_t>(i);
    ws.matrix.values[row_begin] = diagonal;
    ws.divergence[i] = divergence;
    ws.rhs[i] = divergence / cfg.time_step - wall_rhs;
    if (ws.pressure_laplacian) ws.pressure_laplacian[i] = wall_rhs;
  }
  return LSMPS_OK;
}

// ---------------------------------------------------------------------------
// PPE solve: replaces AmgxPpeSolver::solve (ppe_matrix.cu:586-679).

void spmv(const lsmps_csr& a, const double* x, double* y) {
  const int64_t n = static_cast<int64_t>(a.rows);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t c = a.row_offsets[i]; c < a.row_offsets[i + 1]; ++c)
      s += a.values[c] * x[a.col_indices[c]];
    y[i] = s;
  }
}

double dot(const double* a, const double* b, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

double true_residual(const lsmps_csr& a, const double* b, const double* x, double* r) {
  spmv(a, x, r);
  const int64_t n = static_cast<int64_t>(a.rows);
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - r[i];
    s += r[i] * r[i];
  }
  return std::sqrt(s);
}

// Restricted additive Schwarz preconditioner (no reference counterpart: the
// reference hands the PPE to AMGX, ppe_matrix.cu:586-679).  Rows with
// off-diagonal couplings are "eligible"; eligible rows with a positive diagonal
// (wrong-sign LSMPS Laplacian rows at wall corners of the free surface) seed one
// dense block each: the eligible rows among the seed's columns and their
// columns (2 hops), or the 1-hop set alone when the 2-hop set has more than
// kSchwarzMaxBlock rows.  Each block is inverted by LU with partial pivoting;
// a covered row takes the mean of its blocks' values (summed in ascending block
// order), every other row Jacobi.
constexpr int kSchwarzMaxBlock = 256;

struct SchwarzBlocks {
  std::vector<std::vector<int32_t>> members;  // ascending row indices
  std::vector<std::vector<double>> inverse;   // row-major n x n
  std::vector<int32_t> cover;                 // blocks covering each row
};

// LU with partial pivoting (largest |a|, first row on ties) and n column
// solves into an explicit inverse.  A zero pivot is replaced by 1e-300.
void lu_inverse(std::vector<double>& a, int n, std::vector<double>& inv) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[static_cast<size_t>(k) * n + k]);
    for (int r = k + 1; r < n; ++r) {
      const double v = std::fabs(a[static_cast<size_t>(r) * n + k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j)
        std::swap(a[static_cast<size_t>(k) * n + j], a[static_cast<size_t>(p) * n + j]);
    double akk = a[static_cast<size_t>(k) * n + k];
    if (akk == 0.0) akk = 1.0e-300;
    for (int r = k + 1; r < n; ++r) a[static_cast<size_t>(r) * n + k] /= akk;
    for (int r = k + 1; r < n; ++r) {
      const double l = a[static_cast<size_t>(r) * n + k];
      for (int j = k + 1; j < n; ++j)
        a[static_cast<size_t>(r) * n + j] -= l * a[static_cast<size_t>(k) * n + j];
    }
  }
  inv.assign(static_cast<size_t>(n) * n, 0.0);
  std::vector<double> col(n);
  for (int t = 0; t < n; ++t) {
    for (int i = 0; i < n; ++i) col[i] = i == t ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) std::swap(col[k], col[piv[k]]);
    for (int i = 0; i < n; ++i) {
      double s = col[i];
      for (int j = 0; j < i; ++j) s -= a[static_cast<size_t>(i) * n + j] * col[j];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int j = i + 1; j < n; ++j) s -= a[static_cast<size_t>(i) * n + j] * col[j];
      double d = a[static_cast<size_t>(i) * n + i];
      if (d == 0.0) d = 1.0e-300;
      col[i] = s / d;
    }
    for (int i = 0; i < n; ++i) inv[static_cast<size_t>(i) * n + t] = col[i];
  }
}

void schwarz_build(const lsmps_csr& a, SchwarzBlocks& S) {
  const int64_t n = static_cast<int64_t>(a.rows);
  std::vector<char> eligible(n, 0);
  std::vector<int32_t> seeds;
  for (int64_t i = 0; i < n; ++i) {
    double diag = 0.0, off = 0.0;
    for (int32_t c = a.row_offsets[i]; c < a.row_offsets[i + 1]; ++c) {
      if (a.col_indices[c] == i) diag += a.values[c];
      else off += std::fabs(a.values[c]);
    }
    eligible[i] = off > 0.0;
    if (off > 0.0 && diag > 0.0) seeds.push_back(static_cast<int32_t>(i));
  }
  const size_t nb = seeds.size();
  S.members.assign(nb, {});
  S.inverse.assign(nb, {});
#pragma omp parallel for schedule(dynamic)
  for (size_t b = 0; b < nb; ++b) {
    const int32_t seed = seeds[b];
    std::vector<int32_t> one, two;
    for (int32_t c = a.row_offsets[seed]; c < a.row_offsets[seed + 1]; ++c)
      if (eligible[a.col_indices[c]]) one.push_back(a.col_indices[c]);
    std::sort(one.begin(), one.end());
    one.erase(std::unique(one.begin(), one.end()), one.end());
    two = one;
    for (int32_t j : one)
      for (int32_t c = a.row_offsets[j]; c < a.row_offsets[j + 1]; ++c)
        if (eligible[a.col_indices[c]]) two.push_back(a.col_indices[c]);
    std::sort(two.begin(), two.end());
    two.erase(std::unique(two.begin(), two.end()), two.end());
    std::vector<int32_t>& m = S.members[b];
    m = two.size() <= static_cast<size_t>(kSchwarzMaxBlock) ? two : one;
    if (m.size() > static_cast<size_t>(kSchwarzMaxBlock)) m.resize(kSchwarzMaxBlock);
    const int bn = static_cast<int>(m.size());
    std::vector<double> dense(static_cast<size_t>(bn) * bn, 0.0);
    for (int r = 0; r < bn; ++r) {
      const int32_t row = m[r];
      for (int32_t c = a.row_offsets[row]; c < a.row_offsets[row + 1]; ++c) {
        const auto it = std::lower_bound(m.begin(), m.end(), a.col_indices[c]);
        if (it != m.end() && *it == a.col_indices[c])
          dense[static_cast<size_t>(r) * bn + (it - m.begin())] += a.values[c];
      }
    }
    lu_inverse(dense, bn, S.inverse[b]);
  }
  S.cover.assign(n, 0);
  for (const auto& m : S.members)
    for (int32_t r : m) ++S.cover[r];
}

// z = M^-1 v (z and v distinct).
void schwarz_apply(const SchwarzBlocks& S, const std::vector<double>& dinv, const double* v,
                   double* z, int64_t n) {
  for (int64_t i = 0; i < n; ++i) z[i] = S.cover[i] ? 0.0 : dinv[i] * v[i];
  std::vector<std::vector<double>> y(S.members.size());
#pragma omp parallel for schedule(dynamic)
  for (size_t b = 0; b < S.members.size(); ++b) {
    const auto& m = S.members[b];
    const int bn = static_cast<int>(m.size());
    y[b].assign(bn, 0.0);
    for (int r = 0; r < bn; ++r) {
      double s = 0.0;
      for (int j = 0; j < bn; ++j) s += S.inverse[b][static_cast<size_t>(r) * bn + j] * v[m[j]];
      y[b][r] = s;
    }
  }
  for (size_t b = 0; b < S.members.size(); ++b)  // ascending block order
    for (size_t r = 0; r < S.members[b].size(); ++r) z[S.members[b][r]] += y[b][r];
  for (int64_t i = 0; i < n; ++i)
    if (S.cover[i]) z[i] /= static_cast<double>(S.cover[i]);
}

double g_solver_setup_seconds = 0.0;

// Coarse space of the SCHWARZ_COARSE preconditioner: the aggregate of each row
// (-1: none), set by orc_set_coarse before a solve.  Ac = P^T A P is summed row
// by row in index order, inverted by Gauss-Jordan with partial pivoting (lowest
// index on ties), and applied additively: out += P Ac^-1 P^T in.  (The library
// forms the same operator with its own summation order.)
struct CoarseOracle {
  std::vector<int32_t> agg;
  int count = 0;
  std::vector<double> ainv;  // count x count, row-major
  bool ready = false;
};
CoarseOracle g_coarse;

bool coarse_build(const lsmps_csr& a, CoarseOracle& C) {
  const int m = C.count;
  const int64_t n = static_cast<int64_t>(a.rows);
  std::vector<double> w(static_cast<size_t>(m) * m, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t I = C.agg[i];
    if (I < 0) continue;
    for (int32_t c = a.row_offsets[i]; c < a.row_offsets[i + 1]; ++c) {
      const int32_t J = C.agg[a.col_indices[c]];
      if (J >= 0) w[static_cast<size_t>(I) * m + J] += a.values[c];
    }
  }
  std::vector<int32_t> perm(m);
  std::vector<double> colk(m);
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = std::fabs(w[static_cast<size_t>(k) * m + k]);
    for (int r = k + 1; r < m; ++r) {
      const double v = std::fabs(w[static_cast<size_t>(r) * m + k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (!(best > 0.0) || !std::isfinite(best)) return false;
    perm[k] = p;
    if (p != k)
      for (int j = 0; j < m; ++j)
        std::swap(w[static_cast<size_t>(k) * m + j], w[static_cast<size_t>(p) * m + j]);
    double* rk = w.data() + static_cast<size_t>(k) * m;
    const double inv = 1.0 / rk[k];
    rk[k] = 1.0;
    for (int j = 0; j < m; ++j) rk[j] *= inv;
    for (int r = 0; r < m; ++r) colk[r] = r == k ? 0.0 : w[static_cast<size_t>(r) * m + k];
#pragma omp parallel for schedule(static)
    for (int r = 0; r < m; ++r) {
      if (r == k) continue;
      double* rr = w.data() + static_cast<size_t>(r) * m;
      const double f = colk[r];
      for (int j = 0; j < m; ++j) rr[j] = j == k ? -f * rk[k] : rr[j] - f * rk[j];
    }
  }
  for (int k = m - 1; k >= 0; --k)  // undo the pivoting as column swaps
    if (perm[k] != k)
      for (int r = 0; r < m; ++r)
        std::swap(w[static_cast<size_t>(r) * m + k], w[static_cast<size_t>(r) * m + perm[k]]);
  C.ainv.swap(w);
  return true;
}

void coarse_apply_add(const CoarseOracle& C, const double* in, double* out, int64_t n) {
  const int m = C.count;
  std::vector<double> rc(m, 0.0), yc(m, 0.0);
  for (int64_t i = 0; i < n; ++i)
    if (C.agg[i] >= 0) rc[C.agg[i]] += in[i];
#pragma omp parallel for schedule(static)
  for (int I = 0; I < m; ++I) {
    double t = 0.0;
    for (int J = 0; J < m; ++J) t += C.ainv[static_cast<size_t>(I) * m + J] * rc[J];
    yc[I] = t;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    if (C.agg[i] >= 0) out[i] += yc[C.agg[i]];
}

int ppe_solve(const lsmps_csr& a, const double* b, double* x, const lsmps_solver_options& opt,
              lsmps_solve_report* rep) {
  const int64_t n = static_cast<int64_t>(a.rows);
  lsmps_solve_report report{};
  report.method = opt.method;
  if (n == 0) {
    report.converged = 1;
    if (rep) *rep = report;
    return LSMPS_OK;
  }
  std::vector<double> dinv(n), r(n), z(n), p(n), q(n), s(n), t(n), rhat(n), phat(n), shat(n);
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t c = a.row_offsets[i]; c < a.row_offsets[i + 1]; ++c)
      if (a.col_indices[c] == i) d += a.values[c];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    x[i] = 0.0;
  }
  const double bnorm = std::sqrt(dot(b, b, n));
  report.rhs_norm = bnorm;
  if (bnorm == 0.0) {
    report.converged = 1;
    if (rep) *rep = report;
    return LSMPS_OK;
  }
  const double target = opt.tolerance * bnorm;
  double rnorm = true_residual(a, b, x, r.data());
  int it = 0;
  const int check = opt.check_interval > 0 ? opt.check_interval : 1;
  if (opt.method == LSMPS_SOLVER_IDRS) {
    // IDR(s) with biorthogonalisation (Sonneveld & van Gijzen 2008; van Gijzen &
    // Sonneveld 2011).  Shadow space: P_0 = r_0, P_q (q >= 1) = hashed signs.
    // s = 0: 14 when the coarse space is requested, 16 otherwise
    const bool coarse_requested = opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE &&
                                  g_coarse.count > 1 &&
                                  static_cast<int64_t>(g_coarse.agg.size()) == n;
    int s = opt.shadow_dimension > 0 ? opt.shadow_dimension : (coarse_requested ? 14 : 16);
    if (s > 16) s = 16;
    if (s > n) s = static_cast<int>(n);
    auto sign = [](uint32_t row, int q) {
      uint32_t h = row * 0x9E3779B1u + static_cast<uint32_t>(q) * 0x85EBCA77u + 0x5eedu;
      h ^= h >> 16;
      h *= 0x7FEB352Du;
      h ^= h >> 15;
      h *= 0x846CA68Bu;
      h ^= h >> 16;
      return (h & 1u) ? 1.0 : -1.0;
    };
    std::vector<double> p0(r), G(static_cast<size_t>(s) * n, 0.0), U(static_cast<size_t>(s) * n, 0.0);
    std::vector<double> M(static_cast<size_t>(s) * s, 0.0), f(s), c(s), v(n), tv(n), d(s);
    // M^-1: Jacobi, or Schwarz blocks on the wrong-sign rows plus Jacobi elsewhere
    SchwarzBlocks blocks;
    const bool use_schwarz = opt.preconditioner == LSMPS_PRECOND_SCHWARZ ||
                             opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE;
    {
      const double t0 = omp_get_wtime();
      if (use_schwarz) schwarz_build(a, blocks);
      g_coarse.ready = false;
      if (opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE && g_coarse.count > 1 &&
          static_cast<int64_t>(g_coarse.agg.size()) == n)
        g_coarse.ready = coarse_build(a, g_coarse);
      report.coarse_unknowns = g_coarse.ready ? g_coarse.count : 0;
      g_solver_setup_seconds = omp_get_wtime() - t0;
      report.schwarz_blocks = static_cast<int32_t>(blocks.members.size());
      for (const auto& m : blocks.members) report.schwarz_rows += static_cast<int32_t>(m.size());
    }
    std::vector<double> w(n);
    auto precondition = [&](const double* in, double* out) {
      if (use_schwarz) {
        schwarz_apply(blocks, dinv, in, out, n);
      } else {
        for (int64_t i = 0; i < n; ++i) out[i] = in[i] * dinv[i];
      }
      if (g_coarse.ready) coarse_apply_add(g_coarse, in, out, n);
    };
    auto pdot = [&](int q, const double* w) {
      double acc = 0.0;
      if (q == 0) {
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (int64_t i = 0; i < n; ++i) acc += p0[i] * w[i];
      } else {
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (int64_t i = 0; i < n; ++i) acc += sign(static_cast<uint32_t>(i), q) * w[i];
      }
      return acc;
    };
    auto reset = [&]() {
      std::fill(G.begin(), G.end(), 0.0);
      std::fill(U.begin(), U.end(), 0.0);
      std::fill(M.begin(), M.end(), 0.0);
      for (int i = 0; i < s; ++i) M[i * s + i] = 1.0;
      p0 = r;
    };
    reset();
    double om = 1.0;
    while (rnorm > target && it < opt.max_iterations) {
      bool breakdown = false;
      for (int q = 0; q < s; ++q) f[q] = pdot(q, r.data());
      for (int k = 0; k < s && !breakdown; ++k) {
        for (int i = k; i < s; ++i) {  // c = M[k:,k:]^-1 f[k:]
          double t = f[i];
          for (int l = k; l < i; ++l) t -= M[i * s + l] * c[l];
          c[i] = t / M[i * s + i];
        }
        double* Uk = U.data() + static_cast<size_t>(k) * n;
        double* Gk = G.data() + static_cast<size_t>(k) * n;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
          double vv = r[i];
          for (int q = k; q < s; ++q) vv -= c[q] * G[static_cast<size_t>(q) * n + i];
          v[i] = vv;
        }
        precondition(v.data(), w.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
          double u = 0.0;
          for (int q = k; q < s; ++q) u += c[q] * U[static_cast<size_t>(q) * n + i];
          Uk[i] = u + om * w[i];
        }
        spmv(a, Uk, Gk);
        for (int q = 0; q < s; ++q) d[q] = pdot(q, Gk);
        double alpha[16];
        for (int i = 0; i < k; ++i) {
          double t = d[i];
          for (int l = 0; l < i; ++l) t -= alpha[l] * M[i * s + l];
          alpha[i] = t / M[i * s + i];
        }
        for (int i = k; i < s; ++i) {
          double t = d[i];
          for (int l = 0; l < k; ++l) t -= alpha[l] * M[i * s + l];
          M[i * s + k] = t;
        }
        if (M[k * s + k] == 0.0) {
          breakdown = true;
          break;
        }
        const double beta = f[k] / M[k * s + k];
        double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
          double g = Gk[i], u = Uk[i];
          for (int q = 0; q < k; ++q) {
            g -= alpha[q] * G[static_cast<size_t>(q) * n + i];
            u -= alpha[q] * U[static_cast<size_t>(q) * n + i];
          }
          Gk[i] = g;
          Uk[i] = u;
          r[i] -= beta * g;
          x[i] += beta * u;
          rr += r[i] * r[i];
        }
        rnorm = std::sqrt(rr);
        for (int i = k + 1; i < s; ++i) f[i] -= beta * M[i * s + k];
        ++it;
      }
      if (!breakdown) {
        precondition(r.data(), v.data());
        spmv(a, v.data(), tv.data());
        const double tr = dot(tv.data(), r.data(), n);
        const double tt = dot(tv.data(), tv.data(), n);
        if (tt > 0.0 && rnorm > 0.0) {
          om = tr / tt;
          const double rho = std::fabs(tr) / (std::sqrt(tt) * rnorm);
          if (rho < 0.7) om *= 0.7 / rho;
        } else {
          om = 0.0;
          breakdown = rnorm > 0.0;
        }
        for (int64_t i = 0; i < n; ++i) {
          x[i] += om * v[i];
          r[i] -= om * tv[i];
        }
        ++it;
        rnorm = std::sqrt(dot(r.data(), r.data(), n));
      }
      if (rnorm <= target) {
        rnorm = true_residual(a, b, x, r.data());
        ++report.refinements;
      }
      if (breakdown) {
        reset();
        om = 1.0;
      }
      if (!std::isfinite(rnorm)) break;
    }
  } else if (opt.method == LSMPS_SOLVER_BICGSTAB) {
    // van der Vorst (1992), right Jacobi preconditioning.
    rhat = r;
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    std::fill(p.begin(), p.end(), 0.0);
    std::fill(q.begin(), q.end(), 0.0);  // q holds v
    while (it < opt.max_iterations && rnorm > target) {
      const double rho_new = dot(rhat.data(), r.data(), n);
      if (rho_new == 0.0) {
        rhat = r;
        rho = 1.0; alpha = 1.0; omega = 1.0;
        std::fill(p.begin(), p.end(), 0.0);
        std::fill(q.begin(), q.end(), 0.0);
        continue;
      }
      const double beta = (rho_new / rho) * (alpha / omega);
      rho = rho_new;
      for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * q[i]);
      for (int64_t i = 0; i < n; ++i) phat[i] = p[i] * dinv[i];
      spmv(a, phat.data(), q.data());
      alpha = rho / dot(rhat.data(), q.data(), n);
      for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * q[i];
      for (int64_t i = 0; i < n; ++i) shat[i] = s[i] * dinv[i];
      spmv(a, shat.data(), t.data());
      const double tt = dot(t.data(), t.data(), n);
      omega = tt > 0.0 ? dot(t.data(), s.data(), n) / tt : 0.0;
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * phat[i] + omega * shat[i];
        r[i] = s[i] - omega * t[i];
      }
      ++it;
      rnorm = std::sqrt(dot(r.data(), r.data(), n));
      if (rnorm <= target || it % check == 0) {
        rnorm = true_residual(a, b, x, r.data());
        ++report.refinements;
      }
      if (omega == 0.0) break;
    }
  } else {
    // Hestenes-Stiefel CG with Jacobi preconditioning; the recurrence residual is
    // replaced by the true FP64 residual at every convergence check.
    for (int64_t i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
    p = z;
    double rz = dot(r.data(), z.data(), n);
    while (it < opt.max_iterations && rnorm > target) {
      spmv(a, p.data(), q.data());
      const double pq = dot(p.data(), q.data(), n);
      if (pq == 0.0) break;
      const double alpha = rz / pq;
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      ++it;
      rnorm = std::sqrt(dot(r.data(), r.data(), n));
      if (rnorm <= target || it % check == 0) {
        rnorm = true_residual(a, b, x, r.data());
        ++report.refinements;
      }
      for (int64_t i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
      const double rz_new = dot(r.data(), z.data(), n);
      const double beta = rz_new / rz;
      rz = rz_new;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
  }
  rnorm = true_residual(a, b, x, r.data());
  report.iterations = it;
  report.relative_residual = rnorm / bnorm;
  report.converged = rnorm <= target ? 1 : 0;
  if (opt.clamp_negative) {
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] < 0.0 ? 0.0 : x[i];  // clamp_pressure_kernel (ppe_matrix.cu:373-379)
  }
  if (rep) *rep = report;
  if (!report.converged)
    return fail(LSMPS_ERR_NOT_CONVERGED, "PPE solver did not reach the tolerance");
  return LSMPS_OK;
}

// ---------------------------------------------------------------------------
// pressure correction (src/correction/pressure_correction.cu)

// clamp_fluid_pressure_kernel (pressure_correction.cu:188-198)
void clamp_pressure(const lsmps_fluid_soa& fluid, double* p) {
  for (size_t i = 0; i < fluid.count; ++i) {
    const double v = p[i] < 0.0 ? 0.0 : p[i];
    p[i] = v;
    if (fluid.pressure && fluid.pressure != p) fluid.pressure[i] = v;
  }
}

// compute_pressure_gradient_kernel (pressure_correction.cu:200-276)
void pressure_gradient(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const double* p, const double* pre_inverse,
                       const lsmps_vector3& g) {
  const int64_t n = static_cast<int64_t>(fluid.count);
  const double re = cfg.support_radius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (fluid.surface_type && fluid.surface_type[i] == LSMPS_SURFACE_SPLASH) {
      g.x[i] = 0.0; g.y[i] = 0.0; g.z[i] = 0.0;
      continue;
    }
    const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
    double rhs[9] = {0}, b[9], q[9];
    for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
      const int32_t j = fn.indices[c];
      const double dx = fluid.x[j] - xi, dy = fluid.y[j] - yi, dz = fluid.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = correction_weight_fn(distance, re, cfg.weight_kernel);
      if (w <= 0.0) continue;
      type_a_basis_mul(dx, dy, dz, re, b);
      const double dp = p[j] - p[i];
      for (int r = 0; r < 9; ++r) rhs[r] += w * b[r] * dp;
    }
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double dx = walls.x[j] - xi, dy = walls.y[j] - yi, dz = walls.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = cfg.lsmps_wall_weight_scale * correction_weight_fn(distance, re, cfg.weight_kernel);
      if (w <= 0.0) continue;
      wall_neumann_vector(dx, dy, dz, walls.normal_x[j], walls.normal_y[j], walls.normal_z[j],
                          LSMPS_BASIS_TYPE_A, re, q);
      const double ng = cfg.gravity[0] * walls.normal_x[j] + cfg.gravity[1] * walls.normal_y[j] +
                        cfg.gravity[2] * walls.normal_z[j];
      const double wall_rhs = re * cfg.density * ng;
      for (int r = 0; r < 9; ++r) rhs[r] += w * q[r] * wall_rhs;
    }
    const double* inv = pre_inverse + i * 81;
    auto coeff = [&](int row) {
      double s = 0.0;
      for (int c = 0; c < 9; ++c) s += inv[row + c * 9] * rhs[c];
      return s;
    };
    g.x[i] = coeff(0) / re;
    g.y[i] = coeff(1) / re;
    g.z[i] = coeff(2) / re;
  }
}

// compute_particle_shifting_kernel (pressure_correction.cu:278-367)
void particle_shifting(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const lsmps_vector3& ps) {
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t st = fluid.surface_type ? fluid.surface_type[i] : 0;
    if (cfg.ps_displacement_scale <= 0.0 || st == LSMPS_SURFACE_SPLASH) {
      ps.x[i] = 0.0; ps.y[i] = 0.0; ps.z[i] = 0.0;
      continue;
    }
    const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
    const double min_distance = cfg.ps_min_distance_ratio * cfg.particle_spacing;
    const double max_shift = cfg.ps_max_displacement_ratio * cfg.particle_spacing;
    double sx = 0, sy = 0, sz = 0, nx = 0, ny = 0, nz = 0;
    for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
      const int32_t j = fn.indices[c];
      const double dx = xi - fluid.x[j], dy = yi - fluid.y[j], dz = zi - fluid.z[j];
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (distance <= 0.0 || distance >= min_distance) continue;
      const double push = cfg.ps_displacement_scale * (min_distance - distance) / distance;
      sx += push * dx;
      sy += push * dy;
      sz += push * dz;
      nx += dx / distance;
      ny += dy / distance;
      nz += dz / distance;
    }
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double dx = xi - walls.x[j], dy = yi - walls.y[j], dz = zi - walls.z[j];
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (distance <= 0.0 || distance >= min_distance) continue;
      const double push = cfg.ps_displacement_scale * (min_distance - distance);
      sx += push * walls.normal_x[j];
      sy += push * walls.normal_y[j];
      sz += push * walls.normal_z[j];
    }
    if (st == LSMPS_SURFACE_SURFACE || st == LSMPS_SURFACE_SPLASH) {
      const double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (nn > 0.0) {
        nx /= nn; ny /= nn; nz /= nn;
        const double comp = sx * nx + sy * ny + sz * nz;
        sx -= comp * nx;
        sy -= comp * ny;
        sz -= comp * nz;
      }
    }
    const double norm = std::sqrt(sx * sx + sy * sy + sz * sz);
    if (norm > max_shift && norm > 0.0) {
      const double scale = max_shift / norm;
      sx *= scale; sy *= scale; sz *= scale;
    }
    ps.x[i] = sx; ps.y[i] = sy; ps.z[i] = sz;
  }
}

// update_velocity_position_kernel (pressure_correction.cu:369-402)
void update_velocity_position(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                              const lsmps_fluid_soa& tmp, const lsmps_vector3& g,
                              const lsmps_vector3& ps) {
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const bool splash = fluid.surface_type && fluid.surface_type[i] == LSMPS_SURFACE_SPLASH;
    const double ovx = fluid.vx[i], ovy = fluid.vy[i], ovz = fluid.vz[i];
    double nvx = tmp.vx[i], nvy = tmp.vy[i], nvz = tmp.vz[i];
    if (!splash) {
      const double s = cfg.time_step / cfg.density;
      nvx -= s * g.x[i];
      nvy -= s * g.y[i];
      nvz -= s * g.z[i];
    }
    fluid.x[i] += 0.5 * cfg.time_step * (ovx + nvx) + ps.x[i];
    fluid.y[i] += 0.5 * cfg.time_step * (ovy + nvy) + ps.y[i];
    fluid.z[i] += 0.5 * cfg.time_step * (ovz + nvz) + ps.z[i];
    fluid.vx[i] = nvx;
    fluid.vy[i] = nvy;
    fluid.vz[i] = nvz;
  }
}

// anti_penetration_kernel (pressure_correction.cu:404-444)
void anti_penetration(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                      const lsmps_wall_soa& walls, const lsmps_neighbor_list& wn) {
  const double clearance = cfg.wall_clearance_ratio * cfg.particle_spacing;
  if (clearance <= 0.0) return;
  const int64_t n = static_cast<int64_t>(fluid.count);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t c = wn.offsets[i]; c < wn.offsets[i + 1]; ++c) {
      const int32_t j = wn.indices[c];
      const double nx = walls.normal_x[j], ny = walls.normal_y[j], nz = walls.normal_z[j];
      const double dx = fluid.x[i] - walls.x[j];
      const double dy = fluid.y[i] - walls.y[j];
      const double dz = fluid.z[i] - walls.z[j];
      const double sd = dx * nx + dy * ny + dz * nz;
      if (sd >= clearance) continue;
      const double corr = clearance - sd;
      fluid.x[i] += corr * nx;
      fluid.y[i] += corr * ny;
      fluid.z[i] += corr * nz;
      const double vn = fluid.vx[i] * nx + fluid.vy[i] * ny + fluid.vz[i] * nz;
      if (vn < 0.0) {
        fluid.vx[i] -= vn * nx;
        fluid.vy[i] -= vn * ny;
        fluid.vz[i] -= vn * nz;
      }
    }
  }
}

// smooth_velocity_kernel + copy_smoothed_velocity_kernel (pressure_correction.cu:446-515)
void smooth_velocity(const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_neighbor_list& fn, const lsmps_vector3& sm) {
  const int64_t n = static_cast<int64_t>(fluid.count);
  auto splash = [&](int64_t k) {
    return fluid.surface_type && fluid.surface_type[k] == LSMPS_SURFACE_SPLASH;
  };
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (cfg.velocity_smoothing_strength <= 0.0 || splash(i)) {
      sm.x[i] = fluid.vx[i]; sm.y[i] = fluid.vy[i]; sm.z[i] = fluid.vz[i];
      continue;
    }
    double ws = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    const double xi = fluid.x[i], yi = fluid.y[i], zi = fluid.z[i];
    for (int32_t c = fn.offsets[i]; c < fn.offsets[i + 1]; ++c) {
      const int32_t j = fn.indices[c];
      if (splash(j)) continue;
      const double dx = fluid.x[j] - xi, dy = fluid.y[j] - yi, dz = fluid.z[j] - zi;
      const double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = correction_weight_fn(distance, cfg.support_radius, cfg.weight_kernel);
      if (w <= 0.0) continue;
      ws += w;
      ax += w * fluid.vx[j];
      ay += w * fluid.vy[j];
      az += w * fluid.vz[j];
    }
    if (ws <= 0.0) {
      sm.x[i] = fluid.vx[i]; sm.y[i] = fluid.vy[i]; sm.z[i] = fluid.vz[i];
      continue;
    }
    double alpha = cfg.velocity_smoothing_strength;
    alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
    sm.x[i] = (1.0 - alpha) * fluid.vx[i] + alpha * ax / ws;
    sm.y[i] = (1.0 - alpha) * fluid.vy[i] + alpha * ay / ws;
    sm.z[i] = (1.0 - alpha) * fluid.vz[i] + alpha * az / ws;
  }
  for (int64_t i = 0; i < n; ++i) {
    fluid.vx[i] = sm.x[i];
    fluid.vy[i] = sm.y[i];
    fluid.vz[i] = sm.z[i];
  }
}

}  // namespace

// ===========================================================================
// extern "C" surface of the oracle (host pointers everywhere).
extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

int orc_set_threads(int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  return omp_get_max_threads();
}

int orc_build_cell_list(size_t count, const double* x, const double* y, const double* z,
                        const lsmps_cell_grid* grid, const lsmps_cell_list* cells) {
  return build_cell_list(count, x, y, z, *grid, *cells);
}

// build_neighbor_lists (neighbor_search.cu:345-367)
int orc_build_neighbor_lists(const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                             const lsmps_config* cfg, const lsmps_cell_list* fc,
                             const lsmps_cell_list* wc, const lsmps_neighbor_list* fn,
                             const lsmps_neighbor_list* wn) {
  lsmps_cell_grid grid;
  for (int k = 0; k < 3; ++k) {
    grid.origin[k] = cfg->cell_origin[k];
    grid.dims[k] = cfg->cell_dims[k];
  }
  grid.cell_size = cfg->cell_size;
  const double radius = cfg->support_radius;
  if (radius <= 0.0) return fail(LSMPS_ERR_INVALID_ARGUMENT, "Neighbor search radius must be positive");
  if (radius > grid.cell_size)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "Neighbor search currently requires radius <= cell_size");
  int rc = build_cell_list(fluid->count, fluid->x, fluid->y, fluid->z, grid, *fc);
  if (rc) return rc;
  rc = build_cell_list(walls->count, walls->x, walls->y, walls->z, grid, *wc);
  if (rc) return rc;
  rc = build_one_neighbor_list(*fluid, fluid->count, fluid->x, fluid->y, fluid->z, *fc, grid,
                               radius, true, *fn);
  if (rc) return rc;
  return build_one_neighbor_list(*fluid, walls->count, walls->x, walls->y, walls->z, *wc, grid,
                                 radius, false, *wn);
}

double orc_reference_number_density(double l0, double re, int kernel) {
  return reference_number_density(l0, re, kernel);
}

int orc_classify_surface(const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                         const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                         const lsmps_config* cfg, const lsmps_surface_diagnostics* diag) {
  return classify_surface(*fluid, *walls, *fn, *wn, *cfg, diag);
}

int orc_virtual_light(const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                      const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                      const lsmps_config* cfg, int32_t* open_count, double* open_fraction) {
  return virtual_light(*fluid, *walls, *fn, *wn, *cfg, open_count, open_fraction);
}

// One moment kind: matrices (may be NULL) and inverses are N*n*n column-major.
int orc_moment_prepare(int kind, const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                       const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                       const lsmps_config* cfg, double* matrices, double* inverses, double* trace,
                       double* reg, int32_t* info) {
  return moment_prepare(kind, *fluid, *walls, *fn, *wn, *cfg, matrices, inverses, trace, reg, info);
}

// compute_temporary_velocity (explicit_update.cu:366-399) given the velocity Type-A inverse.
int orc_provision(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                  const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                  const lsmps_neighbor_list* wn, const double* vel_inverse,
                  const lsmps_fluid_soa* tmp, const lsmps_wall_soa* tmpw, double* lap_x,
                  double* lap_y, double* lap_z) {
  std::vector<double> lx(fluid->count), ly(fluid->count), lz(fluid->count);
  velocity_laplacian(*cfg, *fluid, *walls, *fn, *wn, vel_inverse, lx.data(), ly.data(), lz.data());
  combine_velocity(*cfg, *fluid, lx.data(), ly.data(), lz.data(), *tmp);
  combine_wall_velocity(*cfg, *walls, *tmpw);
  if (lap_x) std::copy(lx.begin(), lx.end(), lap_x);
  if (lap_y) std::copy(ly.begin(), ly.end(), lap_y);
  if (lap_z) std::copy(lz.begin(), lz.end(), lap_z);
  return LSMPS_OK;
}

int orc_ppe_assemble(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                     const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                     const lsmps_neighbor_list* wn, const lsmps_fluid_soa* tmp,
                     const lsmps_wall_soa* tmpw, const double* vel_inverse,
                     const double* pre_inverse, const lsmps_ppe_workspace* ws) {
  return ppe_assemble(*cfg, *fluid, *walls, *fn, *wn, *tmp, *tmpw, vel_inverse, pre_inverse, *ws);
}

int orc_ppe_solve(const lsmps_csr* a, const double* b, double* x,
                  const lsmps_solver_options* opt, lsmps_solve_report* rep) {
  return ppe_solve(*a, b, x, *opt, rep);
}

// Seconds spent building the preconditioner in the last IDR(s) solve.
double orc_solver_setup_seconds(void) { return g_solver_setup_seconds; }

// Aggregation used by SCHWARZ_COARSE solves from now on (agg NULL or count < 2:
// none).
int orc_set_coarse(const int32_t* agg, int64_t n, int32_t count) {
  g_coarse = CoarseOracle{};
  if (!agg || n <= 0 || count < 2) return LSMPS_OK;
  g_coarse.agg.assign(agg, agg + n);
  g_coarse.count = count;
  return LSMPS_OK;
}

// z = (P Ac^-1 P^T) v for the current aggregation and matrix `a` (tests).
int orc_coarse_apply(const lsmps_csr* a, const double* v, double* z) {
  const int64_t n = static_cast<int64_t>(a->rows);
  if (g_coarse.count < 2 || static_cast<int64_t>(g_coarse.agg.size()) != n)
    return fail(LSMPS_ERR_INVALID_ARGUMENT, "no coarse aggregation for this matrix");
  if (!coarse_build(*a, g_coarse)) return fail(LSMPS_ERR_NUMERICAL, "singular coarse matrix");
  for (int64_t i = 0; i < n; ++i) z[i] = 0.0;
  coarse_apply_add(g_coarse, v, z, n);
  return LSMPS_OK;
}

// z = M^-1 v with the Schwarz preconditioner of `a` (built on each call).
// Returns the number of blocks.
int orc_schwarz_apply(const lsmps_csr* a, const double* v, double* z, int* blocks_out,
                      int* rows_out) {
  const int64_t n = static_cast<int64_t>(a->rows);
  std::vector<double> dinv(n);
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t c = a->row_offsets[i]; c < a->row_offsets[i + 1]; ++c)
      if (a->col_indices[c] == i) d += a->values[c];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
  SchwarzBlocks S;
  schwarz_build(*a, S);
  schwarz_apply(S, dinv, v, z, n);
  int rows = 0;
  for (const auto& m : S.members) rows += static_cast<int>(m.size());
  if (blocks_out) *blocks_out = static_cast<int>(S.members.size());
  if (rows_out) *rows_out = rows;
  return LSMPS_OK;
}

// DevicePressureCorrection::apply (pressure_correction.cu:598-645)
int orc_correction_apply(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                         const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                         const lsmps_neighbor_list* wn, const lsmps_fluid_soa* tmp,
                         double* pressure, const double* pre_inverse, const lsmps_vector3* grad,
                         const lsmps_vector3* ps, const lsmps_vector3* smoothed) {
  clamp_pressure(*fluid, pressure);
  pressure_gradient(*cfg, *fluid, *walls, *fn, *wn, pressure, pre_inverse, *grad);
  particle_shifting(*cfg, *fluid, *walls, *fn, *wn, *ps);
  update_velocity_position(*cfg, *fluid, *tmp, *grad, *ps);
  anti_penetration(*cfg, *fluid, *walls, *wn);
  smooth_velocity(*cfg, *fluid, *fn, *smoothed);
  return LSMPS_OK;
}

int orc_pressure_gradient(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                          const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                          const lsmps_neighbor_list* wn, const double* pressure,
                          const double* pre_inverse, const lsmps_vector3* grad) {
  pressure_gradient(*cfg, *fluid, *walls, *fn, *wn, pressure, pre_inverse, *grad);
  return LSMPS_OK;
}

int orc_particle_shifting(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                          const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                          const lsmps_neighbor_list* wn, const lsmps_vector3* ps) {
  particle_shifting(*cfg, *fluid, *walls, *fn, *wn, *ps);
  return LSMPS_OK;
}

int orc_anti_penetration(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                         const lsmps_wall_soa* walls, const lsmps_neighbor_list* wn) {
  anti_penetration(*cfg, *fluid, *walls, *wn);
  return LSMPS_OK;
}

int orc_smooth_velocity(const lsmps_config* cfg, const lsmps_fluid_soa* fluid,
                        const lsmps_neighbor_list* fn, const lsmps_vector3* smoothed) {
  smooth_velocity(*cfg, *fluid, *fn, *smoothed);
  return LSMPS_OK;
}

// Stage timings of the last orc_step call (seconds): geometry, moments,
// provision, assembly, solve, correction.
static double g_stage_seconds[6];
const double* orc_stage_seconds(void) { return g_stage_seconds; }

// One full step, in the driver's stage order (simulation_driver.cu:328,336-393):
// geometry (neighbor CSR + surface labels), provision (with the velocity and
// pressure Type-A moment matrices), PPE assembly, PPE solve + clamp, correction.
// All work buffers are allocated internally; outputs the caller wants to
// inspect may be passed (any may be NULL).
int orc_step(const lsmps_config* cfg, const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
             size_t max_neighbors, const lsmps_solver_options* opt, lsmps_solve_report* rep,
             int32_t* out_fluid_offsets, int32_t* out_fluid_indices, int32_t* out_wall_offsets,
             int32_t* out_wall_indices, double* out_rhs, double* out_grad_x, double* out_grad_y,
             double* out_grad_z) {
  const size_t nf = fluid->count, nw = walls->count;
  const size_t cells = static_cast<size_t>(cfg->cell_dims[0]) * cfg->cell_dims[1] * cfg->cell_dims[2];
  std::vector<int32_t> fc_ids(nf), fc_sorted(nf), fc_b(cells), fc_e(cells);
  std::vector<int32_t> wc_ids(nw), wc_sorted(nw), wc_b(cells), wc_e(cells);
  lsmps_cell_list fc{nf, cells, fc_ids.data(), fc_sorted.data(), fc_b.data(), fc_e.data()};
  lsmps_cell_list wc{nw, cells, wc_ids.data(), wc_sorted.data(), wc_b.data(), wc_e.data()};
  const size_t cap = nf * max_neighbors;
  std::vector<int32_t> f_off(nf + 1), f_idx(cap), w_off(nf + 1), w_idx(cap);
  lsmps_neighbor_list fn{nf, cap, f_off.data(), f_idx.data()};
  lsmps_neighbor_list wn{nf, cap, w_off.data(), w_idx.data()};
  double t0 = omp_get_wtime();
  int rc = orc_build_neighbor_lists(fluid, walls, cfg, &fc, &wc, &fn, &wn);
  if (rc) return rc;
  rc = classify_surface(*fluid, *walls, fn, wn, *cfg, nullptr);
  if (rc) return rc;
  double t1 = omp_get_wtime();
  std::vector<double> vinv(nf * 81), pinv(nf * 81), tr(nf), rg(nf);
  std::vector<int32_t> info(nf);
  moment_prepare(LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A, *fluid, *walls, fn, wn, *cfg, nullptr,
                 vinv.data(), tr.data(), rg.data(), info.data());
  moment_prepare(LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A, *fluid, *walls, fn, wn, *cfg, nullptr,
                 pinv.data(), tr.data(), rg.data(), info.data());
  double t2 = omp_get_wtime();
  std::vector<double> tvx(nf), tvy(nf), tvz(nf), twx(nw), twy(nw), twz(nw);
  lsmps_fluid_soa tmp{nf, nullptr, nullptr, nullptr, tvx.data(), tvy.data(), tvz.data(), nullptr, nullptr};
  lsmps_wall_soa tmpw{nw, nullptr, nullptr, nullptr, twx.data(), twy.data(), twz.data(), nullptr, nullptr, nullptr};
  orc_provision(cfg, fluid, walls, &fn, &wn, vinv.data(), &tmp, &tmpw, nullptr, nullptr, nullptr);
  double t3 = omp_get_wtime();
  const size_t nnz_cap = nf + cap;
  std::vector<int32_t> rows(nf + 1), cols(std::max<size_t>(nnz_cap, 1));
  std::vector<double> vals(std::max<size_t>(nnz_cap, 1)), rhs(nf), pres(nf), div(nf), lap(nf);
  lsmps_ppe_workspace ws{{nf, nf, nnz_cap, rows.data(), cols.data(), vals.data()},
                         rhs.data(), pres.data(), div.data(), lap.data()};
  rc = ppe_assemble(*cfg, *fluid, *walls, fn, wn, tmp, tmpw, vinv.data(), pinv.data(), ws);
  if (rc) return rc;
  double t4 = omp_get_wtime();
  lsmps_csr a = ws.matrix;
  int solve_rc = ppe_solve(a, rhs.data(), pres.data(), *opt, rep);
  double t5 = omp_get_wtime();
  std::vector<double> gx(nf), gy(nf), gz(nf), px(nf), py(nf), pz(nf), sx(nf), sy(nf), sz(nf);
  lsmps_vector3 g{nf, gx.data(), gy.data(), gz.data()};
  lsmps_vector3 ps{nf, px.data(), py.data(), pz.data()};
  lsmps_vector3 sm{nf, sx.data(), sy.data(), sz.data()};
  orc_correction_apply(cfg, fluid, walls, &fn, &wn, &tmp, pres.data(), pinv.data(), &g, &ps, &sm);
  double t6 = omp_get_wtime();
  g_stage_seconds[0] = t1 - t0;
  g_stage_seconds[1] = t2 - t1;
  g_stage_seconds[2] = t3 - t2;
  g_stage_seconds[3] = t4 - t3;
  g_stage_seconds[4] = t5 - t4;
  g_stage_seconds[5] = t6 - t5;
  if (out_fluid_offsets) std::copy(f_off.begin(), f_off.end(), out_fluid_offsets);
  if (out_fluid_indices) std::copy(f_idx.begin(), f_idx.begin() + f_off[nf], out_fluid_indices);
  if (out_wall_offsets) std::copy(w_off.begin(), w_off.end(), out_wall_offsets);
  if (out_wall_indices) std::copy(w_idx.begin(), w_idx.begin() + w_off[nf], out_wall_indices);
  if (out_rhs) std::copy(rhs.begin(), rhs.end(), out_rhs);
  if (out_grad_x) std::copy(gx.begin(), gx.end(), out_grad_x);
  if (out_grad_y) std::copy(gy.begin(), gy.end(), out_grad_y);
  if (out_grad_z) std::copy(gz.begin(), gz.end(), out_grad_z);
  return solve_rc;
}

}  // extern "C"
