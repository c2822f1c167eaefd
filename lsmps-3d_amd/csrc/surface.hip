// Free-surface classification on gfx950 (Inner / NearSurface / Surface / Splash).
//
// Replaces classify_surface_particles (src/surface/surface_detection.cu:522-553).
// One thread per fluid particle walks its two CSR rows in CSR order, so the FP64
// sums are formed in the same order as the oracle; compiled with
// -ffp-contract=off, the labels are bit-identical.  rsqrt of the reference
// (surface_detection.cu:21-24) is computed as a correctly rounded 1/sqrt.
//
// The reference's in-place NearSurface expansion (expand_near_surface_kernel,
// :322-348) reads labels other threads are writing; here the primary labels go to
// a scratch array and the expansion reads only that array, so the pass is
// race-free by construction and gives the same deterministic result.
#include <cmath>

#include "common.hpp"
#include "device_math.hpp"
#include "stages.hpp"

namespace lsmps {
namespace {

struct SurfaceParams {
  double re;
  double n0;
  double ratio_threshold;
  double open_threshold;
  double aniso_threshold;
  double independence;
  int32_t splash_threshold;
  int32_t include_walls;
  int32_t kernel;
};

// safe_rsqrt (surface_detection.cu:21-24), as a correctly rounded reciprocal square root.
__device__ __forceinline__ double safe_inv_sqrt(double v) {
  const double eps = 1.0e-20;
  return 1.0 / sqrt(v > eps ? v : eps);
}

// accumulate_neighbor_moments (surface_detection.cu:33-59)
__device__ __forceinline__ void accumulate(double dx, double dy, double dz, const SurfaceParams& p,
                                           double& nd, double& sx, double& sy, double& sz,
                                           int& wc) {
  const double d2 = dx * dx + dy * dy + dz * dz;
  if (d2 <= 0.0) return;
  const double d = sqrt(d2);
  const double w = weight_of(d, p.re, p.kernel);
  if (w <= 0.0) return;
  const double inv = 1.0 / d;
  nd += w;
  sx += w * dx * inv;
  sy += w * dy * inv;
  sz += w * dz * inv;
  ++wc;
}

__global__ void __launch_bounds__(kBlock) classify_kernel(
    size_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ wx, const double* __restrict__ wy,
    const double* __restrict__ wz, const double* __restrict__ wnx, const double* __restrict__ wny,
    const double* __restrict__ wnz, const int32_t* __restrict__ f_off,
    const int32_t* __restrict__ f_idx, const int32_t* __restrict__ w_off,
    const int32_t* __restrict__ w_idx, SurfaceParams p, int32_t* __restrict__ labels,
    lsmps_surface_diagnostics diag) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t fb = f_off[i], fe = f_off[i + 1];
  const int32_t wb = w_off[i], we = w_off[i + 1];
  const int32_t fluid_count = fe - fb;
  const int32_t wall_count = we - wb;
  const int32_t classification_count = p.include_walls ? fluid_count + wall_count : fluid_count;
  double nd = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  int wc = 0;
  const double px = x[i], py = y[i], pz = z[i];
  // Four neighbours' loads are issued before their sums, which still run in
  // neighbour order (the labels depend on the exact rounding of these sums).
  // The indices come 16 at a time: each lane streams its own row of the CSR,
  // and a 64-byte piece read at once is fetched once, where four 16-byte
  // pieces read between gathers found their cache line evicted again and again.
  int32_t c = fb;
  for (; c + 16 <= fe; c += 16) {
    int32_t j[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) j[u] = f_idx[c + u];
#pragma unroll
    for (int h = 0; h < 16; h += 4) {
      double dx[4], dy[4], dz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        dx[u] = x[j[h + u]] - px;
        dy[u] = y[j[h + u]] - py;
        dz[u] = z[j[h + u]] - pz;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) accumulate(dx[u], dy[u], dz[u], p, nd, sx, sy, sz, wc);
    }
  }
  for (; c + 4 <= fe; c += 4) {
    int32_t j[4];
    double dx[4], dy[4], dz[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = f_idx[c + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dx[u] = x[j[u]] - px;
      dy[u] = y[j[u]] - py;
      dz[u] = z[j[u]] - pz;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) accumulate(dx[u], dy[u], dz[u], p, nd, sx, sy, sz, wc);
  }
  for (; c < fe; ++c) {
    const int32_t j = f_idx[c];
    accumulate(x[j] - px, y[j] - py, z[j] - pz, p, nd, sx, sy, sz, wc);
  }
  if (p.include_walls) {
    for (int32_t c = wb; c < we; ++c) {
      const int32_t j = w_idx[c];
      accumulate(wx[j] - px, wy[j] - py, wz[j] - pz, p, nd, sx, sy, sz, wc);
    }
  }
  double anisotropy = 0.0;
  if (wc > 0) anisotropy = sqrt(sx * sx + sy * sy + sz * sz) / static_cast<double>(wc);
  const double ratio = nd / p.n0;
  const double mx = -sx, my = -sy, mz = -sz;
  const double missing2 = mx * mx + my * my + mz * mz;
  double ax = mx, ay = my, az = mz;
  if (missing2 > 0.0 && p.include_walls) {
    // build_wall_basis (surface_detection.cu:151-197) + add_wall_basis_direction (:123-149)
    double bx[3], by[3], bz[3];
    int count = 0;
    for (int32_t c = wb; c < we; ++c) {
      const int32_t j = w_idx[c];
      const double dx = wx[j] - px, dy = wy[j] - py, dz = wz[j] - pz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= 0.0) continue;
      const double d = sqrt(d2);
      if (weight_of(d, p.re, p.kernel) <= 0.0) continue;
      const double nx = wnx[j], ny = wny[j], nz = wnz[j];
      const double nl2 = nx * nx + ny * ny + nz * nz;
      if (nl2 <= 0.0) continue;
      const double inl = safe_inv_sqrt(nl2);
      double rx = -nx * inl, ry = -ny * inl, rz = -nz * inl;
      for (int k = 0; k < count; ++k) {
        const double proj = rx * bx[k] + ry * by[k] + rz * bz[k];
        rx -= proj * bx[k];
        ry -= proj * by[k];
        rz -= proj * bz[k];
      }
      const double l2 = rx * rx + ry * ry + rz * rz;
      if (l2 <= p.independence * p.independence || count >= 3) continue;
      const double il = safe_inv_sqrt(l2);
      bx[count] = rx * il;
      by[count] = ry * il;
      bz[count] = rz * il;
      ++count;
    }
    for (int k = 0; k < count; ++k) {
      const double proj = ax * bx[k] + ay * by[k] + az * bz[k];
      if (proj > 0.0) {
        ax -= proj * bx[k];
        ay -= proj * by[k];
        az -= proj * bz[k];
      }
    }
  }
  const double air2 = ax * ax + ay * ay + az * az;
  const double air = sqrt(air2);
  const double missing = sqrt(missing2);
  const double open_ratio = missing > 0.0 ? air / missing : 0.0;
  const double air_aniso = wc > 0 ? air / static_cast<double>(wc) : 0.0;
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if (air2 > 0.0) {
    const double inv = safe_inv_sqrt(air2);
    nx = ax * inv;
    ny = ay * inv;
    nz = az * inv;
  }
  int32_t type = LSMPS_SURFACE_INNER;
  if (classification_count < p.splash_threshold) {
    type = LSMPS_SURFACE_SPLASH;
  } else if (ratio < p.ratio_threshold && open_ratio > p.open_threshold &&
             air_aniso > p.aniso_threshold) {
    type = LSMPS_SURFACE_SURFACE;
  }
  labels[i] = type;
  if (diag.fluid_neighbor_count) diag.fluid_neighbor_count[i] = fluid_count;
  if (diag.wall_neighbor_count) diag.wall_neighbor_count[i] = wall_count;
  if (diag.number_density) diag.number_density[i] = nd;
  if (diag.number_density_ratio) diag.number_density_ratio[i] = ratio;
  if (diag.anisotropy) diag.anisotropy[i] = anisotropy;
  if (diag.air_open_ratio) diag.air_open_ratio[i] = open_ratio;
  if (diag.air_anisotropy) diag.air_anisotropy[i] = air_aniso;
  if (diag.surface_normal_x) diag.surface_normal_x[i] = nx;
  if (diag.surface_normal_y) diag.surface_normal_y[i] = ny;
  if (diag.surface_normal_z) diag.surface_normal_z[i] = nz;
}

// expand_near_surface_kernel (surface_detection.cu:322-348), reading primary labels only.
// kExpandLanes adjacent lanes per particle, each testing neighbours l,
// l + kExpandLanes, ...: the result is an "any" over the neighbours, so the
// split does not change it.
constexpr int kExpandLanes = 8;

__global__ void __launch_bounds__(kBlock) expand_kernel(
    size_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int32_t* __restrict__ f_off,
    const int32_t* __restrict__ f_idx, const int32_t* __restrict__ primary, double r2,
    int32_t* __restrict__ labels) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / kExpandLanes;
  const int sub = static_cast<int>(t % kExpandLanes);
  if (i >= n) return;  // whole groups
  int32_t type = primary[i];
  if (type == LSMPS_SURFACE_INNER && r2 > 0.0) {
    const double px = x[i], py = y[i], pz = z[i];
    int near = 0;
    for (int32_t c = f_off[i] + sub; c < f_off[i + 1]; c += kExpandLanes) {
      const int32_t j = f_idx[c];
      if (primary[j] != LSMPS_SURFACE_SURFACE) continue;
      const double dx = x[j] - px, dy = y[j] - py, dz = z[j] - pz;
      if (dx * dx + dy * dy + dz * dz <= r2) {
        near = 1;
        break;
      }
    }
#pragma unroll
    for (int off = kExpandLanes / 2; off > 0; off >>= 1) near |= __shfl_xor(near, off, 64);
    if (near) type = LSMPS_SURFACE_NEAR_SURFACE;
  }
  if (sub == 0) labels[i] = type;
}

// The 14 virtual-light directions in the reference's order
// (virtual_light_direction, surface_detection.cu:350-383): the six axis
// directions, then the diagonals of the xy, xz and yz planes it lists.
constexpr int kLightDirections = 14;
constexpr unsigned kAllBlocked = (1u << kLightDirections) - 1u;
constexpr double kLight[kLightDirections][3] = {
    {1.0, 0.0, 0.0},
    {-1.0, 0.0, 0.0},
    {0.0, 1.0, 0.0},
    {0.0, -1.0, 0.0},
    {0.0, 0.0, 1.0},
    {0.0, 0.0, -1.0},
    {0.7071067811865475, 0.7071067811865475, 0.0},
    {0.7071067811865475, -0.7071067811865475, 0.0},
    {-0.7071067811865475, 0.7071067811865475, 0.0},
    {-0.7071067811865475, -0.7071067811865475, 0.0},
    {0.7071067811865475, 0.0, 0.7071067811865475},
    {0.7071067811865475, 0.0, -0.7071067811865475},
    {0.0, 0.7071067811865475, 0.7071067811865475},
    {0.0, 0.7071067811865475, -0.7071067811865475},
};

// Adds to `blocked` the directions whose cone holds the neighbor at (dx, dy, dz):
// the projection (d . u) / |d| >= cone cosine, for 0 < |d|^2 <= re^2
// (fluid_neighbor_blocks_light / wall_neighbor_blocks_light,
// surface_detection.cu:385-447).  One pass over a row tests all 14 directions,
// which gives the same set as the reference's one pass per direction.
__device__ __forceinline__ unsigned light_blocked(double dx, double dy, double dz, double re2,
                                                  double cone) {
  const double d2 = dx * dx + dy * dy + dz * dz;
  if (d2 <= 0.0 || d2 > re2) return 0u;
  const double inv = safe_inv_sqrt(d2);
  unsigned m = 0u;
#pragma unroll
  for (int k = 0; k < kLightDirections; ++k) {
    const double proj = (dx * kLight[k][0] + dy * kLight[k][1] + dz * kLight[k][2]) * inv;
    if (proj >= cone) m |= 1u << k;
  }
  return m;
}

// virtual_light_kernel (surface_detection.cu:449-475): one thread per fluid
// particle; the wall row is read only while some direction is still open.
__global__ void __launch_bounds__(kBlock) virtual_light_kernel(
    size_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ wx, const double* __restrict__ wy,
    const double* __restrict__ wz, const int32_t* __restrict__ f_off,
    const int32_t* __restrict__ f_idx, const int32_t* __restrict__ w_off,
    const int32_t* __restrict__ w_idx, double re2, double cone, int32_t include_walls,
    int32_t* __restrict__ open_count, double* __restrict__ open_fraction) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = x[i], py = y[i], pz = z[i];
  unsigned blocked = 0u;
  for (int32_t c = f_off[i], e = f_off[i + 1]; c < e && blocked != kAllBlocked; ++c) {
    const int32_t j = f_idx[c];
    blocked |= light_blocked(x[j] - px, y[j] - py, z[j] - pz, re2, cone);
  }
  if (include_walls) {
    for (int32_t c = w_off[i], e = w_off[i + 1]; c < e && blocked != kAllBlocked; ++c) {
      const int32_t j = w_idx[c];
      blocked |= light_blocked(wx[j] - px, wy[j] - py, wz[j] - pz, re2, cone);
    }
  }
  const int open = kLightDirections - __popc(blocked);
  if (open_count) open_count[i] = open;
  if (open_fraction)
    open_fraction[i] = static_cast<double>(open) / static_cast<double>(kLightDirections);
}

}  // namespace

// compute_uniform_reference_number_density (surface_detection.cu:494-520), host FP64.
double reference_number_density(double l0, double re, int kernel) {
  require(l0 > 0.0 && re > 0.0,
          "Reference number density requires positive spacing and support radius");
  const int32_t extent = static_cast<int32_t>(std::ceil(re / l0));
  double density = 0.0;
  for (int32_t iz = -extent; iz <= extent; ++iz)
    for (int32_t iy = -extent; iy <= extent; ++iy)
      for (int32_t ix = -extent; ix <= extent; ++ix) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        const double dx = static_cast<double>(ix) * l0;
        const double dy = static_cast<double>(iy) * l0;
        const double dz = static_cast<double>(iz) * l0;
        const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (d >= re) continue;
        const double t = 1.0 - d / re;
        density += kernel == LSMPS_WEIGHT_QUADRATIC ? t * t : t;
      }
  return density;
}

void classify_surface(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                      const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                      const lsmps_config& cfg, const lsmps_surface_diagnostics* diagnostics) {
  // validate_surface_inputs (surface_detection.cu:477-490) and classify preconditions (:528-538)
  require(cfg.support_radius > 0.0, "Surface detection support radius must be positive");
  require(fluid.count <= fn.particle_count && fluid.count <= wn.particle_count,
          "Surface detection neighbor lists are smaller than fluid particle count");
  require(cfg.near_surface_radius >= 0.0, "Near-surface radius must be non-negative");
  const double n0 = reference_number_density(cfg.particle_spacing, cfg.support_radius,
                                             cfg.weight_kernel);
  require(n0 > 0.0, "Reference number density must be positive");
  if (fluid.count == 0) return;
  require(fluid.surface_type != nullptr, "Surface detection requires a surface_type buffer");
  SurfaceParams p;
  p.re = cfg.support_radius;
  p.n0 = n0;
  p.ratio_threshold = cfg.number_density_ratio_threshold;
  p.open_threshold = cfg.air_open_ratio_threshold;
  p.aniso_threshold = cfg.air_anisotropy_threshold;
  p.independence = cfg.wall_normal_independence_threshold;
  p.splash_threshold = cfg.splash_neighbor_threshold;
  p.include_walls = cfg.include_wall_neighbors ? 1 : 0;
  p.kernel = cfg.weight_kernel;
  lsmps_surface_diagnostics diag{};
  if (diagnostics) diag = *diagnostics;
  ctx.tmp_primary.reserve(fluid.count);
  int32_t* primary = ctx.tmp_primary.data();
  classify_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.x, fluid.y, fluid.z, walls.x, walls.y, walls.z, walls.normal_x,
      walls.normal_y, walls.normal_z, fn.offsets, fn.indices, wn.offsets, wn.indices, p, primary,
      diag);
  LSMPS_LAUNCH_CHECK();
  const double r2 = cfg.near_surface_radius > 0.0
                        ? cfg.near_surface_radius * cfg.near_surface_radius
                        : 0.0;
  expand_kernel<<<blocks_for(fluid.count * kExpandLanes), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.x, fluid.y, fluid.z, fn.offsets, fn.indices, primary, r2,
      fluid.surface_type);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps

namespace lsmps {

// compute_virtual_light_diagnostics (surface_detection.cu:555-578): the same
// input checks and messages, then one launch.
void virtual_light(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                   const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                   const lsmps_config& cfg, const lsmps_virtual_light_diagnostics& out) {
  require(cfg.support_radius > 0.0, "Surface detection support radius must be positive");
  require(fluid.count <= fn.particle_count && fluid.count <= wn.particle_count,
          "Surface detection neighbor lists are smaller than fluid particle count");
  // written as the reference's test, so a NaN cosine passes it as it does there
  require(!(cfg.virtual_light_cone_cosine < -1.0 || cfg.virtual_light_cone_cosine > 1.0),
          "Virtual light cone cosine must be in [-1, 1]");
  require(out.open_direction_count != nullptr || out.open_fraction != nullptr,
          "Virtual light diagnostics require at least one output buffer");
  if (fluid.count == 0) return;
  const double re2 = cfg.support_radius * cfg.support_radius;
  virtual_light_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.x, fluid.y, fluid.z, walls.x, walls.y, walls.z, fn.offsets, fn.indices,
      wn.offsets, wn.indices, re2, cfg.virtual_light_cone_cosine,
      cfg.include_wall_neighbors ? 1 : 0, out.open_direction_count, out.open_fraction);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps
