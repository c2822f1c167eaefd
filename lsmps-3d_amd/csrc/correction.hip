// Pressure correction on gfx950.
//
// Replaces DevicePressureCorrection::apply (src/correction/pressure_correction.cu:598-645)
// and its free functions (:647-734).  Kernel order is the reference's:
// clamp -> pressure gradient -> particle shifting -> velocity/position update ->
// anti-penetration -> velocity smoothing (+ copy back).
#include <type_traits>

#include "common.hpp"
#include "device_math.hpp"
#include "stages.hpp"
#include "tune.hpp"

namespace lsmps {
namespace {

__device__ __forceinline__ bool is_splash(const int32_t* st, size_t i) {
  return st && st[i] == LSMPS_SURFACE_SPLASH;
}

// clamp_fluid_pressure_kernel (pressure_correction.cu:188-198)
__global__ void clamp_kernel(size_t n, double* p, double* fluid_p) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = p[i] < 0.0 ? 0.0 : p[i];
  p[i] = v;
  if (fluid_p && fluid_p != p) fluid_p[i] = v;
}

struct GradArgs {
  size_t n;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* st;
  const double* wx;
  const double* wy;
  const double* wz;
  const double* wnx;
  const double* wny;
  const double* wnz;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  const double* p;
  const double* ops;
  double re;
  double wall_scale;
  double density;
  double g[3];
  int32_t kernel;
  double* gx;
  double* gy;
  double* gz;
};

// compute_pressure_gradient_kernel (pressure_correction.cu:200-276)
// The neighbour-sum kernels give each particle L adjacent lanes that take its
// neighbours round-robin (contiguous index loads per wavefront, a particle's
// gathers in flight together); the sums follow the 16-slot tree of SlotSums, so
// every L gives the same result, and lane 0 finishes the particle.
constexpr int kCorrSlots = 16;

template <int L>
__global__ void __launch_bounds__(kBlock) gradient_kernel(GradArgs a) {
  using Sums = SlotSums<L, kCorrSlots>;
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / L;
  const int sub = static_cast<int>(t % L);
  if (i >= a.n) return;  // whole groups
  if (is_splash(a.st, i)) {
    if (sub != 0) return;
    a.gx[i] = 0.0;
    a.gy[i] = 0.0;
    a.gz[i] = 0.0;
    return;
  }
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i], pi = a.p[i];
  const double inv_re = 1.0 / a.re;
  double rhs[Sums::M][9];
#pragma unroll
  for (int m = 0; m < Sums::M; ++m)
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[m][k] = 0.0;
  Sums::for_each(a.f_off[i], a.f_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
    const double w = correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) return;
    double b[9];
    type_a_basis(dx, dy, dz, inv_re, b);
    const double dp = a.p[j] - pi;
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[m][k] += w * b[k] * dp;
  });
  Sums::for_each(a.w_off[i], a.w_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
    const double w =
        a.wall_scale * correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) return;
    const double nx = a.wnx[j], ny = a.wny[j], nz = a.wnz[j];
    double b[9];
    wall_neumann_q(dx, dy, dz, nx, ny, nz, inv_re, b);
    const double ng = a.g[0] * nx + a.g[1] * ny + a.g[2] * nz;
    const double wall_rhs = a.re * a.density * ng;
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[m][k] += w * b[k] * wall_rhs;
  });
  Sums::reduce(rhs);
  if (sub != 0) return;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    r0 += a.ops[static_cast<size_t>(k) * a.n + i] * rhs[0][k];
    r1 += a.ops[static_cast<size_t>(9 + k) * a.n + i] * rhs[0][k];
    r2 += a.ops[static_cast<size_t>(18 + k) * a.n + i] * rhs[0][k];
  }
  a.gx[i] = r0 / a.re;
  a.gy[i] = r1 / a.re;
  a.gz[i] = r2 / a.re;
}

struct ShiftArgs {
  size_t n;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* st;
  const double* wx;
  const double* wy;
  const double* wz;
  const double* wnx;
  const double* wny;
  const double* wnz;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  double scale;
  double min_distance;
  double max_shift;
  double* sx;
  double* sy;
  double* sz;
};

// compute_particle_shifting_kernel (pressure_correction.cu:278-367)
template <int L>
__global__ void __launch_bounds__(kBlock) shifting_kernel(ShiftArgs a) {
  using Sums = SlotSums<L, kCorrSlots>;
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / L;
  const int sub = static_cast<int>(t % L);
  if (i >= a.n) return;  // whole groups
  const int32_t st = a.st ? a.st[i] : LSMPS_SURFACE_INNER;
  if (a.scale <= 0.0 || st == LSMPS_SURFACE_SPLASH) {
    if (sub != 0) return;
    a.sx[i] = 0.0;
    a.sy[i] = 0.0;
    a.sz[i] = 0.0;
    return;
  }
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  // per slot: shift x, y, z, then the fluid normal sum x, y, z
  double acc[Sums::M][6];
#pragma unroll
  for (int m = 0; m < Sums::M; ++m)
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[m][k] = 0.0;
  Sums::for_each(a.f_off[i], a.f_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.f_idx[c];
    const double dx = xi - a.x[j], dy = yi - a.y[j], dz = zi - a.z[j];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d <= 0.0 || d >= a.min_distance) return;
    const double push = a.scale * (a.min_distance - d) / d;
    acc[m][0] += push * dx;
    acc[m][1] += push * dy;
    acc[m][2] += push * dz;
    acc[m][3] += dx / d;
    acc[m][4] += dy / d;
    acc[m][5] += dz / d;
  });
  Sums::for_each(a.w_off[i], a.w_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.w_idx[c];
    const double dx = xi - a.wx[j], dy = yi - a.wy[j], dz = zi - a.wz[j];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d <= 0.0 || d >= a.min_distance) return;
    const double push = a.scale * (a.min_distance - d);
    acc[m][0] += push * a.wnx[j];
    acc[m][1] += push * a.wny[j];
    acc[m][2] += push * a.wnz[j];
  });
  Sums::reduce(acc);
  if (sub != 0) return;
  double shx = acc[0][0], shy = acc[0][1], shz = acc[0][2];
  double nx = acc[0][3], ny = acc[0][4], nz = acc[0][5];
  if (st == LSMPS_SURFACE_SURFACE || st == LSMPS_SURFACE_SPLASH) {
    const double nn = sqrt(nx * nx + ny * ny + nz * nz);
    if (nn > 0.0) {
      nx /= nn;
      ny /= nn;
      nz /= nn;
      const double comp = shx * nx + shy * ny + shz * nz;
      shx -= comp * nx;
      shy -= comp * ny;
      shz -= comp * nz;
    }
  }
  const double norm = sqrt(shx * shx + shy * shy + shz * shz);
  if (norm > a.max_shift && norm > 0.0) {
    const double s = a.max_shift / norm;
    shx *= s;
    shy *= s;
    shz *= s;
  }
  a.sx[i] = shx;
  a.sy[i] = shy;
  a.sz[i] = shz;
}

// update_velocity_position_kernel (pressure_correction.cu:369-402)
__global__ void update_kernel(size_t n, double* x, double* y, double* z, double* vx, double* vy,
                              double* vz, const int32_t* st, const double* tvx,
                              const double* tvy, const double* tvz, const double* gx,
                              const double* gy, const double* gz, const double* px,
                              const double* py, const double* pz, double dt, double density) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ovx = vx[i], ovy = vy[i], ovz = vz[i];
  double nvx = tvx[i], nvy = tvy[i], nvz = tvz[i];
  if (!is_splash(st, i)) {
    const double s = dt / density;
    nvx -= s * gx[i];
    nvy -= s * gy[i];
    nvz -= s * gz[i];
  }
  x[i] += 0.5 * dt * (ovx + nvx) + px[i];
  y[i] += 0.5 * dt * (ovy + nvy) + py[i];
  z[i] += 0.5 * dt * (ovz + nvz) + pz[i];
  vx[i] = nvx;
  vy[i] = nvy;
  vz[i] = nvz;
}

// anti_penetration_kernel (pressure_correction.cu:404-444): sequential over the
// particle's wall neighbors, each correction sees the previous ones.
__global__ void antipen_kernel(size_t n, double* x, double* y, double* z, double* vx, double* vy,
                               double* vz, const double* wx, const double* wy, const double* wz,
                               const double* wnx, const double* wny, const double* wnz,
                               const int32_t* w_off, const int32_t* w_idx, double clearance) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double px = x[i], py = y[i], pz = z[i];
  double ux = vx[i], uy = vy[i], uz = vz[i];
  bool moved = false;
  for (int32_t c = w_off[i]; c < w_off[i + 1]; ++c) {
    const int32_t j = w_idx[c];
    const double nx = wnx[j], ny = wny[j], nz = wnz[j];
    const double sd = (px - wx[j]) * nx + (py - wy[j]) * ny + (pz - wz[j]) * nz;
    if (sd >= clearance) continue;
    const double corr = clearance - sd;
    px += corr * nx;
    py += corr * ny;
    pz += corr * nz;
    const double vn = ux * nx + uy * ny + uz * nz;
    if (vn < 0.0) {
      ux -= vn * nx;
      uy -= vn * ny;
      uz -= vn * nz;
    }
    moved = true;
  }
  if (moved) {
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    vx[i] = ux;
    vy[i] = uy;
    vz[i] = uz;
  }
}

// smooth_velocity_kernel (pressure_correction.cu:446-504)
template <int L>
__global__ void __launch_bounds__(kBlock) smooth_kernel(
    size_t n, const double* x, const double* y, const double* z, const double* vx,
    const double* vy, const double* vz, const int32_t* st, const int32_t* f_off,
    const int32_t* f_idx, double re, double strength, int32_t kernel, double* ox, double* oy,
    double* oz) {
  using Sums = SlotSums<L, kCorrSlots>;
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / L;
  const int sub = static_cast<int>(t % L);
  if (i >= n) return;  // whole groups
  if (strength <= 0.0 || is_splash(st, i)) {
    if (sub != 0) return;
    ox[i] = vx[i];
    oy[i] = vy[i];
    oz[i] = vz[i];
    return;
  }
  // per slot: weight sum, then the weighted velocity sums
  double acc[Sums::M][4];
#pragma unroll
  for (int m = 0; m < Sums::M; ++m) acc[m][0] = acc[m][1] = acc[m][2] = acc[m][3] = 0.0;
  const double xi = x[i], yi = y[i], zi = z[i];
  Sums::for_each(f_off[i], f_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = f_idx[c];
    if (is_splash(st, j)) return;
    const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
    const double w = correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), re, kernel);
    if (w <= 0.0) return;
    acc[m][0] += w;
    acc[m][1] += w * vx[j];
    acc[m][2] += w * vy[j];
    acc[m][3] += w * vz[j];
  });
  Sums::reduce(acc);
  if (sub != 0) return;
  const double ws = acc[0][0], ax = acc[0][1], ay = acc[0][2], az = acc[0][3];
  if (ws <= 0.0) {
    ox[i] = vx[i];
    oy[i] = vy[i];
    oz[i] = vz[i];
    return;
  }
  const double alpha = strength < 0.0 ? 0.0 : (strength > 1.0 ? 1.0 : strength);
  ox[i] = (1.0 - alpha) * vx[i] + alpha * ax / ws;
  oy[i] = (1.0 - alpha) * vy[i] + alpha * ay / ws;
  oz[i] = (1.0 - alpha) * vz[i] + alpha * az / ws;
}

// copy_smoothed_velocity_kernel (pressure_correction.cu:506-515)
__global__ void copy3_kernel(size_t n, const double* sx, const double* sy, const double* sz,
                             double* vx, double* vy, double* vz) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = sx[i];
  vy[i] = sy[i];
  vz[i] = sz[i];
}

void require_fluid_arrays(const lsmps_fluid_soa& f) {
  require(f.count == 0 || (f.x && f.y && f.z && f.vx && f.vy && f.vz),
          "Pressure correction requires fluid position and velocity arrays");
}
void require_vector(const lsmps_fluid_soa& f, const lsmps_vector3& v, const char* what) {
  if (v.count < f.count) raise(LSMPS_ERR_INVALID_ARGUMENT,
                               std::string(what) + " workspace is smaller than fluid particle count");
  if (f.count > 0 && !(v.x && v.y && v.z))
    raise(LSMPS_ERR_INVALID_ARGUMENT, std::string(what) + " workspace requires x/y/z arrays");
}

}  // namespace

void clamp_pressure(Context& ctx, const lsmps_fluid_soa& fluid, double* pressure) {
  require(fluid.count == 0 || pressure != nullptr, "Pressure correction requires a pressure array");
  if (fluid.count == 0) return;
  clamp_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(fluid.count, pressure,
                                                                   fluid.pressure);
  LSMPS_LAUNCH_CHECK();
}

void pressure_gradient(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const double* pressure,
                       const double* pre_ops, const lsmps_vector3& grad) {
  require_fluid_arrays(fluid);
  require(fluid.count == 0 || (fn.particle_count >= fluid.count && fn.offsets),
          "Pressure correction requires a valid fluid-neighbor CSR table");
  require(fluid.count == 0 || pressure != nullptr, "Pressure correction requires a pressure array");
  require_vector(fluid, grad, "Pressure gradient");
  if (fluid.count == 0) return;
  GradArgs a;
  a.n = fluid.count;
  a.x = fluid.x;
  a.y = fluid.y;
  a.z = fluid.z;
  a.st = fluid.surface_type;
  a.wx = walls.x;
  a.wy = walls.y;
  a.wz = walls.z;
  a.wnx = walls.normal_x;
  a.wny = walls.normal_y;
  a.wnz = walls.normal_z;
  a.f_off = fn.offsets;
  a.f_idx = fn.indices;
  a.w_off = wn.offsets;
  a.w_idx = wn.indices;
  a.p = pressure;
  a.ops = pre_ops;
  a.re = cfg.support_radius;
  a.wall_scale = cfg.lsmps_wall_weight_scale;
  a.density = cfg.density;
  for (int k = 0; k < 3; ++k) a.g[k] = cfg.gravity[k];
  a.kernel = cfg.weight_kernel;
  a.gx = grad.x;
  a.gy = grad.y;
  a.gz = grad.z;
  const size_t n = fluid.count;
  launch_tuned(ctx, kTuneGradient, {4, 8, 16}, 8, n, fn.offsets, [&](int lanes) {
    switch (lanes) {
      case 4: gradient_kernel<4><<<blocks_for(n * 4), kBlock, 0, ctx.stream>>>(a); break;
      case 8: gradient_kernel<8><<<blocks_for(n * 8), kBlock, 0, ctx.stream>>>(a); break;
      default: gradient_kernel<16><<<blocks_for(n * 16), kBlock, 0, ctx.stream>>>(a); break;
    }
  });
}

void particle_shifting(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const lsmps_vector3& ps) {
  require_fluid_arrays(fluid);
  require(fluid.count == 0 || (fn.particle_count >= fluid.count && fn.offsets),
          "Pressure correction requires a valid fluid-neighbor CSR table");
  require_vector(fluid, ps, "PS displacement");
  if (fluid.count == 0) return;
  ShiftArgs a;
  a.n = fluid.count;
  a.x = fluid.x;
  a.y = fluid.y;
  a.z = fluid.z;
  a.st = fluid.surface_type;
  a.wx = walls.x;
  a.wy = walls.y;
  a.wz = walls.z;
  a.wnx = walls.normal_x;
  a.wny = walls.normal_y;
  a.wnz = walls.normal_z;
  a.f_off = fn.offsets;
  a.f_idx = fn.indices;
  a.w_off = wn.offsets;
  a.w_idx = wn.indices;
  a.scale = cfg.ps_displacement_scale;
  a.min_distance = cfg.ps_min_distance_ratio * cfg.particle_spacing;
  a.max_shift = cfg.ps_max_displacement_ratio * cfg.particle_spacing;
  a.sx = ps.x;
  a.sy = ps.y;
  a.sz = ps.z;
  const size_t n = fluid.count;
  launch_tuned(ctx, kTuneShifting, {4, 8, 16}, 8, n, fn.offsets, [&](int lanes) {
    switch (lanes) {
      case 4: shifting_kernel<4><<<blocks_for(n * 4), kBlock, 0, ctx.stream>>>(a); break;
      case 8: shifting_kernel<8><<<blocks_for(n * 8), kBlock, 0, ctx.stream>>>(a); break;
      default: shifting_kernel<16><<<blocks_for(n * 16), kBlock, 0, ctx.stream>>>(a); break;
    }
  });
}

void update_velocity_position(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                              const lsmps_fluid_soa& tmp, const lsmps_vector3& grad,
                              const lsmps_vector3& ps) {
  require_fluid_arrays(fluid);
  require(tmp.count >= fluid.count, "Pressure correction temporary velocity target is too small");
  require(fluid.count == 0 || (tmp.vx && tmp.vy && tmp.vz),
          "Pressure correction requires temporary vx/vy/vz arrays");
  require_vector(fluid, grad, "Pressure gradient");
  require_vector(fluid, ps, "PS displacement");
  if (fluid.count == 0) return;
  update_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.x, fluid.y, fluid.z, fluid.vx, fluid.vy, fluid.vz, fluid.surface_type,
      tmp.vx, tmp.vy, tmp.vz, grad.x, grad.y, grad.z, ps.x, ps.y, ps.z, cfg.time_step,
      cfg.density);
  LSMPS_LAUNCH_CHECK();
}

void anti_penetration(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                      const lsmps_wall_soa& walls, const lsmps_neighbor_list& wn) {
  const double clearance = cfg.wall_clearance_ratio * cfg.particle_spacing;
  if (clearance <= 0.0 || fluid.count == 0) return;
  antipen_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.x, fluid.y, fluid.z, fluid.vx, fluid.vy, fluid.vz, walls.x, walls.y,
      walls.z, walls.normal_x, walls.normal_y, walls.normal_z, wn.offsets, wn.indices, clearance);
  LSMPS_LAUNCH_CHECK();
}

void smooth_velocity(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_neighbor_list& fn, const lsmps_vector3& smoothed) {
  require_fluid_arrays(fluid);
  require(fluid.count == 0 || (fn.particle_count >= fluid.count && fn.offsets),
          "Pressure correction requires a valid fluid-neighbor CSR table");
  require_vector(fluid, smoothed, "Velocity smoothing");
  if (fluid.count == 0) return;
  const size_t n = fluid.count;
  auto smooth = [&](auto lanes_tag) {
    constexpr int L = decltype(lanes_tag)::value;
    smooth_kernel<L><<<blocks_for(n * L), kBlock, 0, ctx.stream>>>(
        n, fluid.x, fluid.y, fluid.z, fluid.vx, fluid.vy, fluid.vz, fluid.surface_type,
        fn.offsets, fn.indices, cfg.support_radius, cfg.velocity_smoothing_strength,
        cfg.weight_kernel, smoothed.x, smoothed.y, smoothed.z);
  };
  launch_tuned(ctx, kTuneSmoothing, {4, 8, 16}, 8, n, fn.offsets, [&](int lanes) {
    switch (lanes) {
      case 4: smooth(std::integral_constant<int, 4>{}); break;
      case 8: smooth(std::integral_constant<int, 8>{}); break;
      default: smooth(std::integral_constant<int, 16>{}); break;
    }
  });
  copy3_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, smoothed.x, smoothed.y, smoothed.z, fluid.vx, fluid.vy, fluid.vz);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps
