// Explicit predictor on gfx950.
//
// Replaces DeviceProvisionExplicitUpdate::compute_temporary_velocity
// (src/provision/explicit_update.cu:366-399): compute_velocity_laplacian_kernel
// (:178-248) and combine_explicit_velocity_kernel (:250-270) are fused into one
// pass per particle, and combine_explicit_wall_velocity_kernel (:272-284) keeps
// its own launch over the walls.  The Laplacian uses the velocity Type-A
// operator row 3+4+5 (Laplacian coefficients), so 9 values per particle are
// read instead of the 81-entry inverse.
#include "common.hpp"
#include "device_math.hpp"
#include "stages.hpp"
#include "tune.hpp"

namespace lsmps {
namespace {

struct ProvisionArgs {
  size_t n;
  const double* x;
  const double* y;
  const double* z;
  const double* vx;
  const double* vy;
  const double* vz;
  const int32_t* surface_type;
  const double* wx;
  const double* wy;
  const double* wz;
  const double* wvx;
  const double* wvy;
  const double* wvz;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  const double* ops;  // velocity operator rows, SoA
  double re;
  double wall_scale;
  double nu;
  double dt;
  double g[3];
  int32_t kernel;
  double* out_vx;
  double* out_vy;
  double* out_vz;
  double* lap_x;
  double* lap_y;
  double* lap_z;
};

// L adjacent lanes per particle take its neighbours round-robin (contiguous
// index loads per wavefront, a particle's gathers in flight together); the sums
// follow the 16-slot tree of SlotSums, so every L gives the same result.
constexpr int kProvSlots = 16;

template <int L>
__global__ void __launch_bounds__(kBlock) provision_kernel(ProvisionArgs a) {
  using Sums = SlotSums<L, kProvSlots>;
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / L;
  const int sub = static_cast<int>(t % L);
  if (i >= a.n) return;  // whole groups
  double lap_row[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) lap_row[k] = a.ops[static_cast<size_t>(kOpLap + k) * a.n + i];
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const double vxi = a.vx[i], vyi = a.vy[i], vzi = a.vz[i];
  const double inv_re = 1.0 / a.re;
  // accumulate_velocity_rhs (explicit_update.cu:148-162) contracted with the
  // Laplacian row on the fly: lap = 2 (L . rhs) / re^2 and L . rhs = sum_j w (L . b_j) dv_j.
  double acc[Sums::M][3];
#pragma unroll
  for (int m = 0; m < Sums::M; ++m) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
  struct Nb {
    double x, y, z, vx, vy, vz;
  };
  Sums::template for_each_batched<kNbBatch>(
      a.f_off[i], a.f_off[i + 1], sub,
      [&](int32_t c) {
        const int32_t j = a.f_idx[c];
        return Nb{a.x[j], a.y[j], a.z[j], a.vx[j], a.vy[j], a.vz[j]};
      },
      [&](int32_t, int m, const Nb& q) {
        const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
        const double w = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
        if (w <= 0.0) return;
        double b[9];
        type_a_basis(dx, dy, dz, inv_re, b);
        double lb = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) lb += lap_row[k] * b[k];
        const double wl = w * lb;
        acc[m][0] += wl * (q.vx - vxi);
        acc[m][1] += wl * (q.vy - vyi);
        acc[m][2] += wl * (q.vz - vzi);
      });
  Sums::for_each(a.w_off[i], a.w_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
    const double w = a.wall_scale * weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) return;
    double b[9];
    type_a_basis(dx, dy, dz, inv_re, b);
    double lb = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) lb += lap_row[k] * b[k];
    const double wl = w * lb;
    acc[m][0] += wl * (a.wvx[j] - vxi);
    acc[m][1] += wl * (a.wvy[j] - vyi);
    acc[m][2] += wl * (a.wvz[j] - vzi);
  });
  Sums::reduce(acc);
  if (sub != 0) return;
  const double sx = acc[0][0], sy = acc[0][1], sz = acc[0][2];
  // laplacian_from_inverse_rhs (explicit_update.cu:164-176)
  const double scale = 2.0 / (a.re * a.re);
  const double lx = scale * sx, ly = scale * sy, lz = scale * sz;
  if (a.lap_x) {
    a.lap_x[i] = lx;
    a.lap_y[i] = ly;
    a.lap_z[i] = lz;
  }
  // combine_explicit_velocity_kernel (explicit_update.cu:250-270)
  const bool splash = a.surface_type && a.surface_type[i] == LSMPS_SURFACE_SPLASH;
  const double nu = splash ? 0.0 : a.nu;
  a.out_vx[i] = vxi + a.dt * (nu * lx + a.g[0]);
  a.out_vy[i] = vyi + a.dt * (nu * ly + a.g[1]);
  a.out_vz[i] = vzi + a.dt * (nu * lz + a.g[2]);
}

__global__ void combine_kernel(size_t n, const double* vx, const double* vy, const double* vz,
                               const int32_t* surface_type, const double* lx, const double* ly,
                               const double* lz, double nu_base, double dt, double gx, double gy,
                               double gz, double* ox, double* oy, double* oz) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool splash = surface_type && surface_type[i] == LSMPS_SURFACE_SPLASH;
  const double nu = splash ? 0.0 : nu_base;
  ox[i] = vx[i] + dt * (nu * lx[i] + gx);
  oy[i] = vy[i] + dt * (nu * ly[i] + gy);
  oz[i] = vz[i] + dt * (nu * lz[i] + gz);
}

__global__ void wall_kernel(size_t n, const double* vx, const double* vy, const double* vz,
                            double dt, double gx, double gy, double gz, double* ox, double* oy,
                            double* oz) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ox[i] = vx[i] + dt * gx;
  oy[i] = vy[i] + dt * gy;
  oz[i] = vz[i] + dt * gz;
}

}  // namespace

void provision_fluid(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                     const lsmps_neighbor_list& wn, const double* vel_ops,
                     const lsmps_fluid_soa& tmp, const lsmps_vector3* lap) {
  if (fluid.count == 0) return;
  require(fluid.vx && fluid.vy && fluid.vz, "Provision update requires fluid vx/vy/vz arrays");
  require(tmp.count >= fluid.count,
          "Provision temporary velocity target is smaller than fluid particle count");
  require(tmp.vx && tmp.vy && tmp.vz, "Provision temporary velocity target requires vx/vy/vz arrays");
  require(walls.count == 0 || (walls.vx && walls.vy && walls.vz),
          "Provision update requires wall vx/vy/vz arrays");
  ProvisionArgs a;
  a.n = fluid.count;
  a.x = fluid.x;
  a.y = fluid.y;
  a.z = fluid.z;
  a.vx = fluid.vx;
  a.vy = fluid.vy;
  a.vz = fluid.vz;
  a.surface_type = fluid.surface_type;
  a.wx = walls.x;
  a.wy = walls.y;
  a.wz = walls.z;
  a.wvx = walls.vx;
  a.wvy = walls.vy;
  a.wvz = walls.vz;
  a.f_off = fn.offsets;
  a.f_idx = fn.indices;
  a.w_off = wn.offsets;
  a.w_idx = wn.indices;
  a.ops = vel_ops;
  a.re = cfg.support_radius;
  a.wall_scale = cfg.lsmps_wall_weight_scale;
  a.nu = cfg.kinematic_viscosity;
  a.dt = cfg.time_step;
  for (int k = 0; k < 3; ++k) a.g[k] = cfg.gravity[k];
  a.kernel = cfg.weight_kernel;
  a.out_vx = tmp.vx;
  a.out_vy = tmp.vy;
  a.out_vz = tmp.vz;
  a.lap_x = lap ? lap->x : nullptr;
  a.lap_y = lap ? lap->y : nullptr;
  a.lap_z = lap ? lap->z : nullptr;
  const size_t n = fluid.count;
  launch_tuned(ctx, kTuneProvision, {4, 8, 16}, 8, n, fn.offsets, [&](int lanes) {
    switch (lanes) {
      case 4: provision_kernel<4><<<blocks_for(n * 4), kBlock, 0, ctx.stream>>>(a); break;
      case 8: provision_kernel<8><<<blocks_for(n * 8), kBlock, 0, ctx.stream>>>(a); break;
      default: provision_kernel<16><<<blocks_for(n * 16), kBlock, 0, ctx.stream>>>(a); break;
    }
  });
}

// compute_provision_temporary_velocity (explicit_update.cu:401-415)
void provision_combine(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_vector3& lap, const lsmps_fluid_soa& tmp) {
  require(fluid.count == 0 || (fluid.vx && fluid.vy && fluid.vz),
          "Provision update requires fluid vx/vy/vz arrays");
  require(lap.count >= fluid.count, "Provision laplacian workspace is smaller than fluid particle count");
  require(fluid.count == 0 || (lap.x && lap.y && lap.z),
          "Provision laplacian workspace requires vx/vy/vz arrays");
  require(tmp.count >= fluid.count,
          "Provision temporary velocity target is smaller than fluid particle count");
  require(fluid.count == 0 || (tmp.vx && tmp.vy && tmp.vz),
          "Provision temporary velocity target requires vx/vy/vz arrays");
  if (fluid.count == 0) return;
  combine_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(
      fluid.count, fluid.vx, fluid.vy, fluid.vz, fluid.surface_type, lap.x, lap.y, lap.z,
      cfg.kinematic_viscosity, cfg.time_step, cfg.gravity[0], cfg.gravity[1], cfg.gravity[2],
      tmp.vx, tmp.vy, tmp.vz);
  LSMPS_LAUNCH_CHECK();
}

// compute_provision_temporary_wall_velocity (explicit_update.cu:417-429)
void provision_walls(Context& ctx, const lsmps_config& cfg, const lsmps_wall_soa& walls,
                     const lsmps_wall_soa& tmp) {
  require(walls.count == 0 || (walls.vx && walls.vy && walls.vz),
          "Provision update requires wall vx/vy/vz arrays");
  require(tmp.count >= walls.count,
          "Provision temporary wall velocity target is smaller than wall particle count");
  require(walls.count == 0 || (tmp.vx && tmp.vy && tmp.vz),
          "Provision temporary wall velocity target requires vx/vy/vz arrays");
  if (walls.count == 0) return;
  wall_kernel<<<blocks_for(walls.count), kBlock, 0, ctx.stream>>>(
      walls.count, walls.vx, walls.vy, walls.vz, cfg.time_step, cfg.gravity[0], cfg.gravity[1],
      cfg.gravity[2], tmp.vx, tmp.vy, tmp.vz);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps
