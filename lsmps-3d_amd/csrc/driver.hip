// Multi-step driver: the simulation clock, the step-start speed reduction and
// the step loop.
//
// SimulationTimeManager (src/core/time_step.cpp:47-172) is restated as host
// code with the same arithmetic order: dt = clamp(min(dt_prev * growth,
// cfl * l0 / v_max), [dt_min, dt_max]), cut to the remaining time; outputs are
// due at the initial state and whenever the time passes the next multiple of
// the output interval.  max_velocity (src/app/simulation_driver.cu:122-143)
// becomes a device reduction: per-block maxima of |v| (non-finite magnitudes
// skipped, as the reference's host loop does) and one finishing block, so only
// one double crosses to the host per step.  lsmps_stepper_run follows the
// driver loop (simulation_driver.cu:340-391): speed, clock advance, dt into the
// config, one step, and on output steps a geometry refresh (the reference
// rebuilds the lists and labels at the end of every step, before writing) and
// the caller's output callback.
#include <algorithm>
#include <cmath>
#include <limits>

#include "objects.hpp"

namespace lsmps {
namespace {

constexpr int kMaxBlocks = 1024;

__device__ __forceinline__ double block_max(double v) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  return red[0];
}

__global__ void __launch_bounds__(kBlock) speed_max_kernel(size_t n, const double* vx,
                                                           const double* vy, const double* vz,
                                                           double* partial) {
  double m = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double s = sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    if (isfinite(s)) m = fmax(m, s);
  }
  const double b = block_max(m);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

__global__ void __launch_bounds__(kBlock) speed_finish_kernel(int nb, const double* partial,
                                                              double* out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, partial[i]);
  const double b = block_max(m);
  if (threadIdx.x == 0) out[0] = b;
}

}  // namespace

double max_speed(Context& ctx, const lsmps_fluid_soa& fluid, DeviceArray<double>& scratch) {
  if (fluid.count == 0) return 0.0;
  require(fluid.vx && fluid.vy && fluid.vz, "max velocity requires fluid vx/vy/vz arrays");
  const unsigned nb = std::min<unsigned>(blocks_for(fluid.count), kMaxBlocks);
  scratch.reserve(kMaxBlocks + 1);
  speed_max_kernel<<<nb, kBlock, 0, ctx.stream>>>(fluid.count, fluid.vx, fluid.vy, fluid.vz,
                                                   scratch.data());
  speed_finish_kernel<<<1, kBlock, 0, ctx.stream>>>(static_cast<int>(nb), scratch.data(),
                                                     scratch.data() + kMaxBlocks);
  LSMPS_LAUNCH_CHECK();
  LSMPS_HIP(hipMemcpyAsync(ctx.host_f64, scratch.data() + kMaxBlocks, sizeof(double),
                           hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  return ctx.host_f64[0];
}

// ---- clock --------------------------------------------------------------

lsmps_time_limits time_limits_from_config(const lsmps_config& c) {
  lsmps_time_limits l{};
  l.current_time_step = c.time_step;
  l.min_time_step = c.min_time_step;
  l.max_time_step = c.max_time_step;
  l.growth_factor = c.time_step_growth_factor;
  l.cfl = c.cfl;
  l.particle_spacing = c.particle_spacing;
  l.final_time = c.final_time;
  l.output_interval = c.output_interval;
  return l;
}

static void require_positive(double v, const char* name) {
  if (!(std::isfinite(v) && v > 0.0))
    raise(LSMPS_ERR_INVALID_ARGUMENT,
          std::string("TimeStepLimits field '") + name + "' must be positive and finite");
}

void validate_time_limits(const lsmps_time_limits& l) {
  require_positive(l.current_time_step, "current_time_step");
  require_positive(l.min_time_step, "min_time_step");
  require_positive(l.max_time_step, "max_time_step");
  require_positive(l.growth_factor, "growth_factor");
  require_positive(l.cfl, "cfl");
  require_positive(l.particle_spacing, "particle_spacing");
  if (!(std::isfinite(l.final_time) && l.final_time >= 0.0))
    raise(LSMPS_ERR_INVALID_ARGUMENT,
          "TimeStepLimits field 'final_time' must be non-negative and finite");
  require_positive(l.output_interval, "output_interval");
  if (l.growth_factor < 1.0)
    raise(LSMPS_ERR_INVALID_ARGUMENT, "TimeStepLimits field 'growth_factor' must be >= 1");
  if (l.min_time_step > l.max_time_step)
    raise(LSMPS_ERR_INVALID_ARGUMENT, "TimeStepLimits requires min_time_step <= max_time_step");
}

}  // namespace lsmps

using namespace lsmps;

void lsmps_time_manager::reset(const lsmps_time_limits& l) {
  validate_time_limits(l);
  limits = l;
  current_time = 0.0;
  current_time_step = clamp_dt(l.current_time_step);
  next_output_time = 0.0;
  step_index = 0;
  output_index = 0;
  initial_output_pending = true;
}

double lsmps_time_manager::clamp_dt(double v) const {
  return std::clamp(v, limits.min_time_step, limits.max_time_step);
}

double lsmps_time_manager::cfl_dt(double vmax) const {
  if (!(vmax > 0.0) || !std::isfinite(vmax)) return limits.max_time_step;
  return limits.cfl * limits.particle_spacing / vmax;
}

double lsmps_time_manager::next_dt(double vmax) const {
  const double growth = current_time_step * limits.growth_factor;
  double dt = clamp_dt(std::min(growth, cfl_dt(vmax)));
  const double remaining = limits.final_time - current_time;
  if (remaining > 0.0 && dt > remaining) dt = remaining;
  return dt;
}

lsmps_time_status lsmps_time_manager::advance(double vmax) {
  const double previous = current_time;
  const double cdt = cfl_dt(vmax);
  const double dt = next_dt(vmax);
  current_time += dt;
  current_time_step = dt;
  ++step_index;
  bool out = false;
  if (initial_output_pending || current_time >= next_output_time) {
    out = true;
    initial_output_pending = false;
    ++output_index;
    while (next_output_time <= current_time) next_output_time += limits.output_interval;
  }
  return lsmps_time_status{step_index, output_index, previous, current_time, dt, cdt, vmax,
                           out ? 1 : 0, finished() ? 1 : 0};
}

lsmps_time_status lsmps_time_manager::mark_initial_output() {
  bool out = false;
  if (initial_output_pending) {
    out = true;
    initial_output_pending = false;
    ++output_index;
    next_output_time = limits.output_interval;
  }
  return lsmps_time_status{step_index, output_index, current_time, current_time,
                           current_time_step, std::numeric_limits<double>::infinity(), 0.0,
                           out ? 1 : 0, finished() ? 1 : 0};
}

// ---- step loop ----------------------------------------------------------

void lsmps_stepper::refresh_geometry(const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls) {
  require(geometry_ready && geometry_fluid == fluid.count && geometry_walls == walls.count,
          "geometry refresh needs a previous step over the same particles");
  Context& c = *ctx;
  const size_t cells = segment_count > 0
                           ? segment_cells
                           : static_cast<size_t>(config.cell_dims[0]) * config.cell_dims[1] *
                                 config.cell_dims[2];
  lsmps_cell_list fc{fluid_capacity, cells, fc_ids.data(), fc_sorted.data(), fc_begin.data(),
                     fc_end.data()};
  lsmps_cell_list wc{wall_capacity, cells, wc_ids.data(), wc_sorted.data(), wc_begin.data(),
                     wc_end.data()};
  const lsmps_neighbor_list fn = fluid_neighbors();
  const lsmps_neighbor_list wn = wall_neighbors();
  if (segment_count > 0) {
    Segments seg;
    seg.count = segment_count;
    seg.grids = segment_grids.data();
    seg.fluid_seg = fluid_segment.data();
    seg.wall_seg = wall_segment.data();
    seg.cells = segment_cells;
    build_neighbor_lists_segmented(c, fluid, walls, config, seg, fc, wc, fn, wn);
  } else {
    build_neighbor_lists(c, fluid, walls, config, fc, wc, fn, wn, nullptr, nullptr);
  }
  ++generation;
  classify_surface(c, fluid, walls, fn, wn, config, nullptr);
  c.sync();
}

namespace lsmps {

int run_steps(lsmps_stepper& S, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
              lsmps_time_manager& clock, int64_t max_steps, lsmps_output_callback cb, void* user,
              lsmps_time_status* last) {
  Context& c = *S.ctx;
  int rc = LSMPS_OK;
  if (clock.initial_output_pending) {
    const lsmps_time_status st = clock.mark_initial_output();
    if (last) *last = st;
    if (st.should_output && cb && cb(&st, user) != 0) return LSMPS_OK;
  }
  for (int64_t k = 0; !clock.finished() && (max_steps < 0 || k < max_steps); ++k) {
    const double vmax = max_speed(c, fluid, S.speed_scratch);
    const lsmps_time_status st = clock.advance(vmax);
    S.config.time_step = st.time_step;
    try {
      S.step(fluid, walls);
    } catch (const Error& e) {
      if (e.code != LSMPS_ERR_NOT_CONVERGED) throw;
      rc = LSMPS_ERR_NOT_CONVERGED;  // the reference continues with the solver's pressure
    }
    if (last) *last = st;
    if (st.should_output || st.reached_final_time) {
      S.refresh_geometry(fluid, walls);
      if (cb && cb(&st, user) != 0) break;
    }
  }
  return rc;
}

}  // namespace lsmps
