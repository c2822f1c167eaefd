// Opaque objects behind the C-ABI: the moment-matrix cache and the fused stepper.
#pragma once

#include <vector>

#include "common.hpp"
#include "solver.hpp"
#include "stages.hpp"

// DeviceMomentMatrix (include/lsmps3d/moment_matrix/moment_matrix.cuh:36-68).
// The step path keeps only the operator rows of the two consumed kinds; full
// inverses of any kind are materialised on request for MomentMatrixView parity
// from the inputs of the last prepare (which must not have changed since).
struct lsmps_moment_matrix {
  lsmps::Context* ctx = nullptr;
  lsmps_context* owner = nullptr;  // holds a reference (standalone objects only)
  lsmps_config config{};
  size_t capacity = 0;
  size_t count = 0;
  uint64_t generation = 0;
  bool prepared = false;
  lsmps::DeviceArray<double> vel_ops;  // kOpRows x capacity
  lsmps::DeviceArray<double> pre_ops;
  struct Full {
    lsmps::DeviceArray<double> inverse;
    lsmps::DeviceArray<double> trace;
    lsmps::DeviceArray<double> regularization;
    lsmps::DeviceArray<int32_t> info;
    uint64_t generation = 0;
    bool ready = false;
  } full[4];
  lsmps::DeviceArray<int32_t> inversion_count;  // 4 counters, one per kind
  int32_t host_counts[4] = {0, 0, 0, 0};
  lsmps_fluid_soa fluid{};
  lsmps_wall_soa walls{};
  lsmps_neighbor_list fn{};
  lsmps_neighbor_list wn{};

  size_t bytes() const {
    size_t b = vel_ops.bytes() + pre_ops.bytes() + inversion_count.bytes();
    for (const auto& f : full)
      b += f.inverse.bytes() + f.trace.bytes() + f.regularization.bytes() + f.info.bytes();
    return b;
  }
  void resize(size_t cap) {
    if (cap != capacity) {
      vel_ops.release();
      pre_ops.release();
      for (auto& f : full) {
        f.inverse.release();
        f.trace.release();
        f.regularization.release();
        f.info.release();
        f.ready = false;
      }
      capacity = cap;
      prepared = false;
    }
  }
  void prepare(const lsmps_fluid_soa& fluid_in, const lsmps_wall_soa& walls_in,
               const lsmps_neighbor_list& fn_in, const lsmps_neighbor_list& wn_in,
               uint64_t gen);
  void materialize(int kind);
  const double* velocity_ops() const;
  const double* pressure_ops() const;
};

// Owns every intermediate buffer of one time step.
struct lsmps_stepper {
  lsmps::Context* ctx = nullptr;
  lsmps_context* owner = nullptr;  // holds a reference
  lsmps_config config{};
  lsmps_solver_options solver_options{};
  size_t fluid_capacity = 0;
  size_t wall_capacity = 0;
  size_t max_neighbors = 0;
  size_t cell_capacity = 0;
  uint64_t generation = 0;
  lsmps::DeviceArray<int32_t> fc_ids, fc_sorted, fc_begin, fc_end;
  lsmps::DeviceArray<int32_t> wc_ids, wc_sorted, wc_begin, wc_end;
  lsmps::DeviceArray<int32_t> f_off, f_idx, w_off, w_idx;
  lsmps_moment_matrix moments;
  lsmps::DeviceArray<double> tmp_fluid;  // 3 x fluid capacity
  lsmps::DeviceArray<double> tmp_wall;   // 3 x wall capacity
  lsmps::DeviceArray<int32_t> ppe_rows, ppe_cols;
  lsmps::DeviceArray<double> ppe_vals, ppe_vec;  // ppe_vec: rhs, pressure, divergence, wall term
  lsmps::PpeSolver solver;
  lsmps::DeviceArray<double> corr;  // grad, ps, smoothed: 9 x fluid capacity
  // batched tanks (segment_count > 0): per-tank grids and particle tank ids
  int segment_count = 0;
  size_t segment_cells = 0;
  std::vector<int64_t> fluid_offsets, wall_offsets;
  lsmps::DeviceArray<lsmps::SegmentGrid> segment_grids;
  lsmps::DeviceArray<int32_t> fluid_segment, wall_segment;
  lsmps::DeviceArray<int32_t> segment_rows;  // fluid_offsets as int32 (the PPE's segments)
  std::vector<lsmps::CoarseTankGrid> coarse_grids;  // per-tank grids of the coarse space
  hipEvent_t events[8] = {};
  int n_events = 0;
  double stage_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  lsmps_solve_report last_report{};
  // LSMPS_STEP_* mode.  RECORD keeps what the adjoint of the step reads and the
  // step overwrites: positions and velocities at the start of the step and the
  // PPE solution before the clamp.  FROZEN_GEOMETRY reuses the previous step's
  // neighbor lists and surface labels.
  uint32_t mode = 0;
  bool geometry_ready = false;
  size_t geometry_fluid = 0, geometry_walls = 0;
  bool recorded = false;
  size_t recorded_fluid = 0, recorded_walls = 0;
  lsmps::DeviceArray<double> start_state;   // x, y, z, vx, vy, vz: 6 x fluid capacity
  lsmps::DeviceArray<double> raw_pressure;  // fluid capacity
  struct Adjoint {
    lsmps::DeviceArray<int32_t> rev, flags;
    lsmps::DeviceArray<double> pair, work, at_vals, sums;
    lsmps::PpeSolver solver;
    size_t bytes() const {
      return rev.bytes() + flags.bytes() + pair.bytes() + work.bytes() + at_vals.bytes() +
             sums.bytes() + solver.bytes();
    }
  } adjoint;
  lsmps::DeviceArray<double> speed_scratch;  // max-velocity block maxima

  void refresh_geometry(const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls);

  ~lsmps_stepper();
  void allocate(size_t nf, size_t nw, size_t cells);
  void set_segments(int count, const int64_t* fluid_offsets, const int64_t* wall_offsets,
                    const lsmps_cell_grid* grids);
  size_t bytes() const;
  void step(const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls);
  lsmps_neighbor_list fluid_neighbors() const;
  lsmps_neighbor_list wall_neighbors() const;
  lsmps_ppe_workspace ppe_workspace(size_t n) const;
};

// SimulationTimeManager (src/core/time_step.cpp:79-170); driver.hip.
struct lsmps_time_manager {
  lsmps_time_limits limits{};
  double current_time = 0.0;
  double current_time_step = 0.0;
  double next_output_time = 0.0;
  size_t step_index = 0;
  size_t output_index = 0;
  bool initial_output_pending = true;

  void reset(const lsmps_time_limits& l);
  double clamp_dt(double v) const;
  double cfl_dt(double vmax) const;
  double next_dt(double vmax) const;
  bool finished() const { return current_time >= limits.final_time; }
  bool output_due() const { return initial_output_pending || current_time >= next_output_time; }
  lsmps_time_status advance(double vmax);
  lsmps_time_status mark_initial_output();
};

namespace lsmps {
// io.cpp
void default_config_files(lsmps_config_files& f);
void validate_config(const lsmps_config& c, const lsmps_config_files* f);
void load_config(const char* path, lsmps_config& c, lsmps_config_files& f);
void save_config(const char* path, const lsmps_config& c, const lsmps_config_files& f);
size_t read_particle_csv(const char* path, int columns, double* const* out, size_t capacity);
std::string vtk_path(const char* directory, const char* prefix, size_t step);
void write_vtk(const char* directory, const char* prefix, bool point_fields, size_t step,
               size_t count, const double* x, const double* y, const double* z,
               const lsmps_vtk_field* fields, int nfields);
// driver.hip
lsmps_time_limits time_limits_from_config(const lsmps_config& c);
void validate_time_limits(const lsmps_time_limits& l);
double max_speed(Context& ctx, const lsmps_fluid_soa& fluid, DeviceArray<double>& scratch);
int run_steps(lsmps_stepper& stepper, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
              lsmps_time_manager& clock, int64_t max_steps, lsmps_output_callback cb, void* user,
              lsmps_time_status* last);
// Per-particle rows of adjoint work arrays (adjoint.hip).
constexpr int kAdjointWork = 142;
// Reverse mode of the stepper's last (recorded) step; adjoint.hip.
void adjoint_step(Context& ctx, lsmps_stepper& stepper, const lsmps_fluid_soa& fluid,
                  const lsmps_wall_soa& walls, const lsmps_vector3& grad_velocity_out,
                  const lsmps_vector3* grad_position_out, lsmps_adjoint_grads& grads);
}  // namespace lsmps
