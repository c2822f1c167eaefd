// Stage entry points of the per-step pipeline, shared by the C-ABI (capi.cpp)
// and the fused stepper (stepper.cpp).  All functions enqueue on ctx.stream and
// throw lsmps::Error on failure.
#pragma once

#include "common.hpp"
#include "device_math.hpp"

namespace lsmps {

// ---- neighbor.hip
// Scratch of the cell-list / neighbor-table build, as raw device pointers:
// per-cell counts and cursors (cells_capacity entries each, >= cells + 1),
// per-particle cell labels and 32-byte sorted point records
// (particles_capacity, >= the larger of the fluid and wall counts), and the
// partial sums of the device-wide scans.
constexpr size_t kSortedPointBytes = 32;
struct GeometryScratch {
  int32_t* counts = nullptr;
  int32_t* cursor = nullptr;
  size_t cells_capacity = 0;
  int32_t* labels = nullptr;
  char* points = nullptr;
  size_t particles_capacity = 0;
  int32_t* partials = nullptr;
  size_t partials_capacity = 0;
};
// Partial-sum entries the device-wide scan of n values needs.
size_t scan_partials_needed(size_t n);
void exclusive_scan_i32_with(Context& ctx, const int32_t* in, int32_t* out, size_t n,
                             int32_t* partials, size_t partials_capacity);
// The context's own scratch, grown to the given sizes (the only allocation of
// the geometry build).
GeometryScratch context_geometry_scratch(Context& ctx, size_t cells, size_t fluid_count,
                                         size_t wall_count);
lsmps_cell_grid neighbor_grid(const lsmps_config& config);
void build_cell_list(Context& ctx, size_t count, const double* x, const double* y,
                     const double* z, const lsmps_cell_grid& grid, const lsmps_cell_list& cells);
void build_neighbor_lists(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                          const lsmps_config& config, const lsmps_cell_list& fluid_cells,
                          const lsmps_cell_list& wall_cells,
                          const lsmps_neighbor_list& fluid_neighbors,
                          const lsmps_neighbor_list& wall_neighbors, int32_t* fluid_total,
                          int32_t* wall_total);
// The same with caller-provided scratch (no device allocation).
void build_neighbor_lists_with(Context& ctx, const GeometryScratch& scratch,
                               const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                               const lsmps_config& config, const lsmps_cell_list& fluid_cells,
                               const lsmps_cell_list& wall_cells,
                               const lsmps_neighbor_list& fluid_neighbors,
                               const lsmps_neighbor_list& wall_neighbors, int32_t* fluid_total,
                               int32_t* wall_total);

// ---- noalloc.hip
// Looks up this device's copy of the allocation-free build's static scratch
// (once per device; the device must be current).
void noalloc_prepare_device(int device);
// build_neighbor_lists with the static scratch: no device allocation, at most
// 2^20 particles per set and 2^20 cells (LSMPS_ERR_CAPACITY beyond).
void build_neighbor_lists_noalloc(Context& ctx, const lsmps_fluid_soa& fluid,
                                  const lsmps_wall_soa& walls, const lsmps_config& config,
                                  const lsmps_cell_list& fluid_cells,
                                  const lsmps_cell_list& wall_cells,
                                  const lsmps_neighbor_list& fluid_neighbors,
                                  const lsmps_neighbor_list& wall_neighbors);

// Batched independent tanks: tank k's cells occupy [base, base + nx ny nz) of a
// combined cell numbering; fluid_seg / wall_seg give each particle's tank.  All
// arrays are device arrays; neighbors never cross tanks.
struct SegmentGrid {
  double ox, oy, oz, h;
  int32_t nx, ny, nz, base;
};
struct Segments {
  int count = 0;
  const SegmentGrid* grids = nullptr;
  const int32_t* fluid_seg = nullptr;
  const int32_t* wall_seg = nullptr;
  size_t cells = 0;
};
void build_neighbor_lists_segmented(Context& ctx, const lsmps_fluid_soa& fluid,
                                    const lsmps_wall_soa& walls, const lsmps_config& config,
                                    const Segments& segments, const lsmps_cell_list& fluid_cells,
                                    const lsmps_cell_list& wall_cells,
                                    const lsmps_neighbor_list& fluid_neighbors,
                                    const lsmps_neighbor_list& wall_neighbors);

// ---- surface.hip
double reference_number_density(double l0, double re, int kernel);
void classify_surface(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                      const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                      const lsmps_config& config, const lsmps_surface_diagnostics* diagnostics);
void virtual_light(Context& ctx, const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                   const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn,
                   const lsmps_config& cfg, const lsmps_virtual_light_diagnostics& out);

// ---- moment.hip
struct MomentOutputs {
  double* ops = nullptr;             // kOpRows x N operator rows (Type-A kinds)
  double* inverse = nullptr;         // N x n x n column-major
  double* matrices = nullptr;        // N x n x n column-major
  double* trace = nullptr;
  double* regularization = nullptr;
  int32_t* info = nullptr;
};
int moment_size(int kind);
void compute_moments(Context& ctx, int kind, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                     const lsmps_neighbor_list& wn, const MomentOutputs& out);
// Reverse mode of the operator rows of a Type-A kind: ops_bar (kOpRows x N, the
// loss gradient of the rows) is propagated to the fluid positions.  Pair terms
// are added to pair_grad (3 x pair_stride, gradient w.r.t. x_j - x_i of each
// fluid CSR entry), wall terms to xbar (3 x N).
void moment_adjoint(Context& ctx, int kind, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                    const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                    const lsmps_neighbor_list& wn, const double* ops_bar, double* pair_grad,
                    size_t pair_stride, double* xbar);

// Operator rows (kOpRows x n, the consumers' layout) read from the explicit
// inverses of a Type-A MomentMatrixView.
void operator_rows_from_view(Context& ctx, const lsmps_moment_view& view, size_t n,
                             double* ops);

// ---- provision.hip
// velocity Laplacian (rows of the velocity Type-A operator) fused with the
// explicit update u* = u + dt (nu_eff lap + g); `lap` may be null.
void provision_fluid(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                     const lsmps_neighbor_list& wn, const double* vel_ops,
                     const lsmps_fluid_soa& tmp, const lsmps_vector3* lap);
void provision_combine(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_vector3& lap, const lsmps_fluid_soa& tmp);
void provision_walls(Context& ctx, const lsmps_config& cfg, const lsmps_wall_soa& walls,
                     const lsmps_wall_soa& tmp);

// ---- ppe.hip
void ppe_assemble(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                  const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                  const lsmps_neighbor_list& wn, const lsmps_fluid_soa& tmp,
                  const lsmps_wall_soa& tmpw, const double* vel_ops, const double* pre_ops,
                  const lsmps_ppe_workspace& ws);

// ---- solver.hip
class PpeSolver;  // owns Krylov workspaces; defined in solver.hpp
void sell_spmv(Context& ctx, const lsmps_csr& A, const double* x, double* y, bool fp32_values);
void ppe_solve(Context& ctx, PpeSolver& solver, const lsmps_csr& a, const double* b, double* x,
               const lsmps_solver_options& opt, lsmps_solve_report* report);

// ---- correction.hip
void clamp_pressure(Context& ctx, const lsmps_fluid_soa& fluid, double* pressure);
void pressure_gradient(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const double* pressure,
                       const double* pre_ops, const lsmps_vector3& grad);
void particle_shifting(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                       const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                       const lsmps_neighbor_list& wn, const lsmps_vector3& ps);
void update_velocity_position(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                              const lsmps_fluid_soa& tmp, const lsmps_vector3& grad,
                              const lsmps_vector3& ps);
void anti_penetration(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                      const lsmps_wall_soa& walls, const lsmps_neighbor_list& wn);
void smooth_velocity(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_neighbor_list& fn, const lsmps_vector3& smoothed);

}  // namespace lsmps
