// One time step through the C++ module API of include/lsmps3d_amd.hpp, in the
// stage order of the reference driver (src/app/simulation_driver.cu:336-393):
// cell lists -> neighbor CSR -> surface labels -> moment matrices -> provision ->
// PPE assembly -> PPE solve -> pressure correction.  It then replays the
// correction through the free functions (clamp, gradient, shifting, update,
// anti-penetration, smoothing) on a second copy of the state, as
// DevicePressureCorrection::apply sequences them (pressure_correction.cu:598-645).
// The reference-signature free functions that read MomentMatrixViews
// (assemble_ppe_matrix_and_rhs, compute_pressure_gradient) run beside the
// module classes, and all four moment views are written out.
//
// Usage: module_step <dir>.  <dir>/sizes.txt holds "n_fluid n_wall
// max_neighbors precision"; <dir>/config.bin the raw lsmps_config; <dir>/<name>.bin
// the input arrays (fluid_x ... fluid_surface_type, wall_x ... wall_normal_z).
// Outputs are written to <dir>/out_<name>.bin.  Exit status 0 on success.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "lsmps3d_amd.hpp"

using namespace lsmps3d;

namespace {

std::string g_dir;

template <typename T>
std::vector<T> read_array(const std::string& name, size_t count) {
  std::vector<T> v(count);
  std::ifstream f(g_dir + "/" + name + ".bin", std::ios::binary);
  if (!f) throw std::runtime_error("missing input " + name);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(count * sizeof(T)));
  if (static_cast<size_t>(f.gcount()) != count * sizeof(T)) throw std::runtime_error("short input " + name);
  return v;
}

template <typename T>
T* to_device(const std::vector<T>& h) {
  T* d = nullptr;
  detail::hip_check(hipMalloc(reinterpret_cast<void**>(&d), std::max<size_t>(h.size(), 1) * sizeof(T)), "hipMalloc");
  if (!h.empty()) detail::hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

template <typename T>
T* device_zeros(size_t count) {
  T* d = nullptr;
  detail::hip_check(hipMalloc(reinterpret_cast<void**>(&d), std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
  detail::hip_check(hipMemset(d, 0, std::max<size_t>(count, 1) * sizeof(T)), "hipMemset");
  return d;
}

template <typename T>
void write_device(const std::string& name, const T* d, size_t count) {
  std::vector<T> h(count);
  if (count) detail::hip_check(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  std::ofstream f(g_dir + "/out_" + name + ".bin", std::ios::binary);
  f.write(reinterpret_cast<const char*>(h.data()), static_cast<std::streamsize>(count * sizeof(T)));
}

SimulationConfig from_c(const lsmps_config& c) {
  SimulationConfig s;
  s.particle_spacing = c.particle_spacing;
  s.support_radius = c.support_radius;
  s.near_surface_radius = c.near_surface_radius;
  s.cell_origin = Vec3{c.cell_origin[0], c.cell_origin[1], c.cell_origin[2]};
  s.cell_size = c.cell_size;
  s.cell_dims = Int3{c.cell_dims[0], c.cell_dims[1], c.cell_dims[2]};
  s.time_step = c.time_step;
  s.density = c.density;
  s.kinematic_viscosity = c.kinematic_viscosity;
  s.gravity = Vec3{c.gravity[0], c.gravity[1], c.gravity[2]};
  s.splash_neighbor_threshold = c.splash_neighbor_threshold;
  s.include_wall_neighbors = c.include_wall_neighbors != 0;
  s.number_density_ratio_threshold = c.number_density_ratio_threshold;
  s.air_open_ratio_threshold = c.air_open_ratio_threshold;
  s.air_anisotropy_threshold = c.air_anisotropy_threshold;
  s.wall_normal_independence_threshold = c.wall_normal_independence_threshold;
  s.virtual_light_cone_cosine = c.virtual_light_cone_cosine;
  s.lsmps_regularization = c.lsmps_regularization;
  s.lsmps_wall_weight_scale = c.lsmps_wall_weight_scale;
  s.ps_displacement_scale = c.ps_displacement_scale;
  s.ps_min_distance_ratio = c.ps_min_distance_ratio;
  s.ps_max_displacement_ratio = c.ps_max_displacement_ratio;
  s.wall_clearance_ratio = c.wall_clearance_ratio;
  s.velocity_smoothing_strength = c.velocity_smoothing_strength;
  s.weight_kernel = c.weight_kernel;
  return s;
}

struct FluidBuffers {
  std::vector<double*> f;  // x y z vx vy vz pressure
  int32_t* surface_type = nullptr;
  FluidParticleSoA soa(size_t n) const {
    return FluidParticleSoA{n, f[0], f[1], f[2], f[3], f[4], f[5], f[6], surface_type};
  }
};

const char* kFluidNames[] = {"x", "y", "z", "vx", "vy", "vz", "pressure"};
const char* kWallNames[] = {"x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z"};

FluidBuffers upload_fluid(size_t n) {
  FluidBuffers b;
  for (const char* k : kFluidNames) b.f.push_back(to_device(read_array<double>(std::string("fluid_") + k, n)));
  b.surface_type = to_device(read_array<int32_t>("fluid_surface_type", n));
  return b;
}

int run() {
  size_t n = 0, m = 0, max_neighbors = 0;
  int precision = 0;
  {
    std::ifstream f(g_dir + "/sizes.txt");
    f >> n >> m >> max_neighbors >> precision;
    if (!f) throw std::runtime_error("bad sizes.txt");
  }
  lsmps_config craw{};
  {
    std::ifstream f(g_dir + "/config.bin", std::ios::binary);
    f.read(reinterpret_cast<char*>(&craw), sizeof(craw));
    if (static_cast<size_t>(f.gcount()) != sizeof(craw)) throw std::runtime_error("bad config.bin");
  }
  set_precision(precision);
  const SimulationConfig config = from_c(craw);

  FluidBuffers fb = upload_fluid(n);
  FluidBuffers fb2 = upload_fluid(n);  // second copy for the free-function replay
  std::vector<double*> w;
  for (const char* k : kWallNames) w.push_back(to_device(read_array<double>(std::string("wall_") + k, m)));
  const FluidParticleSoA fluid = fb.soa(n);
  const FluidParticleSoA fluid2 = fb2.soa(n);
  const WallParticleSoA walls{m, w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8]};

  // cell lists and neighbor tables (neighbor_search.cuh:10-24)
  const size_t cells = config.cell_capacity();
  CellListView fc{n, cells, device_zeros<index_t>(n), device_zeros<index_t>(n),
                  device_zeros<index_t>(cells), device_zeros<index_t>(cells)};
  CellListView wc{m, cells, device_zeros<index_t>(m), device_zeros<index_t>(m),
                  device_zeros<index_t>(cells), device_zeros<index_t>(cells)};
  build_cell_list(fluid, config.cell_grid(), fc);
  build_cell_list(walls, config.cell_grid(), wc);
  const size_t cap = n * max_neighbors;
  NeighborListView fn{n, cap, device_zeros<index_t>(n + 1), device_zeros<index_t>(cap)};
  NeighborListView wn{n, cap, device_zeros<index_t>(n + 1), device_zeros<index_t>(cap)};
  build_neighbor_lists(fluid, walls, config, fc, wc, fn, wn);

  // surface labels (surface_detection.cuh:31-36), on both state copies
  classify_surface_particles(fluid, walls, fn, wn, config);
  classify_surface_particles(fluid2, walls, fn, wn, config);

  // moment matrices, generation 1 (moment_matrix.cuh:36-68)
  const unsigned long long generation = 1;
  DeviceMomentMatrix moments(n, config);
  moments.prepare_matrices(fluid, walls, fn, wn, generation);
  const MomentMatrixView vview = moments.velocity_type_a();
  if (!vview.is_ready || vview.matrix_size != 9 || vview.particle_count < n)
    throw std::runtime_error("velocity view not ready");
  // all four views (moment_matrix.cuh:60-63), taken before anything moves the
  // particles: explicit inverses, traces, regularisations and factor status
  const MomentMatrixView pview = moments.pressure_type_a();
  const MomentMatrixView views[4] = {vview, pview, moments.pressure_type_b(),
                                     moments.fluid_only_type_a()};
  const char* view_names[4] = {"velocity_a", "pressure_a", "pressure_b", "fluid_only_a"};
  for (int k = 0; k < 4; ++k) {
    const MomentMatrixView& v = views[k];
    if (!v.is_ready || v.particle_count < n || static_cast<int>(v.kind) != k)
      throw std::runtime_error(std::string("view not ready: ") + view_names[k]);
    const size_t nn = static_cast<size_t>(v.matrix_size) * v.matrix_size;
    write_device(std::string("view_") + view_names[k] + "_inverse", v.inverse_matrices, n * nn);
    write_device(std::string("view_") + view_names[k] + "_trace", v.moment_trace, n);
    write_device(std::string("view_") + view_names[k] + "_regularization",
                 v.regularization_added, n);
    write_device(std::string("view_") + view_names[k] + "_info", v.info, n);
    const int32_t sizes[2] = {v.matrix_size, static_cast<int32_t>(v.basis_kind)};
    std::ofstream f(g_dir + "/out_view_" + view_names[k] + "_shape.bin", std::ios::binary);
    f.write(reinterpret_cast<const char*>(sizes), sizeof(sizes));
  }

  // provision (explicit_update.cuh:58-82)
  double* tmp = device_zeros<double>(3 * n);
  double* tmpw = device_zeros<double>(3 * m);
  const FluidParticleSoA temporary{n, nullptr, nullptr, nullptr, tmp, tmp + n, tmp + 2 * n, nullptr, nullptr};
  const WallParticleSoA temporary_wall{m, nullptr, nullptr, nullptr, tmpw, tmpw + m, tmpw + 2 * m,
                                       nullptr, nullptr, nullptr};
  DeviceProvisionExplicitUpdate provision(n, config);
  provision.compute_temporary_velocity(fluid, walls, fn, wn, moments, generation, temporary,
                                       temporary_wall);

  // PPE assembly and solve (ppe_matrix.cuh:64-118)
  DevicePpeMatrixAssembler assembler(n, n + cap, config);
  assembler.assemble(fluid, walls, fn, wn, temporary, temporary_wall, moments, generation);
  const PpeWorkspaceView ws = assembler.workspace();
  // the reference's free assembly (ppe_matrix.cuh:120-129), reading the views
  DevicePpeWorkspace view_ws(n, n + cap);
  assemble_ppe_matrix_and_rhs(fluid, walls, config, fn, wn, temporary, temporary_wall, vview,
                              pview, view_ws.view());
  AmgxPpeSolver solver;
  solver.solve(ws.matrix, ws.rhs, ws.pressure);
  std::printf("solve: %d iterations, relative residual %.3e, %d Schwarz blocks\n",
              solver.last_report().iterations, solver.last_report().relative_residual,
              solver.last_report().schwarz_blocks);

  // geometry + PPE outputs before the correction moves the particles
  int32_t used = 0;
  detail::hip_check(hipMemcpy(&used, ws.matrix.row_offsets + n, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
  int32_t fused = 0, wused = 0;
  detail::hip_check(hipMemcpy(&fused, fn.offsets + n, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
  detail::hip_check(hipMemcpy(&wused, wn.offsets + n, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
  write_device("fluid_cell_ids", fc.cell_ids, n);
  write_device("fluid_sorted_indices", fc.sorted_indices, n);
  write_device("fluid_cell_begin", fc.cell_begin, cells);
  write_device("fluid_cell_end", fc.cell_end, cells);
  write_device("fluid_offsets", fn.offsets, n + 1);
  write_device("fluid_indices", fn.indices, static_cast<size_t>(fused));
  write_device("wall_offsets", wn.offsets, n + 1);
  write_device("wall_indices", wn.indices, static_cast<size_t>(wused));
  write_device("surface_type", fb.surface_type, n);
  write_device("temporary_velocity", tmp, 3 * n);
  write_device("temporary_wall_velocity", tmpw, 3 * m);
  write_device("ppe_row_offsets", ws.matrix.row_offsets, n + 1);
  write_device("ppe_col_indices", ws.matrix.col_indices, static_cast<size_t>(used));
  write_device("ppe_values", ws.matrix.values, static_cast<size_t>(used));
  write_device("ppe_rhs", ws.rhs, n);
  write_device("pressure", ws.pressure, n);
  {
    int32_t vused = 0;
    detail::hip_check(hipMemcpy(&vused, view_ws.view().matrix.row_offsets + n, sizeof(int32_t),
                                hipMemcpyDeviceToHost), "hipMemcpy");
    write_device("view_ppe_row_offsets", view_ws.view().matrix.row_offsets, n + 1);
    write_device("view_ppe_col_indices", view_ws.view().matrix.col_indices, static_cast<size_t>(vused));
    write_device("view_ppe_values", view_ws.view().matrix.values, static_cast<size_t>(vused));
    write_device("view_ppe_rhs", view_ws.view().rhs, n);
  }

  // free-function replay of the correction on the second copy, from the same pressure
  double* p2 = device_zeros<double>(n);
  detail::hip_check(hipMemcpy(p2, ws.pressure, n * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
  double* vec2 = device_zeros<double>(12 * n);
  CorrectionVectorView grad2{n, vec2, vec2 + n, vec2 + 2 * n};
  CorrectionVectorView ps2{n, vec2 + 3 * n, vec2 + 4 * n, vec2 + 5 * n};
  CorrectionVectorView smooth2{n, vec2 + 6 * n, vec2 + 7 * n, vec2 + 8 * n};
  clamp_particle_pressure(fluid2, p2);
  compute_pressure_gradient(fluid2, walls, fn, wn, p2, config, moments, grad2);
  // the reference's signature (pressure_correction.cuh:93-100), reading the view
  CorrectionVectorView grad_view{n, vec2 + 9 * n, vec2 + 10 * n, vec2 + 11 * n};
  compute_pressure_gradient(fluid2, walls, fn, wn, p2, config, pview, grad_view);
  compute_particle_shifting_displacement(fluid2, walls, fn, wn, config, ps2);
  update_velocity_and_position(fluid2, temporary, config, grad2, ps2);
  const lsmps_config cc = detail::to_c(config);
  detail::check(lsmps_apply_anti_penetration(default_context(), &fluid2, &walls, &wn, &cc),
                "lsmps_apply_anti_penetration");
  smooth_particle_velocity(fluid2, fn, config, smooth2);

  // pressure correction (pressure_correction.cuh:57-89)
  DevicePressureCorrection correction(n, config);
  correction.apply(fluid, walls, fn, wn, temporary, ws.pressure, moments, generation);
  const CorrectionVectorView grad = correction.pressure_gradient();
  const CorrectionVectorView ps = correction.ps_displacement();
  write_device("grad", grad.x, 3 * n);  // x, y, z contiguous in the workspace
  write_device("ps", ps.x, 3 * n);
  for (int k = 0; k < 7; ++k) {
    write_device(std::string("fluid_") + kFluidNames[k], fb.f[k], n);
    write_device(std::string("replay_fluid_") + kFluidNames[k], fb2.f[k], n);
  }
  write_device("replay_grad", vec2, 3 * n);
  write_device("replay_ps", vec2 + 3 * n, 3 * n);
  write_device("view_grad", vec2 + 9 * n, 3 * n);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  try {
    return run();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "module_step: %s\n", e.what());
    return 1;
  }
}
