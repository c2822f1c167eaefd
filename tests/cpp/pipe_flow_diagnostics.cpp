// Poiseuille pipe fixture over the library's C++ module API (the reference's
// tests/reference/pipe_flow_diagnostics.cu).
//
//   pipe_flow_diagnostics [output_dir]
//
// A 0.6 m pipe of radius 0.2 m at 0.025 m spacing: fluid on the lattice
// (x at cell centres, y and z at nodes) inside the radius with the parabolic
// profile u_x = U (1 - r^2 / R^2), and wall rings at x = cell centres with
// inward normals.  The program builds the neighbor lists for the pipe's cell
// grid, writes the profile and its analytic gradient and Laplacian
// (-4 U / R^2) to VTK, and prints the case summary.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <iostream>
#include <vector>

#include "lsmps3d_amd.hpp"

namespace {

using lsmps3d::real;

struct Pipe {
  std::vector<real> x, y, z, u, gx, gy, gz, lap;
  std::vector<real> wx, wy, wz, wnx, wny, wnz, wu;
};

Pipe make_pipe(real h, real length, real radius, real u_max) {
  Pipe p;
  const int nx = static_cast<int>(std::lround(length / h));
  const int reach = static_cast<int>(std::ceil(radius / h));
  const real r2max = radius * radius;
  const real inv_r2 = 1.0 / r2max;
  for (int i = 0; i < nx; ++i) {
    const real x = (static_cast<real>(i) + 0.5) * h;
    for (int j = -reach; j <= reach; ++j) {
      const real y = static_cast<real>(j) * h;
      for (int k = -reach; k <= reach; ++k) {
        const real z = static_cast<real>(k) * h;
        const real r2 = y * y + z * z;
        if (r2 > r2max) continue;
        p.x.push_back(x);
        p.y.push_back(y);
        p.z.push_back(z);
        p.u.push_back(u_max * (1.0 - r2 * inv_r2));
        p.gx.push_back(0.0);
        p.gy.push_back(-2.0 * u_max * y * inv_r2);
        p.gz.push_back(-2.0 * u_max * z * inv_r2);
        p.lap.push_back(-4.0 * u_max * inv_r2);
      }
    }
  }
  const int segments = std::max(24, static_cast<int>(std::ceil(2 * M_PI * radius / h)));
  for (int i = 0; i < nx; ++i) {
    const real x = (static_cast<real>(i) + 0.5) * h;
    for (int s = 0; s < segments; ++s) {
      const real theta = static_cast<real>(2.0 * M_PI) * static_cast<real>(s) / static_cast<real>(segments);
      p.wx.push_back(x);
      p.wy.push_back(radius * std::cos(theta));
      p.wz.push_back(radius * std::sin(theta));
      p.wnx.push_back(0.0);
      p.wny.push_back(-std::cos(theta));
      p.wnz.push_back(-std::sin(theta));
      p.wu.push_back(0.0);
    }
  }
  return p;
}

template <typename T>
void upload(T* device, const std::vector<T>& host) {
  lsmps3d::detail::hip_check(
      hipMemcpy(device, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

int run(const std::filesystem::path& out_dir) {
  std::filesystem::create_directories(out_dir);
  const real h = 0.025, length = 0.6, radius = 0.2, u_max = 1.0, re = 3.1 * h;
  const Pipe p = make_pipe(h, length, radius, u_max);
  const size_t nf = p.x.size(), nw = p.wx.size();
  const lsmps3d::Vec3 lo{-re, -radius - re, -radius - re};
  const lsmps3d::Vec3 hi{length + re, radius + re, radius + re};
  const lsmps3d::Int3 dims{static_cast<lsmps3d::index_t>(std::ceil((hi.x - lo.x) / re)),
                           static_cast<lsmps3d::index_t>(std::ceil((hi.y - lo.y) / re)),
                           static_cast<lsmps3d::index_t>(std::ceil((hi.z - lo.z) / re))};
  const size_t cells = static_cast<size_t>(dims.x) * dims.y * dims.z;
  lsmps3d::SimulationWorkspace ws(lsmps3d::WorkspaceSpec{nf, nw, 256, 128, cells});
  auto v = ws.view();
  upload(v.fluid.x, p.x);
  upload(v.fluid.y, p.y);
  upload(v.fluid.z, p.z);
  upload(v.fluid.vx, p.u);
  upload(v.walls.x, p.wx);
  upload(v.walls.y, p.wy);
  upload(v.walls.z, p.wz);
  upload(v.walls.normal_x, p.wnx);
  upload(v.walls.normal_y, p.wny);
  upload(v.walls.normal_z, p.wnz);
  upload(v.walls.vx, p.wu);

  lsmps3d::SimulationConfig cfg;
  cfg.support_radius = re;
  cfg.cell_origin = lo;
  cfg.cell_size = re;
  cfg.cell_dims = dims;
  cfg.density = 1.0;
  cfg.gravity = {};
  lsmps3d::build_neighbor_lists(v.fluid, v.walls, cfg, v.fluid_cells, v.wall_cells,
                                v.fluid_neighbors, v.wall_neighbors);

  std::vector<real> gmag(nf);
  for (size_t i = 0; i < nf; ++i)
    gmag[i] = std::sqrt(p.gx[i] * p.gx[i] + p.gy[i] * p.gy[i] + p.gz[i] * p.gz[i]);
  lsmps3d::HostVtkPointFields f;
  f.add_scalar("velocity_x", p.u);
  f.add_scalar("analytic_velocity_gradient_magnitude", gmag);
  f.add_scalar("analytic_velocity_laplacian", p.lap);
  f.add_vector("velocity", p.u, std::vector<real>(nf), std::vector<real>(nf));
  f.add_vector("analytic_velocity_gradient", p.gx, p.gy, p.gz);
  lsmps3d::SimulationConfig out = cfg;
  out.output_directory = out_dir;
  out.vtk_file_prefix = "pipe_flow";
  out.vtk_write_point_fields = true;
  const lsmps3d::LegacyVtkWriter writer(out);
  writer.write(0, {p.x, p.y, p.z}, f);
  std::cout << "Pipe flow diagnostics\n"
            << "  fluid particles: " << nf << '\n'
            << "  wall particles: " << nw << '\n'
            << "  radius: " << radius << " m\n"
            << "  spacing: " << h << " m\n"
            << "  support radius: " << re << " m\n"
            << "  analytic laplacian: " << -4.0 * u_max / (radius * radius) << '\n'
            << "  VTK: " << writer.make_path(0) << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc > 1 ? std::filesystem::path(argv[1])
                        : std::filesystem::path("output/pipe_flow_diagnostics"));
  } catch (const std::exception& e) {
    std::cerr << "pipe_flow_diagnostics failed: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
