// Surface-detection acceptance driver on five free-surface geometries, over the
// library's C++ module API (the reference's tests/reference/
// complex_surface_diagnostics.cu:151-632).
//
//   complex_surface_diagnostics [case|all] [output_dir]
//
// Cases: inclined_plane, sine_wave, droplet, cylinder, stepped_box, each a
// 1 m box at 0.02 m spacing with its analytic free surface.  Fluid particles
// within one spacing of that surface are the expected Surface set, with the
// analytic outward normal.  For each case the program classifies the fluid with
// every diagnostic, writes the fluid and wall VTK files and a per-particle CSV,
// and prints the label counts, the detected fraction of the expected surface,
// the false Surface count and the mean angle between detected and analytic
// normals.  The geometry, file layout and printed lines follow the reference
// program so that its outputs can be compared line by line.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsmps3d_amd.hpp"

namespace {

using lsmps3d::real;
constexpr real kPi = 3.14159265358979323846;
constexpr real kSpacing = 0.02;
constexpr real kBox = 1.0;

struct PointSet {
  std::vector<real> x, y, z, nx, ny, nz;
  void add(real px, real py, real pz, real ux, real uy, real uz) {
    x.push_back(px);
    y.push_back(py);
    z.push_back(pz);
    nx.push_back(ux);
    ny.push_back(uy);
    nz.push_back(uz);
  }
  size_t size() const { return x.size(); }
};

struct Geometry {
  std::string name;
  PointSet fluid;          // normals = expected surface normal (0 when not expected)
  std::vector<int> expected;
  PointSet walls;
  void fluid_at(real px, real py, real pz, bool on_surface, real ux = 0, real uy = 0, real uz = 0) {
    fluid.add(px, py, pz, ux, uy, uz);
    expected.push_back(on_surface ? 1 : 0);
  }
};

void unit(real& a, real& b, real& c) {
  const real len = std::sqrt(a * a + b * b + c * c);
  if (len <= 0) return;
  a /= len;
  b /= len;
  c /= len;
}

// cell centre of lattice index i
real centre(int i, real h) { return (static_cast<real>(i) + 0.5) * h; }

// Side walls x = 0 / x = n h (all of y, z), y = 0 / y = n h (interior x), then the
// floor (interior x, y), each normal pointing into the box.
void box_walls(PointSet& w, int n, real h) {
  for (int k = 0; k <= n; ++k) {
    for (int j = 0; j <= n; ++j) {
      w.add(0.0, j * h, k * h, 1, 0, 0);
      w.add(n * h, j * h, k * h, -1, 0, 0);
    }
    for (int i = 1; i < n; ++i) {
      w.add(i * h, 0.0, k * h, 0, 1, 0);
      w.add(i * h, n * h, k * h, 0, -1, 0);
    }
  }
  for (int j = 1; j < n; ++j)
    for (int i = 1; i < n; ++i) w.add(i * h, j * h, 0.0, 0, 0, 1);
}

int cells_per_side(real h) { return static_cast<int>(std::lround(kBox / h)); }

// Visits the lattice cell centres in z, y, x order.
void for_each_centre(int n, real h, const std::function<void(real, real, real)>& f) {
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) f(centre(i, h), centre(j, h), centre(k, h));
}

Geometry inclined_plane(real h) {
  Geometry g{"inclined_plane", {}, {}, {}};
  const int n = cells_per_side(h);
  box_walls(g.walls, n, h);
  const real slope = 0.25;
  real ux = -slope, uy = 0, uz = 1;
  unit(ux, uy, uz);
  for_each_centre(n, h, [&](real x, real y, real z) {
    const real top = 0.35 + slope * x;
    if (z <= top) g.fluid_at(x, y, z, top - z <= h, ux, uy, uz);
  });
  return g;
}

Geometry sine_wave(real h) {
  Geometry g{"sine_wave", {}, {}, {}};
  const int n = cells_per_side(h);
  box_walls(g.walls, n, h);
  const real a = 0.08;
  for_each_centre(n, h, [&](real x, real y, real z) {
    const real sx = std::sin(2 * kPi * x), sy = std::sin(2 * kPi * y);
    const real cx = std::cos(2 * kPi * x), cy = std::cos(2 * kPi * y);
    const real top = 0.5 + a * sx * sy;
    if (z > top) return;
    real ux = -a * 2 * kPi * cx * sy, uy = -a * 2 * kPi * sx * cy, uz = 1;
    unit(ux, uy, uz);
    g.fluid_at(x, y, z, top - z <= h, ux, uy, uz);
  });
  return g;
}

Geometry droplet(real h) {
  Geometry g{"droplet", {}, {}, {}};
  const int n = cells_per_side(h);
  box_walls(g.walls, n, h);
  const real c0 = 0.5, c1 = 0.5, c2 = 0.25, radius = 0.28;
  for_each_centre(n, h, [&](real x, real y, real z) {
    const real dx = x - c0, dy = y - c1, dz = z - c2;
    const real r = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (!(r <= radius && z >= 0)) return;
    real ux = dx, uy = dy, uz = dz;
    unit(ux, uy, uz);
    g.fluid_at(x, y, z, radius - r <= h, ux, uy, uz);
  });
  return g;
}

Geometry cylinder(real h) {
  Geometry g{"cylinder", {}, {}, {}};
  const int n = cells_per_side(h);
  box_walls(g.walls, n, h);
  const real c0 = 0.5, c1 = 0.5, radius = 0.12, level = 0.5;
  for (int k = 0; k <= n; ++k) {  // 48 wall particles per ring, normals outward from the axis
    const real z = k * h;
    for (int s = 0; s < 48; ++s) {
      const real theta = s * 2 * kPi / 48;
      const real ux = std::cos(theta), uy = std::sin(theta);
      g.walls.add(c0 + radius * ux, c1 + radius * uy, z, ux, uy, 0);
    }
  }
  for_each_centre(n, h, [&](real x, real y, real z) {
    if (z > level) return;
    const real r = std::sqrt((x - c0) * (x - c0) + (y - c1) * (y - c1));
    if (r >= radius + 0.5 * h) g.fluid_at(x, y, z, level - z <= h, 0, 0, 1);
  });
  return g;
}

Geometry stepped_box(real h) {
  Geometry g{"stepped_box", {}, {}, {}};
  const int n = cells_per_side(h);
  box_walls(g.walls, n, h);
  const real level = 0.5, step_x = 0.5, step_top = 0.18;
  for (int j = 0; j <= n; ++j)  // riser at x = step_x
    for (int k = 0; k <= static_cast<int>(std::lround(step_top / h)); ++k)
      g.walls.add(step_x, j * h, k * h, 1, 0, 0);
  for (int i = 0; i <= static_cast<int>(std::lround(step_x / h)); ++i)  // tread at z = step_top
    for (int j = 0; j <= n; ++j) g.walls.add(i * h, j * h, step_top, 0, 0, 1);
  for_each_centre(n, h, [&](real x, real y, real z) {
    const real bottom = x < step_x ? step_top : 0.0;
    if (z >= bottom && z <= level) g.fluid_at(x, y, z, level - z <= h, 0, 0, 1);
  });
  return g;
}

Geometry make_geometry(const std::string& name, real h) {
  if (name == "inclined_plane") return inclined_plane(h);
  if (name == "sine_wave") return sine_wave(h);
  if (name == "droplet") return droplet(h);
  if (name == "cylinder") return cylinder(h);
  if (name == "stepped_box") return stepped_box(h);
  throw std::invalid_argument("Unknown case: " + name);
}

struct Detection {
  std::vector<int> fluid_count, wall_count, label;
  std::vector<real> density, ratio, anisotropy, open_ratio, air_anisotropy, nx, ny, nz;
};

template <typename T>
T* upload(const std::vector<T>& host, T* device) {
  lsmps3d::detail::hip_check(
      hipMemcpy(device, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return device;
}
template <typename T>
std::vector<T> download(const T* device, size_t n) {
  std::vector<T> h(n);
  lsmps3d::detail::hip_check(hipMemcpy(h.data(), device, n * sizeof(T), hipMemcpyDeviceToHost),
                             "hipMemcpy");
  return h;
}

Detection detect(const Geometry& g, real h) {
  const real re = 3.1 * h;
  const int dim = static_cast<int>(std::ceil((kBox + 2.0 * re) / re));
  const size_t nf = g.fluid.size(), nw = g.walls.size();
  lsmps3d::SimulationWorkspace ws(lsmps3d::WorkspaceSpec{nf, nw, 256, 256,
                                                         static_cast<size_t>(dim) * dim * dim});
  auto v = ws.view();
  upload(g.fluid.x, v.fluid.x);
  upload(g.fluid.y, v.fluid.y);
  upload(g.fluid.z, v.fluid.z);
  upload(g.walls.x, v.walls.x);
  upload(g.walls.y, v.walls.y);
  upload(g.walls.z, v.walls.z);
  upload(g.walls.nx, v.walls.normal_x);
  upload(g.walls.ny, v.walls.normal_y);
  upload(g.walls.nz, v.walls.normal_z);

  lsmps3d::SimulationConfig cfg;
  cfg.particle_spacing = h;
  cfg.support_radius = re;
  cfg.near_surface_radius = 2.0 * h;
  cfg.cell_origin = lsmps3d::Vec3{-re, -re, -re};
  cfg.cell_size = re;
  cfg.cell_dims = lsmps3d::Int3{dim, dim, dim};
  cfg.splash_neighbor_threshold = 12;
  cfg.number_density_ratio_threshold = 0.85;
  cfg.air_open_ratio_threshold = 0.33;
  cfg.air_anisotropy_threshold = 0.033;
  cfg.include_wall_neighbors = true;
  cfg.wall_normal_independence_threshold = 0.25;
  lsmps3d::build_neighbor_lists(v.fluid, v.walls, cfg, v.fluid_cells, v.wall_cells,
                                v.fluid_neighbors, v.wall_neighbors);

  // Diagnostic outputs in owned device arrays (the reference borrows spare
  // particle and neighbor-list buffers for them).
  lsmps3d::DeviceNeighborList counts(nf, nf);
  lsmps3d::DeviceFluidParticles reals(nf), normals(nf);
  const auto c = counts.view();
  const auto r = reals.view();
  const auto u = normals.view();
  lsmps3d::classify_surface_particles(
      v.fluid, v.walls, v.fluid_neighbors, v.wall_neighbors, cfg,
      lsmps3d::SurfaceDetectionDiagnosticsView{c.offsets, c.indices, r.x, r.y, r.z, r.vx, r.vy,
                                               u.x, u.y, u.z});
  Detection d;
  d.fluid_count = download(c.offsets, nf);
  d.wall_count = download(c.indices, nf);
  d.density = download(r.x, nf);
  d.ratio = download(r.y, nf);
  d.anisotropy = download(r.z, nf);
  d.open_ratio = download(r.vx, nf);
  d.air_anisotropy = download(r.vy, nf);
  d.nx = download(u.x, nf);
  d.ny = download(u.y, nf);
  d.nz = download(u.z, nf);
  d.label = download(v.fluid.surface_type, nf);
  return d;
}

const char* label_name(int t) {
  switch (t) {
    case 0: return "Inner";
    case 1: return "NearSurface";
    case 2: return "Surface";
    case 3: return "Splash";
    default: return "Unknown";
  }
}

void write_outputs(const std::filesystem::path& dir, const Geometry& g, const Detection& d) {
  std::filesystem::create_directories(dir);
  lsmps3d::HostVtkPointFields f;
  f.add_scalar("fluid_neighbor_count", d.fluid_count);
  f.add_scalar("wall_neighbor_count", d.wall_count);
  f.add_scalar("number_density", d.density);
  f.add_scalar("number_density_ratio", d.ratio);
  f.add_scalar("anisotropy", d.anisotropy);
  f.add_scalar("air_open_ratio", d.open_ratio);
  f.add_scalar("air_anisotropy", d.air_anisotropy);
  f.add_scalar("surface_type", d.label);
  f.add_scalar("expected_surface", g.expected);
  f.add_vector("surface_normal", d.nx, d.ny, d.nz);
  f.add_vector("expected_normal", g.fluid.nx, g.fluid.ny, g.fluid.nz);
  lsmps3d::SimulationConfig out;
  out.output_directory = dir;
  out.vtk_file_prefix = g.name;
  out.vtk_write_point_fields = true;
  lsmps3d::LegacyVtkWriter(out).write(0, {g.fluid.x, g.fluid.y, g.fluid.z}, f);
  lsmps3d::HostVtkPointFields wf;
  wf.add_vector("wall_normal", g.walls.nx, g.walls.ny, g.walls.nz);
  out.vtk_file_prefix = g.name + "_walls";
  lsmps3d::LegacyVtkWriter(out).write(0, {g.walls.x, g.walls.y, g.walls.z}, wf);

  const auto csv = dir / (g.name + "_debug.csv");
  std::ofstream o(csv);
  if (!o) throw std::runtime_error("Failed to open CSV output: " + csv.string());
  o << std::setprecision(9);
  o << "particle_id,x,y,z,fluid_neighbor_count,wall_neighbor_count,number_density,"
       "number_density_ratio,anisotropy,air_open_ratio,air_anisotropy,surface_normal_x,"
       "surface_normal_y,surface_normal_z,surface_type,surface_type_name,expected_surface,"
       "expected_normal_x,expected_normal_y,expected_normal_z\n";
  for (size_t i = 0; i < g.fluid.size(); ++i)
    o << i << ',' << g.fluid.x[i] << ',' << g.fluid.y[i] << ',' << g.fluid.z[i] << ','
      << d.fluid_count[i] << ',' << d.wall_count[i] << ',' << d.density[i] << ',' << d.ratio[i]
      << ',' << d.anisotropy[i] << ',' << d.open_ratio[i] << ',' << d.air_anisotropy[i] << ','
      << d.nx[i] << ',' << d.ny[i] << ',' << d.nz[i] << ',' << d.label[i] << ','
      << label_name(d.label[i]) << ',' << g.expected[i] << ',' << g.fluid.nx[i] << ','
      << g.fluid.ny[i] << ',' << g.fluid.nz[i] << '\n';
}

void report(const std::filesystem::path& dir, const Geometry& g, const Detection& d) {
  std::array<size_t, 4> by_label{};
  size_t expected = 0, found = 0, false_surface = 0, angles = 0;
  real angle_sum = 0;
  for (size_t i = 0; i < d.label.size(); ++i) {
    const int t = d.label[i];
    if (t >= 0 && t < 4) ++by_label[static_cast<size_t>(t)];
    const bool want = g.expected[i] != 0, got = t == 2;
    if (want) {
      ++expected;
      if (got) ++found;
    } else if (got) {
      ++false_surface;
    }
    if (want && got) {
      const real dot = d.nx[i] * g.fluid.nx[i] + d.ny[i] * g.fluid.ny[i] + d.nz[i] * g.fluid.nz[i];
      angle_sum += std::acos(std::max<real>(-1, std::min<real>(1, dot))) * 180 / kPi;
      ++angles;
    }
  }
  const real mean_angle = angles > 0 ? angle_sum / static_cast<real>(angles) : 0;
  const double detected = expected > 0 ? static_cast<double>(found) / static_cast<double>(expected) : 0.0;
  std::cout << "Complex surface diagnostics: " << g.name << '\n'
            << "  fluid particles: " << g.fluid.size() << '\n'
            << "  wall particles: " << g.walls.size() << '\n'
            << "  Inner: " << by_label[0] << '\n'
            << "  NearSurface: " << by_label[1] << '\n'
            << "  Surface: " << by_label[2] << '\n'
            << "  Splash: " << by_label[3] << '\n'
            << "  Expected-surface particles: " << expected << '\n'
            << "  Expected-surface detected: " << found << " (" << detected << ")\n"
            << "  False Surface particles: " << false_surface << '\n'
            << "  Mean normal angle error: " << mean_angle << " deg\n"
            << "  VTK: " << (dir / (g.name + "_000000.vtk")) << '\n'
            << "  Wall VTK: " << (dir / (g.name + "_walls_000000.vtk")) << '\n'
            << "  CSV: " << (dir / (g.name + "_debug.csv")) << std::endl;
}

void run_case(const std::string& name, const std::filesystem::path& base) {
  const Geometry g = make_geometry(name, kSpacing);
  const Detection d = detect(g, kSpacing);
  const auto dir = base / g.name;
  write_outputs(dir, g, d);
  report(dir, g, d);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string which = argc > 1 ? argv[1] : "inclined_plane";
    const std::filesystem::path base =
        argc > 2 ? std::filesystem::path(argv[2]) : std::filesystem::path("output/complex_surface_diagnostics");
    if (which == "all") {
      for (const char* name : {"inclined_plane", "sine_wave", "droplet", "cylinder", "stepped_box"})
        run_case(name, base);
    } else {
      run_case(which, base);
    }
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "complex_surface_diagnostics failed: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
