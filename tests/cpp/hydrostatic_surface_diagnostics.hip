// Hydrostatic acceptance driver over the library's C++ module API (the
// reference's tests/reference/hydrostatic_surface_diagnostics.cu:489-820).
//
//   hydrostatic_surface_diagnostics [output_dir]
//
// A 1 m box filled to 0.5 m at 0.02 m spacing (62,500 fluid particles) with
// five wall faces.  The program labels the surface with every diagnostic,
// builds the moment operators, the provisional velocity and the PPE, solves it
// with the library's solver, evaluates the LSMPS pressure gradient and Laplacian
// of the solution (a kernel of this program, as in the reference), and compares
// them with the hydrostatic field p = rho g (H - l0/2 - z): the gradient should
// be (0, 0, -rho g) and the Laplacian 0.  It also evaluates the residual of the
// assembled system at the hydrostatic pressure, with the wall Neumann term added
// and removed once and twice.  Outputs: one VTK file with every per-particle
// field, a CSV of the surface diagnostics, and the summary lines the reference
// prints.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "lsmps3d_amd.hpp"

namespace {

using lsmps3d::real;
constexpr real kBox = 1.0;
constexpr real kLevel = 0.5;
constexpr real kSpacing = 0.02;
constexpr real kRe = 3.1 * kSpacing;
constexpr real kNearRadius = 2.0 * kSpacing;
constexpr real kRho = 1000.0;
constexpr real kG = 9.81;

struct Box {
  std::vector<real> fx, fy, fz;
  std::vector<real> wx, wy, wz, wnx, wny, wnz;
};

// Fluid at cell centres below kLevel (z-major, then y, then x); walls at the
// lattice nodes of the faces x = 0, x = 1, y = 0, y = 1 and z = 0 (normals into
// the box), side faces first, ring by ring in z, then the floor's interior.
Box half_full_box(real h, real size, real level) {
  Box b;
  const int n = static_cast<int>(std::lround(size / h));
  const int layers = static_cast<int>(std::lround(level / h));
  for (int k = 0; k < layers; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        b.fx.push_back((static_cast<real>(i) + 0.5) * h);
        b.fy.push_back((static_cast<real>(j) + 0.5) * h);
        b.fz.push_back((static_cast<real>(k) + 0.5) * h);
      }
  auto wall = [&](int i, int j, int k, real ux, real uy, real uz) {
    b.wx.push_back(static_cast<real>(i) * h);
    b.wy.push_back(static_cast<real>(j) * h);
    b.wz.push_back(static_cast<real>(k) * h);
    b.wnx.push_back(ux);
    b.wny.push_back(uy);
    b.wnz.push_back(uz);
  };
  for (int k = 0; k <= n; ++k) {
    for (int j = 0; j <= n; ++j) {
      wall(0, j, k, 1, 0, 0);
      wall(n, j, k, -1, 0, 0);
    }
    for (int i = 1; i < n; ++i) {
      wall(i, 0, k, 0, 1, 0);
      wall(i, n, k, 0, -1, 0);
    }
  }
  for (int j = 1; j < n; ++j)
    for (int i = 1; i < n; ++i) wall(i, j, 0, 0, 0, 1);
  return b;
}

// ---- the pressure operator of a given field (gradient and Laplacian) --------
// LSMPS Type-A reconstruction from the pressure operator's inverse moment
// matrices: rhs = sum_j w_ij b(d_ij / re) (p_j - p_i) over fluid neighbors plus
// sum_k s_w w_ik q_k (re rho g . n_k) over wall neighbors; c = M^-1 rhs;
// grad p = c[0..2] / re, lap p = 2 (c3 + c4 + c5) / re^2.

struct OperatorArgs {
  size_t n;
  const real *x, *y, *z, *wx, *wy, *wz, *wnx, *wny, *wnz;
  const int32_t *f_off, *f_idx, *w_off, *w_idx;
  const real* inverse;  // 81 per particle, column-major
  const real* p;
  real re, wall_scale, rho;
  real gx, gy, gz;
  real *grad_x, *grad_y, *grad_z, *lap;
};

__global__ void __launch_bounds__(256) pressure_operator_kernel(OperatorArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const real px = a.x[i], py = a.y[i], pz = a.z[i], pi = a.p[i];
  const real inv_re = 1.0 / a.re;
  real rhs[9] = {};
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const real dx = a.x[j] - px, dy = a.y[j] - py, dz = a.z[j] - pz;
    const real r = sqrt(dx * dx + dy * dy + dz * dz);
    const real w = r >= a.re ? 0.0 : 1.0 - r / a.re;
    if (w <= 0.0) continue;
    const real sx = dx * inv_re, sy = dy * inv_re, sz = dz * inv_re;
    const real b[9] = {sx, sy, sz, sx * sx, sy * sy, sz * sz, sx * sy, sy * sz, sz * sx};
    const real dp = a.p[j] - pi;
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[k] += w * b[k] * dp;
  }
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const real dx = a.wx[j] - px, dy = a.wy[j] - py, dz = a.wz[j] - pz;
    const real r = sqrt(dx * dx + dy * dy + dz * dz);
    const real w = a.wall_scale * (r >= a.re ? 0.0 : 1.0 - r / a.re);
    if (w <= 0.0) continue;
    const real nx = a.wnx[j], ny = a.wny[j], nz = a.wnz[j];
    const real q[9] = {nx,
                       ny,
                       nz,
                       2.0 * dx * nx * inv_re,
                       2.0 * dy * ny * inv_re,
                       2.0 * dz * nz * inv_re,
                       (dy * nx + dx * ny) * inv_re,
                       (dz * ny + dy * nz) * inv_re,
                       (dx * nz + dz * nx) * inv_re};
    const real s = a.re * a.rho * (a.gx * nx + a.gy * ny + a.gz * nz);
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[k] += w * q[k] * s;
  }
  const real* m = a.inverse + i * 81;
  real c[6];
#pragma unroll
  for (int row = 0; row < 6; ++row) {
    real v = 0.0;
#pragma unroll
    for (int col = 0; col < 9; ++col) v += m[row + col * 9] * rhs[col];
    c[row] = v;
  }
  a.grad_x[i] = c[0] / a.re;
  a.grad_y[i] = c[1] / a.re;
  a.grad_z[i] = c[2] / a.re;
  a.lap[i] = 2.0 * (c[3] + c[4] + c[5]) / (a.re * a.re);
}

template <typename T>
void upload(T* device, const std::vector<T>& host) {
  lsmps3d::detail::hip_check(
      hipMemcpy(device, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}
template <typename T>
std::vector<T> download(const T* device, size_t n) {
  std::vector<T> h(n);
  lsmps3d::detail::hip_check(hipMemcpy(h.data(), device, n * sizeof(T), hipMemcpyDeviceToHost),
                             "hipMemcpy");
  return h;
}
void zero(real* device, size_t n) {
  lsmps3d::detail::hip_check(hipMemset(device, 0, n * sizeof(real)), "hipMemset");
}

std::vector<real> minus(const std::vector<real>& a, const std::vector<real>& b) {
  std::vector<real> o(a.size());
  for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] - b[i];
  return o;
}
std::vector<real> abs_of(const std::vector<real>& a) {
  std::vector<real> o(a.size());
  for (size_t i = 0; i < a.size(); ++i) o[i] = std::abs(a[i]);
  return o;
}
real max_of(const std::vector<real>& a) {
  real m = 0.0;
  for (real v : a) m = std::max(m, v);
  return m;
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

int run(const std::filesystem::path& out_dir) {
  std::filesystem::create_directories(out_dir);
  const Box box = half_full_box(kSpacing, kBox, kLevel);
  const size_t nf = box.fx.size(), nw = box.wx.size();
  const int dim = static_cast<int>(std::ceil((kBox + 2.0 * kRe) / kRe));
  lsmps3d::SimulationWorkspace ws(
      lsmps3d::WorkspaceSpec{nf, nw, 256, 256, static_cast<size_t>(dim) * dim * dim});
  auto v = ws.view();
  upload(v.fluid.x, box.fx);
  upload(v.fluid.y, box.fy);
  upload(v.fluid.z, box.fz);
  zero(v.fluid.vx, nf);
  zero(v.fluid.vy, nf);
  zero(v.fluid.vz, nf);
  std::vector<real> p_true(nf);
  const real surface_z = kLevel - 0.5 * kSpacing;  // centre of the top layer
  for (size_t i = 0; i < nf; ++i) p_true[i] = kRho * kG * std::max(surface_z - box.fz[i], 0.0);
  upload(v.fluid.pressure, p_true);
  upload(v.walls.x, box.wx);
  upload(v.walls.y, box.wy);
  upload(v.walls.z, box.wz);
  upload(v.walls.normal_x, box.wnx);
  upload(v.walls.normal_y, box.wny);
  upload(v.walls.normal_z, box.wnz);
  zero(v.walls.vx, nw);
  zero(v.walls.vy, nw);
  zero(v.walls.vz, nw);

  lsmps3d::SimulationConfig cfg;
  cfg.particle_spacing = kSpacing;
  cfg.support_radius = kRe;
  cfg.near_surface_radius = kNearRadius;
  cfg.cell_origin = lsmps3d::Vec3{-kRe, -kRe, -kRe};
  cfg.cell_size = kRe;
  cfg.cell_dims = lsmps3d::Int3{dim, dim, dim};
  cfg.splash_neighbor_threshold = 12;
  cfg.number_density_ratio_threshold = 0.85;
  cfg.air_open_ratio_threshold = 0.33;
  cfg.air_anisotropy_threshold = 0.05;
  cfg.include_wall_neighbors = true;
  cfg.wall_normal_independence_threshold = 0.25;
  cfg.density = kRho;
  cfg.gravity = lsmps3d::Vec3{0.0, 0.0, -kG};
  lsmps3d::build_neighbor_lists(v.fluid, v.walls, cfg, v.fluid_cells, v.wall_cells,
                                v.fluid_neighbors, v.wall_neighbors);

  lsmps3d::DeviceNeighborList counts(nf, nf);
  lsmps3d::DeviceFluidParticles reals(nf), normals(nf);
  const auto cd = counts.view();
  const auto rd = reals.view();
  const auto nd = normals.view();
  lsmps3d::classify_surface_particles(
      v.fluid, v.walls, v.fluid_neighbors, v.wall_neighbors, cfg,
      lsmps3d::SurfaceDetectionDiagnosticsView{cd.offsets, cd.indices, rd.x, rd.y, rd.z, rd.vx,
                                               rd.vy, nd.x, nd.y, nd.z});

  lsmps3d::DeviceMomentMatrix moments(nf, cfg);
  moments.prepare_matrices(v.fluid, v.walls, v.fluid_neighbors, v.wall_neighbors, 1);
  lsmps3d::DeviceFluidParticles tmp(nf);
  lsmps3d::DeviceWallParticles tmp_walls(nw);
  lsmps3d::DeviceProvisionExplicitUpdate provision(nf, cfg);
  provision.compute_temporary_velocity(v.fluid, v.walls, v.fluid_neighbors, v.wall_neighbors,
                                       moments, 1, tmp.view(), tmp_walls.view());
  lsmps3d::DevicePpeMatrixAssembler assembler(nf, nf + v.fluid_neighbors.neighbor_count, cfg);
  assembler.assemble(v.fluid, v.walls, v.fluid_neighbors, v.wall_neighbors, tmp.view(),
                     tmp_walls.view(), moments, 1);
  auto ppe = assembler.workspace();
  lsmps3d::AmgxPpeSolver solver(cfg.amgx_config_path);
  solver.solve(ppe.matrix, ppe.rhs, ppe.pressure);
  lsmps3d::detail::hip_check(
      hipMemcpy(v.fluid.pressure, ppe.pressure, nf * sizeof(real), hipMemcpyDeviceToDevice),
      "hipMemcpy");

  // gradient and Laplacian of the solved pressure
  const lsmps3d::MomentMatrixView pm = moments.pressure_type_a();
  if (!pm.is_ready || pm.matrix_size != 9 || pm.particle_count < nf || pm.inverse_matrices == nullptr)
    throw std::runtime_error("pressure operator diagnostics require a prepared pressure Type-A moment matrix");
  lsmps3d::DeviceFluidParticles op(nf);
  const auto ov = op.view();
  OperatorArgs oa{nf, v.fluid.x, v.fluid.y, v.fluid.z, v.walls.x, v.walls.y, v.walls.z,
                  v.walls.normal_x, v.walls.normal_y, v.walls.normal_z,
                  v.fluid_neighbors.offsets, v.fluid_neighbors.indices,
                  v.wall_neighbors.offsets, v.wall_neighbors.indices, pm.inverse_matrices,
                  ppe.pressure, pm.support_radius, pm.wall_weight_scale, cfg.density,
                  cfg.gravity.x, cfg.gravity.y, cfg.gravity.z, ov.x, ov.y, ov.z, ov.pressure};
  pressure_operator_kernel<<<static_cast<unsigned>((nf + 255) / 256), 256>>>(oa);
  lsmps3d::detail::hip_check(hipGetLastError(), "pressure_operator_kernel");
  lsmps3d::detail::hip_check(hipDeviceSynchronize(), "pressure_operator_kernel");

  // host copies
  const auto fluid_count = download(cd.offsets, nf);
  const auto wall_count = download(cd.indices, nf);
  const auto density = download(rd.x, nf);
  const auto ratio = download(rd.y, nf);
  const auto anisotropy = download(rd.z, nf);
  const auto open_ratio = download(rd.vx, nf);
  const auto air_anisotropy = download(rd.vy, nf);
  const auto snx = download(nd.x, nf);
  const auto sny = download(nd.y, nf);
  const auto snz = download(nd.z, nf);
  const auto label = download(v.fluid.surface_type, nf);
  const auto p_ppe = download(ppe.pressure, nf);
  const auto gxv = download(ov.x, nf);
  const auto gyv = download(ov.y, nf);
  const auto gzv = download(ov.z, nf);
  const auto lap = download(ov.pressure, nf);
  const auto rhs = download(ppe.rhs, nf);
  const auto divergence = download(ppe.divergence, nf);
  const auto wall_rhs = download(ppe.pressure_laplacian, nf);
  const auto rows = download(ppe.matrix.row_offsets, nf + 1);
  const size_t used = static_cast<size_t>(rows[nf]);
  const auto cols = download(ppe.matrix.col_indices, used);
  const auto vals = download(ppe.matrix.values, used);

  // matrix rows and the residual at the hydrostatic pressure
  std::vector<real> diag(nf, 0.0), row_sum(nf, 0.0), resid(nf), rp1(nf), rm1(nf), rp2(nf), rm2(nf);
  std::vector<int> row_nnz(nf);
  for (size_t r = 0; r < nf; ++r) {
    real ap = 0.0;
    row_nnz[r] = rows[r + 1] - rows[r];
    for (int32_t c = rows[r]; c < rows[r + 1]; ++c) {
      row_sum[r] += vals[c];
      if (cols[c] == static_cast<int32_t>(r)) diag[r] = vals[c];
      ap += vals[c] * p_true[static_cast<size_t>(cols[c])];
    }
    const real res = ap - rhs[r];
    resid[r] = res;
    rp1[r] = res + wall_rhs[r];
    rm1[r] = res - wall_rhs[r];
    rp2[r] = res + 2.0 * wall_rhs[r];
    rm2[r] = res - 2.0 * wall_rhs[r];
  }
  const auto p_err = minus(p_ppe, p_true);
  const auto p_abs = abs_of(p_err);
  const auto gx_err = abs_of(gxv);
  const auto gy_err = abs_of(gyv);
  std::vector<real> gz_err(nf), lap_err(nf);
  for (size_t i = 0; i < nf; ++i) {
    gz_err[i] = std::abs(gzv[i] - (-kRho * kG));
    lap_err[i] = std::abs(lap[i] - 0.0);
  }

  lsmps3d::HostVtkPointFields f;
  f.add_scalar("fluid_neighbor_count", fluid_count);
  f.add_scalar("wall_neighbor_count", wall_count);
  f.add_scalar("hydrostatic_pressure", p_true);
  f.add_scalar("ppe_pressure_gradient_x", gxv);
  f.add_scalar("ppe_pressure_gradient_y", gyv);
  f.add_scalar("ppe_pressure_gradient_z", gzv);
  f.add_scalar("ppe_pressure_laplacian", lap);
  f.add_scalar("ppe_pressure_gradient_x_abs_error", gx_err);
  f.add_scalar("ppe_pressure_gradient_y_abs_error", gy_err);
  f.add_scalar("ppe_pressure_gradient_z_abs_error", gz_err);
  f.add_scalar("ppe_pressure_laplacian_abs_error", lap_err);
  f.add_scalar("ppe_pressure", p_ppe);
  f.add_scalar("ppe_pressure_error", p_err);
  f.add_scalar("ppe_pressure_abs_error", p_abs);
  f.add_scalar("ppe_rhs", rhs);
  f.add_scalar("ppe_wall_neumann_rhs", wall_rhs);
  f.add_scalar("ppe_velocity_divergence", divergence);
  f.add_scalar("ppe_matrix_diagonal", diag);
  f.add_scalar("ppe_matrix_row_sum", row_sum);
  f.add_scalar("ppe_matrix_row_nnz", row_nnz);
  f.add_scalar("ppe_hydrostatic_residual", resid);
  f.add_scalar("ppe_hydrostatic_residual_plus_wall_rhs", rp1);
  f.add_scalar("ppe_hydrostatic_residual_minus_wall_rhs", rm1);
  f.add_scalar("ppe_hydrostatic_residual_plus_two_wall_rhs", rp2);
  f.add_scalar("ppe_hydrostatic_residual_minus_two_wall_rhs", rm2);
  f.add_scalar("surface_type", label);
  lsmps3d::SimulationConfig out = cfg;
  out.output_directory = out_dir;
  out.vtk_file_prefix = "hydrostatic_surface";
  out.vtk_write_point_fields = true;
  const lsmps3d::LegacyVtkWriter writer(out);
  writer.write(0, {box.fx, box.fy, box.fz}, f);

  const auto csv = out_dir / "hydrostatic_surface_debug.csv";
  {
    std::ofstream o(csv);
    if (!o) throw std::runtime_error("Failed to open CSV output: " + csv.string());
    o << std::setprecision(9);
    o << "particle_id,x,y,z,fluid_neighbor_count,wall_neighbor_count,number_density,"
         "number_density_ratio,anisotropy,air_open_ratio,air_anisotropy,surface_normal_x,"
         "surface_normal_y,surface_normal_z,surface_type,surface_type_name\n";
    for (size_t i = 0; i < nf; ++i)
      o << i << ',' << box.fx[i] << ',' << box.fy[i] << ',' << box.fz[i] << ',' << fluid_count[i]
        << ',' << wall_count[i] << ',' << density[i] << ',' << ratio[i] << ',' << anisotropy[i]
        << ',' << open_ratio[i] << ',' << air_anisotropy[i] << ',' << snx[i] << ',' << sny[i]
        << ',' << snz[i] << ',' << label[i] << ',' << label_name(label[i]) << '\n';
  }

  std::array<size_t, 4> by_label{};
  size_t top = 0, top_surface = 0, below_surface = 0;
  for (size_t i = 0; i < nf; ++i) {
    if (label[i] >= 0 && label[i] < 4) ++by_label[static_cast<size_t>(label[i])];
    const bool in_top = box.fz[i] >= kLevel - 0.75 * kSpacing;
    if (in_top) ++top;
    if (label[i] == 2) (in_top ? top_surface : below_surface) += 1;
  }
  const real n0 = lsmps3d::compute_uniform_reference_number_density(kSpacing, kRe);
  std::cout << "Hydrostatic half-full box diagnostics\n"
            << "  fluid particles: " << nf << '\n'
            << "  wall particles: " << nw << '\n'
            << "  spacing: " << kSpacing << " m\n"
            << "  support radius: " << kRe << " m\n"
            << "  reference number density n0: " << n0 << '\n'
            << "  Inner: " << by_label[0] << '\n'
            << "  NearSurface: " << by_label[1] << '\n'
            << "  Surface: " << by_label[2] << '\n'
            << "  Splash: " << by_label[3] << '\n'
            << "  Top-layer fluid particles: " << top << '\n'
            << "  Top-layer Surface particles: " << top_surface << '\n'
            << "  Below-top Surface particles: " << below_surface << '\n'
            << "  Hydrostatic pressure max: " << max_of(p_true) << " Pa\n"
            << "  PPE pressure gradient z max abs error: " << max_of(gz_err) << " Pa/m\n"
            << "  PPE pressure laplacian max abs error: " << max_of(lap_err) << " Pa/m^2\n"
            << "  PPE A*p_true-b max abs: " << max_of(abs_of(resid)) << '\n'
            << "  PPE residual+wall_rhs max abs: " << max_of(abs_of(rp1)) << '\n'
            << "  PPE residual-wall_rhs max abs: " << max_of(abs_of(rm1)) << '\n'
            << "  PPE residual+2*wall_rhs max abs: " << max_of(abs_of(rp2)) << '\n'
            << "  PPE residual-2*wall_rhs max abs: " << max_of(abs_of(rm2)) << '\n'
            << "  PPE pressure solve: library IDR(s) ("
            << solver.last_report().iterations << " iterations, relative residual "
            << solver.last_report().relative_residual << ")\n"
            << "  PPE pressure max abs error: " << max_of(p_abs) << " Pa\n"
            << "  VTK: " << writer.make_path(0) << '\n'
            << "  CSV: " << csv << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc > 1 ? std::filesystem::path(argv[1])
                        : std::filesystem::path("output/hydrostatic_surface_diagnostics"));
  } catch (const std::exception& e) {
    std::cerr << "hydrostatic_surface_diagnostics failed: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
