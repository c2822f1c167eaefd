// extern "C" surface of include/lsmps3d_amd.h.  Every entry converts C++
// exceptions into status codes and a thread-local message; nothing exits the
// process (the reference called std::exit from library code,
// include/lsmps3d/core/cuda_check.cuh:11-20).
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>

#include "objects.hpp"
#include "tune.hpp"

using namespace lsmps;

namespace {

thread_local std::string g_last_error;

int handle(const char* where) {
  try {
    throw;
  } catch (const Error& e) {
    g_last_error = std::string(where) + ": " + e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = std::string(where) + ": host allocation failed";
    return LSMPS_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = std::string(where) + ": " + e.what();
    return LSMPS_ERR_DEVICE;
  } catch (...) {
    g_last_error = std::string(where) + ": unknown error";
    return LSMPS_ERR_DEVICE;
  }
}

// Every entry starts by clearing the thread's last HIP error, so a launch check
// inside it reports this call's work and not an error some other HIP user of
// the thread (the caller, torch) left behind.
#define API_BEGIN       \
  (void)hipGetLastError(); \
  try {
#define API_END(name)      \
  }                        \
  catch (...) {            \
    return handle(name);   \
  }                        \
  return LSMPS_OK;

Context& C(lsmps_context* ctx) {
  require(ctx != nullptr, "null context");
  ctx->impl.activate();
  return ctx->impl;
}

template <typename T>
const T& deref(const T* p, const char* what) {
  if (!p) raise(LSMPS_ERR_INVALID_ARGUMENT, std::string("null ") + what);
  return *p;
}

}  // namespace

namespace {
lsmps_context* retain_context(lsmps_context* ctx) {
  ctx->refs.fetch_add(1);
  return ctx;
}

void release_context(lsmps_context* ctx) {
  if (!ctx || ctx->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(ctx->impl.device);
  (void)hipStreamSynchronize(ctx->impl.stream);
  if (ctx->impl.timing_ready)
    for (auto& e : ctx->impl.timing_events) (void)hipEventDestroy(e);
  if (ctx->impl.host_i32) (void)hipHostFree(ctx->impl.host_i32);
  if (ctx->impl.host_f64) (void)hipHostFree(ctx->impl.host_f64);
  for (auto& e : ctx->impl.tune.ev)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
}
}  // namespace

extern "C" {

const char* lsmps_last_error(void) { return g_last_error.c_str(); }
const char* lsmps_version(void) { return "lsmps3d_amd 0.1.0 (gfx950)"; }

int lsmps_context_create(int device, int precision, uint32_t flags, lsmps_context** out) {
  API_BEGIN
  require(out != nullptr, "null output pointer");
  *out = nullptr;
  require(precision == LSMPS_PRECISION_FP64 || precision == LSMPS_PRECISION_MIXED,
          "unknown precision mode");
  int count = 0;
  LSMPS_HIP(hipGetDeviceCount(&count));
  require(device >= 0 && device < count, "device index out of range", LSMPS_ERR_DEVICE);
  auto* ctx = new lsmps_context();
  ctx->impl.device = device;
  ctx->impl.precision = precision;
  ctx->impl.flags = flags;
  try {
    LSMPS_HIP(hipSetDevice(device));
    LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->impl.host_i32), 64));
    LSMPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->impl.host_f64), 64));
    noalloc_prepare_device(device);
    tune_prepare(ctx->impl);
  } catch (...) {
    delete ctx;
    throw;
  }
  *out = ctx;
  API_END("lsmps_context_create")
}

int lsmps_context_destroy(lsmps_context* ctx) {
  API_BEGIN
  if (!ctx) return LSMPS_OK;
  release_context(ctx);
  API_END("lsmps_context_destroy")
}

int lsmps_context_set_stream(lsmps_context* ctx, void* hip_stream) {
  API_BEGIN
  C(ctx).stream = static_cast<hipStream_t>(hip_stream);
  API_END("lsmps_context_set_stream")
}

int lsmps_context_lanes(lsmps_context* ctx, int32_t kernel, int32_t* lanes,
                        int64_t* tuned_keys) {
  API_BEGIN
  Context& c = C(ctx);
  require(kernel >= 0 && kernel < kTunedKernels, "unknown tuned kernel");
  if (lanes) *lanes = c.tune.last[kernel];
  if (tuned_keys) *tuned_keys = c.tune.tuned_keys;
  API_END("lsmps_context_lanes")
}

int lsmps_context_force_lanes(lsmps_context* ctx, int32_t kernel, int32_t lanes) {
  API_BEGIN
  Context& c = C(ctx);
  require(kernel >= 0 && kernel < kTunedKernels, "unknown tuned kernel");
  require(lanes >= 0 && lanes <= 16 && (lanes & (lanes - 1)) == 0,
          "lanes per particle must be 0 or a power of two up to 16");
  c.tune.forced[kernel] = lanes;
  API_END("lsmps_context_force_lanes")
}

int lsmps_dense_inverse(lsmps_context* ctx, int32_t n, const double* matrix, double* inverse) {
  API_BEGIN
  dense_inverse(C(ctx), n, matrix, inverse);
  API_END("lsmps_dense_inverse")
}

int lsmps_ppe_spmv(lsmps_context* ctx, const lsmps_csr* matrix, const double* x, double* y,
                   int32_t value_precision) {
  API_BEGIN
  require(matrix != nullptr, "null matrix");
  require(value_precision == LSMPS_PRECISION_FP64 || value_precision == LSMPS_PRECISION_MIXED,
          "value precision must be LSMPS_PRECISION_FP64 or LSMPS_PRECISION_MIXED");
  sell_spmv(C(ctx), *matrix, x, y, value_precision == LSMPS_PRECISION_MIXED);
  API_END("lsmps_ppe_spmv")
}

int lsmps_context_synchronize(lsmps_context* ctx) {
  API_BEGIN
  C(ctx).sync();
  API_END("lsmps_context_synchronize")
}

int lsmps_context_precision(const lsmps_context* ctx) {
  return ctx ? ctx->impl.precision : -1;
}

int lsmps_context_kernel_stats(lsmps_context* ctx, int64_t* launches, double* total_ms,
                               double* total_bytes, int reset) {
  API_BEGIN
  Context& c = C(ctx);
  if (launches) *launches = c.spmv_launches;
  if (total_ms) *total_ms = c.spmv_ms;
  if (total_bytes) *total_bytes = c.spmv_bytes;
  if (reset) {
    c.spmv_launches = 0;
    c.spmv_ms = 0.0;
    c.spmv_bytes = 0.0;
  }
  API_END("lsmps_context_kernel_stats")
}

int lsmps_default_config(lsmps_config* out) {
  API_BEGIN
  require(out != nullptr, "null config");
  lsmps_config c{};
  c.particle_spacing = 0.01;
  c.support_radius = 0.031;
  c.near_surface_radius = 0.015;
  c.cell_size = c.support_radius;
  c.cell_dims[0] = c.cell_dims[1] = c.cell_dims[2] = 1;
  c.time_step = 1.0e-4;
  c.min_time_step = 1.0e-6;
  c.max_time_step = 1.0e-3;
  c.time_step_growth_factor = 1.05;
  c.final_time = 1.0;
  c.output_interval = 0.01;
  c.density = 1000.0;
  c.kinematic_viscosity = 1.0e-6;
  c.cfl = 0.2;
  c.gravity[2] = -9.81;
  c.splash_neighbor_threshold = 4;
  c.include_wall_neighbors = 1;
  c.number_density_ratio_threshold = 0.85;
  c.air_open_ratio_threshold = 0.33;
  c.air_anisotropy_threshold = 0.033;
  c.wall_normal_independence_threshold = 0.25;
  c.virtual_light_cone_cosine = 0.8660254037844386;
  c.lsmps_regularization = 1.0e-8;
  c.lsmps_wall_weight_scale = 1.0;
  c.ps_displacement_scale = 0.05;
  c.ps_min_distance_ratio = 0.85;
  c.ps_max_displacement_ratio = 0.2;
  c.wall_clearance_ratio = 0.25;
  c.velocity_smoothing_strength = 0.1;
  c.weight_kernel = LSMPS_WEIGHT_LINEAR;
  *out = c;
  API_END("lsmps_default_config")
}

int lsmps_derive_cell_grid(const double* fx, const double* fy, const double* fz,
                           size_t nf, const double* wx, const double* wy, const double* wz,
                           size_t nw, lsmps_config* config) {
  API_BEGIN
  require(config != nullptr, "null config");
  bool has = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  auto include = [&](const double* x, const double* y, const double* z, size_t n) {
    for (size_t i = 0; i < n; ++i) {
      const double p[3] = {x[i], y[i], z[i]};
      for (int k = 0; k < 3; ++k)
        require(std::isfinite(p[k]), "Particle coordinates must be finite");
      if (!has) {
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = p[k];
        has = true;
        continue;
      }
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], p[k]);
        hi[k] = std::max(hi[k], p[k]);
      }
    }
  };
  if (nf) include(fx, fy, fz, nf);
  if (nw) include(wx, wy, wz, nw);
  require(has, "Cannot derive cell grid from empty particle input");
  const double pad = config->support_radius;
  for (int k = 0; k < 3; ++k) config->cell_origin[k] = lo[k] - pad;
  for (int k = 0; k < 3; ++k) {
    const double span = (hi[k] + pad) - config->cell_origin[k];
    const long long dim = static_cast<long long>(std::ceil(span / config->cell_size)) + 1;
    require(dim > 0 && dim <= std::numeric_limits<int32_t>::max(),
            "Derived cell-grid dimension is outside supported index range");
    config->cell_dims[k] = static_cast<int32_t>(dim);
  }
  API_END("lsmps_derive_cell_grid")
}

// ---- neighbor ---------------------------------------------------------------
int lsmps_build_cell_list_fluid(lsmps_context* ctx, const lsmps_fluid_soa* p,
                                const lsmps_cell_grid* grid, const lsmps_cell_list* cells) {
  API_BEGIN
  Context& c = C(ctx);
  const auto& f = deref(p, "particles");
  build_cell_list(c, f.count, f.x, f.y, f.z, deref(grid, "grid"), deref(cells, "cells"));
  c.finish();
  API_END("lsmps_build_cell_list_fluid")
}

int lsmps_build_cell_list_wall(lsmps_context* ctx, const lsmps_wall_soa* p,
                               const lsmps_cell_grid* grid, const lsmps_cell_list* cells) {
  API_BEGIN
  Context& c = C(ctx);
  const auto& w = deref(p, "particles");
  build_cell_list(c, w.count, w.x, w.y, w.z, deref(grid, "grid"), deref(cells, "cells"));
  c.finish();
  API_END("lsmps_build_cell_list_wall")
}

int lsmps_build_neighbor_lists(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                               const lsmps_wall_soa* walls, const lsmps_config* config,
                               const lsmps_cell_list* fc, const lsmps_cell_list* wc,
                               const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn) {
  API_BEGIN
  Context& c = C(ctx);
  build_neighbor_lists(c, deref(fluid, "fluid"), deref(walls, "walls"), deref(config, "config"),
                       deref(fc, "fluid cells"), deref(wc, "wall cells"),
                       deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"), nullptr,
                       nullptr);
  c.finish();
  API_END("lsmps_build_neighbor_lists")
}

int lsmps_build_neighbor_lists_noalloc(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_wall_soa* walls, const lsmps_config* config,
                                       const lsmps_cell_list* fc, const lsmps_cell_list* wc,
                                       const lsmps_neighbor_list* fn,
                                       const lsmps_neighbor_list* wn) {
  API_BEGIN
  Context& c = C(ctx);
  build_neighbor_lists_noalloc(c, deref(fluid, "fluid"), deref(walls, "walls"),
                               deref(config, "config"), deref(fc, "fluid cells"),
                               deref(wc, "wall cells"), deref(fn, "fluid neighbors"),
                               deref(wn, "wall neighbors"));
  c.finish();
  API_END("lsmps_build_neighbor_lists_noalloc")
}

// ---- surface ----------------------------------------------------------------
int lsmps_reference_number_density(double l0, double re, int32_t kernel, double* out) {
  API_BEGIN
  require(out != nullptr, "null output");
  *out = reference_number_density(l0, re, kernel);
  API_END("lsmps_reference_number_density")
}

int lsmps_classify_surface(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                           const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                           const lsmps_neighbor_list* wn, const lsmps_config* config,
                           const lsmps_surface_diagnostics* diagnostics) {
  API_BEGIN
  Context& c = C(ctx);
  classify_surface(c, deref(fluid, "fluid"), deref(walls, "walls"), deref(fn, "fluid neighbors"),
                   deref(wn, "wall neighbors"), deref(config, "config"), diagnostics);
  c.finish();
  API_END("lsmps_classify_surface")
}

int lsmps_compute_virtual_light_diagnostics(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                            const lsmps_wall_soa* walls,
                                            const lsmps_neighbor_list* fn,
                                            const lsmps_neighbor_list* wn,
                                            const lsmps_config* config,
                                            const lsmps_virtual_light_diagnostics* diagnostics) {
  API_BEGIN
  Context& c = C(ctx);
  virtual_light(c, deref(fluid, "fluid"), deref(walls, "walls"), deref(fn, "fluid neighbors"),
                deref(wn, "wall neighbors"), deref(config, "config"),
                deref(diagnostics, "virtual light diagnostics"));
  c.finish();
  API_END("lsmps_compute_virtual_light_diagnostics")
}

// ---- moments ----------------------------------------------------------------
int lsmps_moment_create(lsmps_context* ctx, size_t capacity, const lsmps_config* config,
                        lsmps_moment_matrix** out) {
  API_BEGIN
  require(out != nullptr, "null output pointer");
  Context& c = C(ctx);
  const auto& cfg = deref(config, "config");
  auto* m = new lsmps_moment_matrix();
  m->ctx = &c;
  m->config = cfg;
  m->resize(capacity);  // records the capacity; allocates on first prepare
  m->owner = retain_context(ctx);
  *out = m;
  API_END("lsmps_moment_create")
}

int lsmps_moment_destroy(lsmps_moment_matrix* m) {
  API_BEGIN
  if (m) {
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    lsmps_context* owner = m->owner;
    delete m;
    release_context(owner);
  }
  API_END("lsmps_moment_destroy")
}

int lsmps_moment_resize(lsmps_moment_matrix* m, size_t capacity) {
  API_BEGIN
  require(m != nullptr, "null moment matrix");
  m->ctx->activate();
  m->resize(capacity);
  API_END("lsmps_moment_resize")
}

int lsmps_moment_set_config(lsmps_moment_matrix* m, const lsmps_config* config) {
  API_BEGIN
  require(m != nullptr, "null moment matrix");
  // set_config invalidates the prepared generation (moment_matrix.cu:602-605)
  m->config = deref(config, "config");
  m->prepared = false;
  API_END("lsmps_moment_set_config")
}

size_t lsmps_moment_bytes(const lsmps_moment_matrix* m) { return m ? m->bytes() : 0; }

int lsmps_moment_prepare(lsmps_moment_matrix* m, const lsmps_fluid_soa* fluid,
                         const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                         const lsmps_neighbor_list* wn, uint64_t generation) {
  API_BEGIN
  require(m != nullptr, "null moment matrix");
  m->ctx->activate();
  m->prepare(deref(fluid, "fluid"), deref(walls, "walls"), deref(fn, "fluid neighbors"),
             deref(wn, "wall neighbors"), generation);
  m->ctx->finish();
  API_END("lsmps_moment_prepare")
}

int lsmps_moment_get_view(lsmps_moment_matrix* m, int32_t kind, lsmps_moment_view* out) {
  API_BEGIN
  require(m != nullptr && out != nullptr, "null argument");
  require(kind >= 0 && kind <= 3, "Unknown Moment matrix kind");
  require(m->prepared, "Moment operators must prepare matrices before computing operators",
          LSMPS_ERR_NOT_PREPARED);
  m->ctx->activate();
  m->materialize(kind);
  LSMPS_HIP(hipMemcpyAsync(m->inversion_count.data(), m->host_counts, 4 * sizeof(int32_t),
                           hipMemcpyHostToDevice, m->ctx->stream));
  m->ctx->sync();
  lsmps_moment_view v{};
  v.particle_count = m->count;
  v.matrix_size = moment_size(kind);
  v.basis_kind = kind == LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_B ? LSMPS_BASIS_TYPE_B : LSMPS_BASIS_TYPE_A;
  v.kind = kind;
  v.is_ready = 1;
  v.support_radius = m->config.support_radius;
  v.wall_weight_scale = m->config.lsmps_wall_weight_scale;
  v.geometry_generation = m->generation;
  v.inverse_matrices = m->full[kind].inverse.data();
  v.moment_trace = m->full[kind].trace.data();
  v.regularization_added = m->full[kind].regularization.data();
  v.info = m->full[kind].info.data();
  v.inversion_count = m->inversion_count.data() + kind;
  *out = v;
  API_END("lsmps_moment_get_view")
}

// ---- provision --------------------------------------------------------------
int lsmps_provision_compute(lsmps_context* ctx, const lsmps_config* config,
                            const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                            const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                            lsmps_moment_matrix* m, uint64_t generation,
                            const lsmps_fluid_soa* tmp, const lsmps_wall_soa* tmpw,
                            const lsmps_vector3* lap) {
  API_BEGIN
  Context& c = C(ctx);
  require(m != nullptr, "null moment matrix");
  const auto& cfg = deref(config, "config");
  const auto& f = deref(fluid, "fluid");
  const auto& w = deref(walls, "walls");
  // compute_temporary_velocity prepares the moments with its own config (explicit_update.cu:384-387)
  m->config = cfg;
  m->prepare(f, w, deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"), generation);
  provision_fluid(c, cfg, f, w, *fn, *wn, m->velocity_ops(), deref(tmp, "temporary velocity"), lap);
  provision_walls(c, cfg, w, deref(tmpw, "temporary wall velocity"));
  c.finish();
  API_END("lsmps_provision_compute")
}

int lsmps_provision_temporary_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_config* config, const lsmps_vector3* lap,
                                       const lsmps_fluid_soa* tmp) {
  API_BEGIN
  Context& c = C(ctx);
  provision_combine(c, deref(config, "config"), deref(fluid, "fluid"), deref(lap, "laplacian"),
                    deref(tmp, "temporary velocity"));
  c.finish();
  API_END("lsmps_provision_temporary_velocity")
}

int lsmps_provision_temporary_wall_velocity(lsmps_context* ctx, const lsmps_wall_soa* walls,
                                            const lsmps_config* config,
                                            const lsmps_wall_soa* tmpw) {
  API_BEGIN
  Context& c = C(ctx);
  provision_walls(c, deref(config, "config"), deref(walls, "walls"),
                  deref(tmpw, "temporary wall velocity"));
  c.finish();
  API_END("lsmps_provision_temporary_wall_velocity")
}

// ---- PPE --------------------------------------------------------------------
int lsmps_ppe_assemble(lsmps_context* ctx, const lsmps_config* config,
                       const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                       const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                       const lsmps_fluid_soa* tmp, const lsmps_wall_soa* tmpw,
                       lsmps_moment_matrix* m, uint64_t generation,
                       const lsmps_ppe_workspace* ws) {
  API_BEGIN
  Context& c = C(ctx);
  require(m != nullptr, "null moment matrix");
  const auto& cfg = deref(config, "config");
  const auto& f = deref(fluid, "fluid");
  const auto& w = deref(walls, "walls");
  m->config = cfg;
  m->prepare(f, w, deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"), generation);
  ppe_assemble(c, cfg, f, w, *fn, *wn, deref(tmp, "temporary velocity"),
               deref(tmpw, "temporary wall velocity"), m->velocity_ops(), m->pressure_ops(),
               deref(ws, "workspace"));
  c.finish();
  API_END("lsmps_ppe_assemble")
}

namespace {
// validate_ppe_moment_matrix (ppe_matrix.cu:108-122) and
// validate_pressure_moment_matrix (pressure_correction.cu:96-111).
void require_type_a_view(const lsmps_moment_view& v, int kind, size_t count,
                         const std::string& message, const std::string& missing) {
  require(v.is_ready != 0 && v.kind == kind && v.basis_kind == LSMPS_BASIS_TYPE_A &&
              v.matrix_size == 9 && v.particle_count >= count,
          message.c_str(), LSMPS_ERR_NOT_PREPARED);
  require(count == 0 || v.inverse_matrices != nullptr, missing.c_str(), LSMPS_ERR_NOT_PREPARED);
}
}  // namespace

int lsmps_ppe_assemble_views(lsmps_context* ctx, const lsmps_config* config,
                             const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                             const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                             const lsmps_fluid_soa* tmp, const lsmps_wall_soa* tmpw,
                             const lsmps_moment_view* velocity_moment,
                             const lsmps_moment_view* pressure_moment,
                             const lsmps_ppe_workspace* ws) {
  API_BEGIN
  Context& c = C(ctx);
  const auto& cfg = deref(config, "config");
  const auto& f = deref(fluid, "fluid");
  const auto& vm = deref(velocity_moment, "velocity moment view");
  const auto& pm = deref(pressure_moment, "pressure moment view");
  require_type_a_view(vm, LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A, f.count,
                      "PPE requires a prepared velocity Type-A moment matrix",
                      "PPE moment matrix inverse buffer is missing for velocity");
  require_type_a_view(pm, LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A, f.count,
                      "PPE requires a prepared pressure Type-A moment matrix",
                      "PPE moment matrix inverse buffer is missing for pressure");
  const size_t n = f.count;
  c.view_ops.reserve(std::max<size_t>(2 * kOpRows * n, 1));
  double* vel = c.view_ops.data();
  double* pre = vel + kOpRows * n;
  operator_rows_from_view(c, vm, n, vel);
  operator_rows_from_view(c, pm, n, pre);
  ppe_assemble(c, cfg, f, deref(walls, "walls"), deref(fn, "fluid neighbors"),
               deref(wn, "wall neighbors"), deref(tmp, "temporary velocity"),
               deref(tmpw, "temporary wall velocity"), vel, pre, deref(ws, "workspace"));
  c.finish();
  API_END("lsmps_ppe_assemble_views")
}

int lsmps_default_solver_options(lsmps_solver_options* out) {
  API_BEGIN
  require(out != nullptr, "null options");
  lsmps_solver_options o{};
  o.method = LSMPS_SOLVER_IDRS;
  o.max_iterations = 50000;
  o.tolerance = 1.0e-10;
  o.inner_tolerance = 1.0e-3;
  o.check_interval = 16;
  o.clamp_negative = 1;
  o.shadow_dimension = 0;  // 14 with the coarse space, 16 without
  o.preconditioner = LSMPS_PRECOND_SCHWARZ_COARSE;
  o.coarse_cells = 4;
  *out = o;
  API_END("lsmps_default_solver_options")
}


int lsmps_ppe_solve(lsmps_context* ctx, const lsmps_csr* matrix, const double* rhs,
                    double* pressure, const lsmps_solver_options* options,
                    lsmps_solve_report* report) {
  API_BEGIN
  Context& c = C(ctx);
  lsmps_solver_options opt{};
  if (options) {
    opt = *options;
  } else {
    lsmps_default_solver_options(&opt);
  }
  // The free solve keeps its workspace in the context (reused across calls).
  if (!ctx->solver) ctx->solver = std::make_unique<PpeSolver>();
  ppe_solve(c, *ctx->solver, deref(matrix, "matrix"), rhs, pressure, opt, report);
  API_END("lsmps_ppe_solve")
}

// ---- correction -------------------------------------------------------------
int lsmps_correction_apply(lsmps_context* ctx, const lsmps_config* config,
                           const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                           const lsmps_neighbor_list* fn, const lsmps_neighbor_list* wn,
                           const lsmps_fluid_soa* tmp, double* pressure, lsmps_moment_matrix* m,
                           uint64_t generation, const lsmps_vector3* grad,
                           const lsmps_vector3* ps, const lsmps_vector3* smoothed) {
  API_BEGIN
  Context& c = C(ctx);
  require(m != nullptr, "null moment matrix");
  const auto& cfg = deref(config, "config");
  const auto& f = deref(fluid, "fluid");
  const auto& w = deref(walls, "walls");
  const auto& fnl = deref(fn, "fluid neighbors");
  const auto& wnl = deref(wn, "wall neighbors");
  clamp_pressure(c, f, pressure);
  m->config = cfg;
  m->prepare(f, w, fnl, wnl, generation);
  pressure_gradient(c, cfg, f, w, fnl, wnl, pressure, m->pressure_ops(), deref(grad, "gradient"));
  particle_shifting(c, cfg, f, w, fnl, wnl, deref(ps, "shifting"));
  update_velocity_position(c, cfg, f, deref(tmp, "temporary velocity"), *grad, *ps);
  anti_penetration(c, cfg, f, w, wnl);
  smooth_velocity(c, cfg, f, fnl, deref(smoothed, "smoothed velocity"));
  c.finish();
  API_END("lsmps_correction_apply")
}

int lsmps_clamp_particle_pressure(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                  double* pressure) {
  API_BEGIN
  Context& c = C(ctx);
  clamp_pressure(c, deref(fluid, "fluid"), pressure);
  c.finish();
  API_END("lsmps_clamp_particle_pressure")
}

int lsmps_compute_pressure_gradient(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                    const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                                    const lsmps_neighbor_list* wn, const double* pressure,
                                    const lsmps_config* config, lsmps_moment_matrix* m,
                                    const lsmps_vector3* grad) {
  API_BEGIN
  Context& c = C(ctx);
  require(m != nullptr, "null moment matrix");
  require(m->prepared, "Pressure correction requires a prepared pressure Type-A moment matrix",
          LSMPS_ERR_NOT_PREPARED);
  require(m->count >= deref(fluid, "fluid").count,
          "Pressure correction requires a prepared pressure Type-A moment matrix");
  pressure_gradient(c, deref(config, "config"), *fluid, deref(walls, "walls"),
                    deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"), pressure,
                    m->pressure_ops(), deref(grad, "gradient"));
  c.finish();
  API_END("lsmps_compute_pressure_gradient")
}

int lsmps_compute_pressure_gradient_view(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                         const lsmps_wall_soa* walls,
                                         const lsmps_neighbor_list* fn,
                                         const lsmps_neighbor_list* wn, const double* pressure,
                                         const lsmps_config* config,
                                         const lsmps_moment_view* pressure_moment,
                                         const lsmps_vector3* grad) {
  API_BEGIN
  Context& c = C(ctx);
  const auto& f = deref(fluid, "fluid");
  const auto& pm = deref(pressure_moment, "pressure moment view");
  require_type_a_view(pm, LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A, f.count,
                      "Pressure correction requires a prepared pressure Type-A moment matrix",
                      "Pressure correction pressure moment inverse buffer is missing");
  const size_t n = f.count;
  c.view_ops.reserve(std::max<size_t>(kOpRows * n, 1));
  operator_rows_from_view(c, pm, n, c.view_ops.data());
  pressure_gradient(c, deref(config, "config"), f, deref(walls, "walls"),
                    deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"), pressure,
                    c.view_ops.data(), deref(grad, "gradient"));
  c.finish();
  API_END("lsmps_compute_pressure_gradient_view")
}

int lsmps_compute_particle_shifting(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                    const lsmps_wall_soa* walls, const lsmps_neighbor_list* fn,
                                    const lsmps_neighbor_list* wn, const lsmps_config* config,
                                    const lsmps_vector3* ps) {
  API_BEGIN
  Context& c = C(ctx);
  particle_shifting(c, deref(config, "config"), deref(fluid, "fluid"), deref(walls, "walls"),
                    deref(fn, "fluid neighbors"), deref(wn, "wall neighbors"),
                    deref(ps, "shifting"));
  c.finish();
  API_END("lsmps_compute_particle_shifting")
}

int lsmps_update_velocity_and_position(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_fluid_soa* tmp, const lsmps_config* config,
                                       const lsmps_vector3* grad, const lsmps_vector3* ps) {
  API_BEGIN
  Context& c = C(ctx);
  update_velocity_position(c, deref(config, "config"), deref(fluid, "fluid"),
                           deref(tmp, "temporary velocity"), deref(grad, "gradient"),
                           deref(ps, "shifting"));
  c.finish();
  API_END("lsmps_update_velocity_and_position")
}

int lsmps_apply_anti_penetration(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                 const lsmps_wall_soa* walls, const lsmps_neighbor_list* wn,
                                 const lsmps_config* config) {
  API_BEGIN
  Context& c = C(ctx);
  anti_penetration(c, deref(config, "config"), deref(fluid, "fluid"), deref(walls, "walls"),
                   deref(wn, "wall neighbors"));
  c.finish();
  API_END("lsmps_apply_anti_penetration")
}

int lsmps_smooth_particle_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                   const lsmps_neighbor_list* fn, const lsmps_config* config,
                                   const lsmps_vector3* smoothed) {
  API_BEGIN
  Context& c = C(ctx);
  smooth_velocity(c, deref(config, "config"), deref(fluid, "fluid"), deref(fn, "fluid neighbors"),
                  deref(smoothed, "smoothed velocity"));
  c.finish();
  API_END("lsmps_smooth_particle_velocity")
}

// ---- stepper ----------------------------------------------------------------
int lsmps_stepper_create(lsmps_context* ctx, size_t fluid_capacity, size_t wall_capacity,
                         size_t max_neighbors, const lsmps_config* config, lsmps_stepper** out) {
  API_BEGIN
  require(out != nullptr, "null output pointer");
  *out = nullptr;
  Context& c = C(ctx);
  require(max_neighbors > 0, "max_neighbors_per_particle must be positive");
  const auto& cfg = deref(config, "config");
  auto* s = new lsmps_stepper();
  try {
    s->ctx = &c;
    s->config = cfg;
    s->max_neighbors = max_neighbors;
    s->moments.ctx = &c;
    lsmps_default_solver_options(&s->solver_options);
    const size_t cells = static_cast<size_t>(cfg.cell_dims[0]) * cfg.cell_dims[1] * cfg.cell_dims[2];
    s->allocate(fluid_capacity, wall_capacity, cells);
  } catch (...) {
    delete s;
    throw;
  }
  s->owner = retain_context(ctx);
  *out = s;
  API_END("lsmps_stepper_create")
}

int lsmps_stepper_destroy(lsmps_stepper* s) {
  API_BEGIN
  if (s) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    lsmps_context* owner = s->owner;
    delete s;
    release_context(owner);
  }
  API_END("lsmps_stepper_destroy")
}

int lsmps_stepper_set_config(lsmps_stepper* s, const lsmps_config* config) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  s->config = deref(config, "config");
  s->moments.prepared = false;
  API_END("lsmps_stepper_set_config")
}

int lsmps_stepper_set_solver(lsmps_stepper* s, const lsmps_solver_options* options) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  s->solver_options = deref(options, "options");
  API_END("lsmps_stepper_set_solver")
}

int lsmps_stepper_set_segments(lsmps_stepper* s, int32_t segment_count,
                               const int64_t* fluid_offsets, const int64_t* wall_offsets,
                               const lsmps_cell_grid* grids) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  s->ctx->activate();
  s->set_segments(segment_count, fluid_offsets, wall_offsets, grids);
  API_END("lsmps_stepper_set_segments")
}

size_t lsmps_stepper_bytes(const lsmps_stepper* s) { return s ? s->bytes() : 0; }

int lsmps_default_config_files(lsmps_config_files* out) {
  API_BEGIN
  require(out != nullptr, "null output");
  default_config_files(*out);
  API_END("lsmps_default_config_files")
}

int lsmps_config_validate(const lsmps_config* config, const lsmps_config_files* files) {
  API_BEGIN
  validate_config(deref(config, "config"), files);
  API_END("lsmps_config_validate")
}

int lsmps_config_load(const char* path, lsmps_config* config, lsmps_config_files* files) {
  API_BEGIN
  require(path != nullptr && config != nullptr, "null argument");
  lsmps_config c{};
  lsmps_config_files f{};
  load_config(path, c, f);
  *config = c;
  if (files) *files = f;
  API_END("lsmps_config_load")
}

int lsmps_config_save(const char* path, const lsmps_config* config,
                      const lsmps_config_files* files) {
  API_BEGIN
  require(path != nullptr, "null path");
  lsmps_config_files f{};
  if (files) f = *files; else default_config_files(f);
  save_config(path, deref(config, "config"), f);
  API_END("lsmps_config_save")
}

int lsmps_read_fluid_csv(const char* path, const lsmps_fluid_soa* host, size_t* count) {
  API_BEGIN
  require(path != nullptr, "null path");
  const bool fill = host && host->x;
  double* out[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (fill) {
    out[0] = host->x; out[1] = host->y; out[2] = host->z;
    out[3] = host->vx; out[4] = host->vy; out[5] = host->vz;
  }
  const size_t rows = read_particle_csv(path, 6, fill ? out : nullptr, fill ? host->count : 0);
  if (fill) {
    for (size_t i = 0; i < rows; ++i) {
      if (host->pressure) host->pressure[i] = 0.0;
      if (host->surface_type) host->surface_type[i] = LSMPS_SURFACE_INNER;
    }
  }
  if (count) *count = rows;
  API_END("lsmps_read_fluid_csv")
}

int lsmps_read_wall_csv(const char* path, const lsmps_wall_soa* host, size_t* count) {
  API_BEGIN
  require(path != nullptr, "null path");
  const bool fill = host && host->x;
  double* out[9] = {};
  if (fill) {
    out[0] = host->x; out[1] = host->y; out[2] = host->z;
    out[3] = host->vx; out[4] = host->vy; out[5] = host->vz;
    out[6] = host->normal_x; out[7] = host->normal_y; out[8] = host->normal_z;
  }
  const size_t rows = read_particle_csv(path, 9, fill ? out : nullptr, fill ? host->count : 0);
  if (count) *count = rows;
  API_END("lsmps_read_wall_csv")
}

int lsmps_vtk_path(const char* output_directory, const char* prefix, size_t step, char* out,
                   size_t capacity) {
  API_BEGIN
  require(output_directory && prefix && out, "null argument");
  const std::string p = vtk_path(output_directory, prefix, step);
  require(p.size() + 1 <= capacity, "output buffer too small", LSMPS_ERR_CAPACITY);
  std::memcpy(out, p.c_str(), p.size() + 1);
  API_END("lsmps_vtk_path")
}

int lsmps_write_vtk(const char* output_directory, const char* prefix, int32_t write_point_fields,
                    size_t step, size_t count, const double* x, const double* y, const double* z,
                    const lsmps_vtk_field* fields, int32_t field_count) {
  API_BEGIN
  require(output_directory && prefix, "null argument");
  require(field_count == 0 || fields != nullptr, "null fields");
  write_vtk(output_directory, prefix, write_point_fields != 0, step, count, x, y, z, fields,
            field_count);
  API_END("lsmps_write_vtk")
}

int lsmps_time_limits_from_config(const lsmps_config* config, lsmps_time_limits* out) {
  API_BEGIN
  require(out != nullptr, "null output");
  *out = time_limits_from_config(deref(config, "config"));
  API_END("lsmps_time_limits_from_config")
}

int lsmps_time_limits_validate(const lsmps_time_limits* limits) {
  API_BEGIN
  validate_time_limits(deref(limits, "limits"));
  API_END("lsmps_time_limits_validate")
}

int lsmps_time_manager_create(const lsmps_time_limits* limits, lsmps_time_manager** out) {
  API_BEGIN
  require(out != nullptr, "null output");
  auto clock = std::make_unique<lsmps_time_manager>();
  clock->reset(deref(limits, "limits"));
  *out = clock.release();
  API_END("lsmps_time_manager_create")
}

int lsmps_time_manager_destroy(lsmps_time_manager* clock) {
  delete clock;
  return LSMPS_OK;
}

int lsmps_time_manager_reset(lsmps_time_manager* clock, const lsmps_time_limits* limits) {
  API_BEGIN
  require(clock != nullptr, "null clock");
  clock->reset(deref(limits, "limits"));
  API_END("lsmps_time_manager_reset")
}

int lsmps_time_manager_state(const lsmps_time_manager* clock, double* current_time,
                             double* current_time_step, size_t* step_index,
                             size_t* output_index, int32_t* output_due, int32_t* finished) {
  API_BEGIN
  require(clock != nullptr, "null clock");
  if (current_time) *current_time = clock->current_time;
  if (current_time_step) *current_time_step = clock->current_time_step;
  if (step_index) *step_index = clock->step_index;
  if (output_index) *output_index = clock->output_index;
  if (output_due) *output_due = clock->output_due() ? 1 : 0;
  if (finished) *finished = clock->finished() ? 1 : 0;
  API_END("lsmps_time_manager_state")
}

int lsmps_time_manager_next_time_step(const lsmps_time_manager* clock, double max_velocity,
                                      double* time_step, double* cfl_time_step) {
  API_BEGIN
  require(clock != nullptr, "null clock");
  if (time_step) *time_step = clock->next_dt(max_velocity);
  if (cfl_time_step) *cfl_time_step = clock->cfl_dt(max_velocity);
  API_END("lsmps_time_manager_next_time_step")
}

int lsmps_time_manager_advance(lsmps_time_manager* clock, double max_velocity,
                               lsmps_time_status* status) {
  API_BEGIN
  require(clock != nullptr && status != nullptr, "null argument");
  *status = clock->advance(max_velocity);
  API_END("lsmps_time_manager_advance")
}

int lsmps_time_manager_mark_initial_output(lsmps_time_manager* clock, lsmps_time_status* status) {
  API_BEGIN
  require(clock != nullptr && status != nullptr, "null argument");
  *status = clock->mark_initial_output();
  API_END("lsmps_time_manager_mark_initial_output")
}

int lsmps_max_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid, double* out) {
  API_BEGIN
  require(out != nullptr, "null output");
  Context& c = C(ctx);
  *out = max_speed(c, deref(fluid, "fluid"), c.reduce_scratch);
  API_END("lsmps_max_velocity")
}

int lsmps_stepper_run(lsmps_stepper* s, const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                      lsmps_time_manager* clock, int64_t max_steps,
                      lsmps_output_callback callback, void* user, lsmps_time_status* last) {
  API_BEGIN
  require(s != nullptr && clock != nullptr, "null argument");
  s->ctx->activate();
  const int rc = run_steps(*s, deref(fluid, "fluid"), deref(walls, "walls"), *clock, max_steps,
                           callback, user, last);
  if (rc != LSMPS_OK) {
    g_last_error = "lsmps_stepper_run: a PPE solve did not reach the tolerance";
    return rc;
  }
  API_END("lsmps_stepper_run")
}

int lsmps_stepper_refresh_geometry(lsmps_stepper* s, const lsmps_fluid_soa* fluid,
                                   const lsmps_wall_soa* walls) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  s->ctx->activate();
  s->refresh_geometry(deref(fluid, "fluid"), deref(walls, "walls"));
  API_END("lsmps_stepper_refresh_geometry")
}

int lsmps_stepper_set_mode(lsmps_stepper* s, uint32_t mode) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  require((mode & ~static_cast<uint32_t>(LSMPS_STEP_RECORD | LSMPS_STEP_FROZEN_GEOMETRY)) == 0,
          "unknown stepper mode bits");
  s->mode = mode;
  if (!(mode & LSMPS_STEP_RECORD)) s->recorded = false;
  API_END("lsmps_stepper_set_mode")
}

int lsmps_adjoint_step(lsmps_stepper* s, const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                       const lsmps_vector3* grad_velocity_out,
                       const lsmps_vector3* grad_position_out, lsmps_adjoint_grads* grads) {
  API_BEGIN
  require(s != nullptr && grads != nullptr, "null argument");
  s->ctx->activate();
  int rc = LSMPS_OK;
  try {
    adjoint_step(*s->ctx, *s, deref(fluid, "fluid"), deref(walls, "walls"),
                 deref(grad_velocity_out, "grad_velocity_out"), grad_position_out, *grads);
  } catch (const Error& e) {
    if (e.code != LSMPS_ERR_NOT_CONVERGED) throw;
    g_last_error = std::string("lsmps_adjoint_step: ") + e.what();
    rc = e.code;
  }
  return rc;
  API_END("lsmps_adjoint_step")
}

int lsmps_stepper_step(lsmps_stepper* s, const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                       lsmps_solve_report* report) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  s->ctx->activate();
  int rc = LSMPS_OK;
  try {
    s->step(deref(fluid, "fluid"), deref(walls, "walls"));
  } catch (const Error& e) {
    if (e.code != LSMPS_ERR_NOT_CONVERGED) throw;
    g_last_error = std::string("lsmps_stepper_step: ") + e.what();
    rc = e.code;
  }
  if (report) *report = s->last_report;
  return rc;
  API_END("lsmps_stepper_step")
}

int lsmps_stepper_stage_times(const lsmps_stepper* s, double* ms, int count) {
  API_BEGIN
  require(s != nullptr && ms != nullptr, "null argument");
  for (int k = 0; k < count && k < 7; ++k) ms[k] = s->stage_ms[k];
  API_END("lsmps_stepper_stage_times")
}

int lsmps_stepper_neighbor_lists(lsmps_stepper* s, lsmps_neighbor_list* fn,
                                 lsmps_neighbor_list* wn) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  if (fn) *fn = s->fluid_neighbors();
  if (wn) *wn = s->wall_neighbors();
  API_END("lsmps_stepper_neighbor_lists")
}

int lsmps_stepper_ppe_workspace(lsmps_stepper* s, lsmps_ppe_workspace* out) {
  API_BEGIN
  require(s != nullptr && out != nullptr, "null argument");
  *out = s->ppe_workspace(s->moments.count);
  API_END("lsmps_stepper_ppe_workspace")
}

int lsmps_stepper_temporary_velocity(lsmps_stepper* s, lsmps_fluid_soa* f, lsmps_wall_soa* w) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  const size_t nf1 = std::max<size_t>(s->fluid_capacity, 1);
  const size_t nw1 = std::max<size_t>(s->wall_capacity, 1);
  double* tv = s->tmp_fluid.data();
  double* tw = s->tmp_wall.data();
  if (f) *f = lsmps_fluid_soa{s->moments.count, nullptr, nullptr, nullptr, tv, tv + nf1,
                              tv + 2 * nf1, nullptr, nullptr};
  if (w) *w = lsmps_wall_soa{s->wall_capacity, nullptr, nullptr, nullptr, tw, tw + nw1,
                             tw + 2 * nw1, nullptr, nullptr, nullptr};
  API_END("lsmps_stepper_temporary_velocity")
}

int lsmps_stepper_correction_vectors(lsmps_stepper* s, lsmps_vector3* grad, lsmps_vector3* ps) {
  API_BEGIN
  require(s != nullptr, "null stepper");
  const size_t nf1 = std::max<size_t>(s->fluid_capacity, 1);
  double* cv = s->corr.data();
  const size_t n = s->moments.count;
  if (grad) *grad = lsmps_vector3{n, cv, cv + nf1, cv + 2 * nf1};
  if (ps) *ps = lsmps_vector3{n, cv + 3 * nf1, cv + 4 * nf1, cv + 5 * nf1};
  API_END("lsmps_stepper_correction_vectors")
}

lsmps_moment_matrix* lsmps_stepper_moments(lsmps_stepper* s) { return s ? &s->moments : nullptr; }

}  // extern "C"
