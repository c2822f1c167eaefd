/*
 * lsmps3d_amd.h — C-ABI of the MI355X (gfx950) per-step LSMPS particle pipeline.
 *
 * The entry points replace the C++17 module API of lsmps-3D's hot path
 * (include/lsmps3d/{neighbor,surface,moment_matrix,provision,ppe,correction}).
 * Each declaration names the reference function it stands in for.
 *
 * Conventions
 *  - Every pointer inside a particle / neighbor / cell / PPE struct is a DEVICE
 *    pointer (hipMalloc'd or torch-owned HBM) unless a function says "host".
 *  - `real` of the reference is double here (LSMPS3D_USE_DOUBLE=ON, the
 *    reference default, CMakeLists.txt:10).  The precision mode of a context
 *    changes only internal arithmetic (operators, SpMV), never buffer types.
 *  - Every entry returns an int status (0 = LSMPS_OK) and never exits the
 *    process; lsmps_last_error() returns the message of the last failure on the
 *    calling thread.  The reference printed to stderr and called std::exit
 *    (include/lsmps3d/core/cuda_check.cuh:11-20).
 *  - Work is enqueued on the context's stream.  Unless the context was created
 *    with LSMPS_FLAG_ASYNC, each entry synchronises that stream before
 *    returning, matching the reference's synchronous semantics.
 *  - Counts are capacities, as in the reference: NeighborListView::neighbor_count
 *    is the allocated index capacity (src/core/runtime.cu:232), and the PPE
 *    matrix reports nnz = N + neighbor capacity (src/ppe/ppe_matrix.cu:494-499)
 *    while row_offsets[N] holds the used count.
 */
#ifndef LSMPS3D_AMD_H
#define LSMPS3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
  LSMPS_OK = 0,
  LSMPS_ERR_INVALID_ARGUMENT = 1, /* precondition violated (ref: std::exit paths) */
  LSMPS_ERR_CAPACITY = 2,         /* neighbor / nnz capacity exceeded (neighbor_search.cu:22-30) */
  LSMPS_ERR_DEVICE = 3,           /* HIP runtime error */
  LSMPS_ERR_NOT_PREPARED = 4,     /* moment views used before prepare (moment_matrix.cu:571-576) */
  LSMPS_ERR_NOT_CONVERGED = 5,    /* PPE solver hit max_iterations */
  LSMPS_ERR_OUT_OF_MEMORY = 6,
  LSMPS_ERR_NUMERICAL = 7,        /* non-finite values / breakdown */
  LSMPS_ERR_IO = 8,               /* a file could not be opened or written (std::runtime_error paths) */
  LSMPS_ERR_INTERNAL = 9          /* a library invariant failed (e.g. a device allocation where none may happen) */
};

/* SurfaceType (include/lsmps3d/surface/surface_type.cuh:5-10) */
enum {
  LSMPS_SURFACE_INNER = 0,
  LSMPS_SURFACE_NEAR_SURFACE = 1,
  LSMPS_SURFACE_SURFACE = 2,
  LSMPS_SURFACE_SPLASH = 3
};

/* MomentMatrixKind (include/lsmps3d/moment_matrix/moment_matrix.cuh:13-18) */
enum {
  LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A = 0,
  LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A = 1,
  LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_B = 2,
  LSMPS_MOMENT_FLUID_ONLY_A = 3
};

/* MomentBasisKind (include/lsmps3d/moment_matrix/basis.cuh:11-20) */
enum { LSMPS_BASIS_TYPE_A = 0, LSMPS_BASIS_TYPE_B = 1 };

/* Arithmetic mode of a context (no counterpart: the reference picks `real` at build time). */
enum {
  LSMPS_PRECISION_FP64 = 0,  /* every stage in FP64 */
  LSMPS_PRECISION_MIXED = 1  /* FP32 operator rows + FP32 inner SpMV, FP64 refinement/dots/residual */
};

/* Weight kernel w(r).  The reference uses the linear kernel everywhere
 * (surface_detection.cu:26-31, moment_matrix.cu:145-150, ppe_matrix.cu:148-153,
 *  explicit_update.cu:121-126, pressure_correction.cu:127-132). */
enum { LSMPS_WEIGHT_LINEAR = 0, LSMPS_WEIGHT_QUADRATIC = 1 };

/* PPE Krylov method (replaces AmgxPpeSolver's GMRES, ppe_matrix.cu:586-679).
 * The assembled matrix is non-symmetric and indefinite near wall corners of the
 * free surface, so the default is preconditioned IDR(s); BICGSTAB is IDR(1);
 * PCG (Jacobi only) is only meaningful for symmetric positive systems. */
enum { LSMPS_SOLVER_PCG = 0, LSMPS_SOLVER_BICGSTAB = 1, LSMPS_SOLVER_IDRS = 2 };

/* Context flags.  KERNEL_TIMING brackets the PPE SpMV launches of every eighth
 * IDR(s) cycle (the first of each solve included) with hipEvents on the context
 * stream and accumulates their durations and algorithmic bytes
 * (lsmps_context_kernel_stats); each event pair stalls the stream for a few
 * microseconds, which timing every launch would add to every iteration. */
/* LSMPS_FLAG_GRAPH: the solver captures one IDR(s) cycle (its vector kernels,
 * SpMVs, preconditioner, scalar recurrences and the loop's convergence,
 * reliable-update and restart decisions, made on the device) in a HIP graph and
 * replays it, two cycles in flight, instead of launching every kernel and
 * reading the residual on the host after each cycle; results are bit-identical
 * to the launch path.  With LSMPS_FLAG_KERNEL_TIMING the cycles are launched
 * directly (events recorded inside a graph cannot be timed). */
/* LSMPS_FLAG_AUTOTUNE: the neighbour-list kernels (neighbour count and fill
 * passes, provision, PPE assembly, pressure gradient, shifting, smoothing) pick
 * their lanes per particle by timing each candidate once per neighbour-count
 * histogram of the list they read; every choice computes the same result. */
/* LSMPS_FLAG_SCHWARZ_FP32: in mixed precision, apply the Schwarz block
 * inverses from their FP32 values.  By default the application reads FP16
 * copies with one power-of-two scale per block column (half the bytes of the
 * dominant preconditioner pass; the inverses themselves are still computed in
 * FP32).  No effect in FP64 precision. */
enum {
  LSMPS_FLAG_ASYNC = 1,
  LSMPS_FLAG_KERNEL_TIMING = 2,
  LSMPS_FLAG_GRAPH = 4,
  LSMPS_FLAG_AUTOTUNE = 8,
  LSMPS_FLAG_SCHWARZ_FP32 = 16
};

/* Kernels the autotuner shapes (lsmps_context_lanes / lsmps_context_force_lanes). */
enum {
  LSMPS_TUNE_SCAN_COUNT = 0,
  LSMPS_TUNE_SCAN_WRITE = 1,
  LSMPS_TUNE_PROVISION = 2,
  LSMPS_TUNE_ASSEMBLY = 3,
  LSMPS_TUNE_GRADIENT = 4,
  LSMPS_TUNE_SHIFTING = 5,
  LSMPS_TUNE_SMOOTHING = 6,
  LSMPS_TUNE_KERNELS = 7
};

/* ---- data contract (POD mirrors of the reference views) ------------------ */

/* FluidParticleSoA (include/lsmps3d/particle/particle_data.cuh:7-17) */
typedef struct lsmps_fluid_soa {
  size_t count;
  double* x;
  double* y;
  double* z;
  double* vx;
  double* vy;
  double* vz;
  double* pressure;
  int32_t* surface_type;
} lsmps_fluid_soa;

/* WallParticleSoA (particle_data.cuh:19-30) */
typedef struct lsmps_wall_soa {
  size_t count;
  double* x;
  double* y;
  double* z;
  double* vx;
  double* vy;
  double* vz;
  double* normal_x;
  double* normal_y;
  double* normal_z;
} lsmps_wall_soa;

/* NeighborListView (include/lsmps3d/particle/neighbor_list.cuh:7-12) */
typedef struct lsmps_neighbor_list {
  size_t particle_count;
  size_t neighbor_count; /* allocated capacity of `indices` */
  int32_t* offsets;      /* particle_count + 1 */
  int32_t* indices;
} lsmps_neighbor_list;

/* CellGrid (include/lsmps3d/neighbor/cell_list.cuh:7-15) */
typedef struct lsmps_cell_grid {
  double origin[3];
  double cell_size;
  int32_t dims[3];
} lsmps_cell_grid;

/* CellListView (cell_list.cuh:17-24) */
typedef struct lsmps_cell_list {
  size_t particle_count;
  size_t cell_count;
  int32_t* cell_ids;       /* sorted cell id per sorted position */
  int32_t* sorted_indices; /* particle index per sorted position */
  int32_t* cell_begin;     /* -1 = empty */
  int32_t* cell_end;       /* -1 = empty */
} lsmps_cell_list;

/* SurfaceDetectionDiagnosticsView (include/lsmps3d/surface/surface_detection.cuh:10-21);
 * every member may be NULL. */
typedef struct lsmps_surface_diagnostics {
  int32_t* fluid_neighbor_count;
  int32_t* wall_neighbor_count;
  double* number_density;
  double* number_density_ratio;
  double* anisotropy;
  double* air_open_ratio;
  double* air_anisotropy;
  double* surface_normal_x;
  double* surface_normal_y;
  double* surface_normal_z;
} lsmps_surface_diagnostics;

/* CsrMatrixView (include/lsmps3d/ppe/ppe_matrix.cuh:15-22) */
typedef struct lsmps_csr {
  size_t rows;
  size_t cols;
  size_t nnz; /* capacity semantics, see header comment */
  int32_t* row_offsets;
  int32_t* col_indices;
  double* values;
} lsmps_csr;

/* PpeWorkspaceView (ppe_matrix.cuh:24-30) */
typedef struct lsmps_ppe_workspace {
  lsmps_csr matrix;
  double* rhs;
  double* pressure;
  double* divergence;
  double* pressure_laplacian; /* holds the wall Neumann RHS term, as the reference (ppe_matrix.cu:367-369) */
} lsmps_ppe_workspace;

/* CorrectionVectorView (include/lsmps3d/correction/pressure_correction.cuh:11-16) and
 * ProvisionVelocityView (include/lsmps3d/provision/explicit_update.cuh:11-16). */
typedef struct lsmps_vector3 {
  size_t count;
  double* x;
  double* y;
  double* z;
} lsmps_vector3;

/* MomentMatrixView (moment_matrix.cuh:20-34).  inverse_matrices is N * n * n,
 * column-major per particle: entry (row, col) of particle i at
 * inverse_matrices[i*n*n + row + col*n]. */
typedef struct lsmps_moment_view {
  size_t particle_count;
  int32_t matrix_size;
  int32_t basis_kind;
  int32_t kind;
  int32_t is_ready;
  double support_radius;
  double wall_weight_scale;
  uint64_t geometry_generation;
  const double* inverse_matrices;
  const double* moment_trace;
  const double* regularization_added;
  const int32_t* info;
  const int32_t* inversion_count;
} lsmps_moment_view;

/* SimulationConfig (include/lsmps3d/core/config.hpp:12-82), numeric fields in
 * declaration order; file-path / string fields are host IO and are not mirrored.
 * `weight_kernel` is the one addition (see LSMPS_WEIGHT_*). */
typedef struct lsmps_config {
  /* Geometry */
  double particle_spacing;
  double support_radius;
  double near_surface_radius;
  double cell_origin[3];
  double cell_size;
  int32_t cell_dims[3];
  /* Simulation */
  double time_step;
  double min_time_step;
  double max_time_step;
  double time_step_growth_factor;
  double final_time;
  double output_interval;
  double density;
  double kinematic_viscosity;
  double cfl;
  double gravity[3];
  /* Surface and virtual-light diagnostics */
  int32_t splash_neighbor_threshold;
  int32_t include_wall_neighbors;
  double number_density_ratio_threshold;
  double air_open_ratio_threshold;
  double air_anisotropy_threshold;
  double wall_normal_independence_threshold;
  double virtual_light_cone_cosine;
  /* LSMPS */
  double lsmps_regularization;
  double lsmps_wall_weight_scale;
  /* Pressure correction, particle shifting, and smoothing */
  double ps_displacement_scale;
  double ps_min_distance_ratio;
  double ps_max_displacement_ratio;
  double wall_clearance_ratio;
  double velocity_smoothing_strength;
  /* Addition: weight kernel */
  int32_t weight_kernel;
  int32_t reserved0;
} lsmps_config;

/* Options / report of the PPE solve (replaces the AMGX config string,
 * ppe_matrix.cu:604-612).  The tolerance is on the true FP64 relative
 * residual ||b - A p||_2 / ||b||_2. */
typedef struct lsmps_solver_options {
  int32_t method;          /* LSMPS_SOLVER_* */
  int32_t max_iterations;  /* total Krylov iterations */
  double tolerance;        /* relative true residual */
  double inner_tolerance;  /* mixed mode: factor by which the recursive residual falls
                              between replacements by the true FP64 residual (1e-3) */
  int32_t check_interval;  /* PCG: iterations between host convergence checks */
  int32_t clamp_negative;  /* clamp p<0 after the solve, as ppe_matrix.cu:662 (default 1) */
  int32_t shadow_dimension; /* IDR(s): s, 1..16; 0 (default): 14 with the coarse
                               space, 16 without */
  int32_t preconditioner;   /* LSMPS_PRECOND_* (default SCHWARZ_COARSE) */
  int32_t coarse_cells;     /* SCHWARZ_COARSE: aggregates span coarse_cells^3 cells (default 4) */
} lsmps_solver_options;

/* Preconditioner of the Krylov PPE solve.  JACOBI: D^-1.  SCHWARZ: Jacobi plus
 * exact local solves (restricted additive Schwarz) on the 2-hop neighbourhoods
 * of the rows with a positive diagonal, which make the reference's PPE
 * indefinite.  SCHWARZ_COARSE: SCHWARZ plus an additive coarse-space
 * correction P Ac^-1 P^T over aggregates of coarse_cells^3 cells of the fluid
 * cell grid (Ac = P^T A P, inverted once per solve); it needs the geometry, so
 * only a stepper applies it -- lsmps_ppe_solve on its own uses SCHWARZ.  The
 * matrix is always the reference's. */
enum { LSMPS_PRECOND_JACOBI = 0, LSMPS_PRECOND_SCHWARZ = 1, LSMPS_PRECOND_SCHWARZ_COARSE = 2 };

typedef struct lsmps_solve_report {
  int32_t method;
  int32_t iterations;         /* Krylov iterations (inner iterations in mixed mode) */
  int32_t refinements;        /* outer FP64 residual recomputations */
  int32_t converged;
  double relative_residual;   /* true FP64 relative residual at exit */
  double rhs_norm;
  int32_t schwarz_blocks;     /* LSMPS_PRECOND_SCHWARZ: dense blocks (wrong-sign rows) */
  int32_t schwarz_rows;       /* LSMPS_PRECOND_SCHWARZ: rows summed over all blocks */
  int32_t coarse_unknowns;    /* SCHWARZ_COARSE: aggregates of the applied coarse space (0: none) */
  int32_t graph_cycles;       /* LSMPS_FLAG_GRAPH: IDR cycles replayed from a captured HIP graph */
} lsmps_solve_report;

/* Opaque objects */
typedef struct lsmps_context lsmps_context;
typedef struct lsmps_moment_matrix lsmps_moment_matrix;
typedef struct lsmps_stepper lsmps_stepper;

/* ---- context ---------------------------------------------------------- */
/* No reference counterpart: the reference used the legacy default stream of the
 * current device implicitly.  `precision` is LSMPS_PRECISION_*. */
int lsmps_context_create(int device, int precision, uint32_t flags, lsmps_context** out);
/* Drops the caller's handle.  Moment matrices and steppers created on the
 * context keep it alive, so objects and their context may be destroyed in any
 * order; the context is freed with the last of them. */
int lsmps_context_destroy(lsmps_context* ctx);
int lsmps_context_set_stream(lsmps_context* ctx, void* hip_stream);
int lsmps_context_synchronize(lsmps_context* ctx);
int lsmps_context_precision(const lsmps_context* ctx);
/* No reference counterpart (the reference launches 128 threads, one particle
 * each, neighbor_search.cu:16).  Lanes per particle of the latest launch of a
 * LSMPS_TUNE_* kernel on this context (0 before its first launch), and how many
 * neighbour-count histograms the autotuner has timed so far. */
int lsmps_context_lanes(lsmps_context* ctx, int32_t kernel, int32_t* lanes,
                        int64_t* tuned_keys);
/* Pins a LSMPS_TUNE_* kernel to `lanes` per particle (one of its candidates:
 * count pass 1/2/4, fill pass 4/8/16, assembly 4/8, the others 4/8/16); 0
 * returns it to the default or the autotuner.  Results do not depend on it. */
int lsmps_context_force_lanes(lsmps_context* ctx, int32_t kernel, int32_t lanes);
/* No reference counterpart (the reference hands its PPE to AMGX).  Inverse of
 * a dense n x n row-major FP64 matrix (device pointers, 1 <= n <= 4096) by the
 * kernels that invert the PPE's coarse-space matrix: blocked Gauss-Jordan with
 * partial pivoting, rank-b updates on the FP64 matrix cores.  Lets a caller
 * check those kernels directly.  LSMPS_ERR_NUMERICAL on a zero pivot. */
int lsmps_dense_inverse(lsmps_context* ctx, int32_t n, const double* matrix, double* inverse);
/* No reference counterpart (the reference hands its PPE to AMGX).  y = A x for
 * a square CSR matrix (device pointers, FP64 x and y) through the layout and
 * kernel the PPE solve's Krylov SpMVs use: diagonals of 256-row groups read as
 * x windows, the rest in sliced ELLPACK, values stored in FP64
 * (LSMPS_PRECISION_FP64) or rounded to FP32 (LSMPS_PRECISION_MIXED), products
 * accumulated in FP64.  Builds the layout on each call; lets a caller check it
 * against a CSR product. */
int lsmps_ppe_spmv(lsmps_context* ctx, const lsmps_csr* matrix, const double* x, double* y,
                   int32_t value_precision);
/* PPE SpMV statistics gathered under LSMPS_FLAG_KERNEL_TIMING since the last
 * reset: launches, summed duration (ms) and summed algorithmic bytes
 * (values + column indices + row pointers, each vector element read or written
 * once).  No reference counterpart. */
int lsmps_context_kernel_stats(lsmps_context* ctx, int64_t* launches, double* total_ms,
                               double* total_bytes, int reset);
const char* lsmps_last_error(void);
const char* lsmps_version(void);

/* ---- config (host) ---------------------------------------------------- */
/* default_simulation_config / SimulationConfig{} defaults (config.hpp:12-82). */
int lsmps_default_config(lsmps_config* out);
/* derive_cell_grid_from_particles (src/app/simulation_driver.cu:83-108); host pointers. */
int lsmps_derive_cell_grid(const double* fluid_x, const double* fluid_y, const double* fluid_z,
                           size_t fluid_count, const double* wall_x, const double* wall_y,
                           const double* wall_z, size_t wall_count, lsmps_config* config);

/* ---- neighbor (include/lsmps3d/neighbor/neighbor_search.cuh) ------------ */
/* build_cell_list(const FluidParticleSoA&, const CellGrid&, CellListView) — neighbor_search.cuh:10-12 */
int lsmps_build_cell_list_fluid(lsmps_context* ctx, const lsmps_fluid_soa* particles,
                                const lsmps_cell_grid* grid, const lsmps_cell_list* cells);
/* build_cell_list(const WallParticleSoA&, const CellGrid&, CellListView) — neighbor_search.cuh:14-16 */
int lsmps_build_cell_list_wall(lsmps_context* ctx, const lsmps_wall_soa* particles,
                               const lsmps_cell_grid* grid, const lsmps_cell_list* cells);
/* build_neighbor_lists(...) — neighbor_search.cuh:18-24 */
int lsmps_build_neighbor_lists(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                               const lsmps_wall_soa* walls, const lsmps_config* config,
                               const lsmps_cell_list* fluid_cells, const lsmps_cell_list* wall_cells,
                               const lsmps_neighbor_list* fluid_neighbors,
                               const lsmps_neighbor_list* wall_neighbors);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same tables as lsmps_build_neighbor_lists, bit for bit,
 * with every scratch buffer in a static per-GPU device array instead of the
 * context's buffers; it calls no hipMalloc / hipFree.  Limits: at most 2^20
 * fluid and 2^20 wall particles and 2^20 cells (LSMPS_ERR_CAPACITY beyond, before
 * any kernel runs).  All calls on one GPU share that GPU's array: calls that
 * overlap on one GPU race. */
int lsmps_build_neighbor_lists_noalloc(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_wall_soa* walls, const lsmps_config* config,
                                       const lsmps_cell_list* fluid_cells,
                                       const lsmps_cell_list* wall_cells,
                                       const lsmps_neighbor_list* fluid_neighbors,
                                       const lsmps_neighbor_list* wall_neighbors);

/* ---- surface (include/lsmps3d/surface/surface_detection.cuh) ------------ */
/* compute_uniform_reference_number_density(l0, re) — surface_detection.cuh:28-29 (host) */
int lsmps_reference_number_density(double particle_spacing, double support_radius,
                                   int32_t weight_kernel, double* out);
/* classify_surface_particles(...) — surface_detection.cuh:31-36; diagnostics may be NULL */
int lsmps_classify_surface(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                           const lsmps_wall_soa* walls, const lsmps_neighbor_list* fluid_neighbors,
                           const lsmps_neighbor_list* wall_neighbors, const lsmps_config* config,
                           const lsmps_surface_diagnostics* diagnostics);

/* VirtualLightDiagnosticsView (surface_detection.cuh:23-26): per fluid particle,
 * how many of the 14 virtual-light directions no neighbor blocks, as a count and
 * as count / 14.  Either member may be NULL, not both. */
typedef struct lsmps_virtual_light_diagnostics {
  int32_t* open_direction_count;
  double* open_fraction;
} lsmps_virtual_light_diagnostics;
/* compute_virtual_light_diagnostics(...) — surface_detection.cuh:38-43,
 * src/surface/surface_detection.cu:350-475,555-578.  A direction is blocked when
 * some neighbor within the support radius lies inside the cone of half-angle
 * acos(virtual_light_cone_cosine) around it (wall neighbors only when
 * include_wall_neighbors is set).  LSMPS_ERR_INVALID_ARGUMENT with the
 * reference's message for a cone cosine outside [-1, 1] or no output buffer. */
int lsmps_compute_virtual_light_diagnostics(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                            const lsmps_wall_soa* walls,
                                            const lsmps_neighbor_list* fluid_neighbors,
                                            const lsmps_neighbor_list* wall_neighbors,
                                            const lsmps_config* config,
                                            const lsmps_virtual_light_diagnostics* diagnostics);

/* ---- moment matrices (include/lsmps3d/moment_matrix/moment_matrix.cuh) -- */
/* DeviceMomentMatrix(fluid_capacity, config) — moment_matrix.cuh:36-40 */
int lsmps_moment_create(lsmps_context* ctx, size_t fluid_capacity, const lsmps_config* config,
                        lsmps_moment_matrix** out);
int lsmps_moment_destroy(lsmps_moment_matrix* moments);
/* DeviceMomentMatrix::resize / set_config / bytes — moment_matrix.cuh:47-52 */
int lsmps_moment_resize(lsmps_moment_matrix* moments, size_t fluid_capacity);
int lsmps_moment_set_config(lsmps_moment_matrix* moments, const lsmps_config* config);
size_t lsmps_moment_bytes(const lsmps_moment_matrix* moments);
/* DeviceMomentMatrix::prepare_matrices(...) — moment_matrix.cuh:54-58.  Builds the
 * operator rows consumed by provision / PPE / correction for the two consumed kinds;
 * full inverses of any kind are materialised when lsmps_moment_get_view asks for them. */
int lsmps_moment_prepare(lsmps_moment_matrix* moments, const lsmps_fluid_soa* fluid,
                         const lsmps_wall_soa* walls, const lsmps_neighbor_list* fluid_neighbors,
                         const lsmps_neighbor_list* wall_neighbors, uint64_t geometry_generation);
/* velocity_type_a() / pressure_type_a() / pressure_type_b() / fluid_only_type_a() —
 * moment_matrix.cuh:60-63.  `kind` is LSMPS_MOMENT_*. */
int lsmps_moment_get_view(lsmps_moment_matrix* moments, int32_t kind, lsmps_moment_view* out);

/* ---- provision (include/lsmps3d/provision/explicit_update.cuh) --------- */
/* DeviceProvisionExplicitUpdate::compute_temporary_velocity(...) — explicit_update.cuh:71-78.
 * `laplacian_workspace` may be NULL (the library then owns a workspace). */
int lsmps_provision_compute(lsmps_context* ctx, const lsmps_config* config,
                            const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                            const lsmps_neighbor_list* fluid_neighbors,
                            const lsmps_neighbor_list* wall_neighbors, lsmps_moment_matrix* moments,
                            uint64_t geometry_generation, const lsmps_fluid_soa* temporary_velocity,
                            const lsmps_wall_soa* temporary_wall_velocity,
                            const lsmps_vector3* laplacian_workspace);
/* compute_provision_temporary_velocity — explicit_update.cuh:84-87 */
int lsmps_provision_temporary_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_config* config,
                                       const lsmps_vector3* velocity_laplacian,
                                       const lsmps_fluid_soa* temporary_velocity);
/* compute_provision_temporary_wall_velocity — explicit_update.cuh:89-91 */
int lsmps_provision_temporary_wall_velocity(lsmps_context* ctx, const lsmps_wall_soa* walls,
                                            const lsmps_config* config,
                                            const lsmps_wall_soa* temporary_wall_velocity);

/* ---- PPE (include/lsmps3d/ppe/ppe_matrix.cuh) --------------------------- */
/* DevicePpeMatrixAssembler::assemble(...) — ppe_matrix.cuh:82-89; writes into the
 * caller's workspace (matrix.nnz must be >= N + fluid neighbor capacity). */
int lsmps_ppe_assemble(lsmps_context* ctx, const lsmps_config* config,
                       const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                       const lsmps_neighbor_list* fluid_neighbors,
                       const lsmps_neighbor_list* wall_neighbors,
                       const lsmps_fluid_soa* temporary_velocity,
                       const lsmps_wall_soa* temporary_wall_velocity, lsmps_moment_matrix* moments,
                       uint64_t geometry_generation, const lsmps_ppe_workspace* workspace);
/* assemble_ppe_matrix_and_rhs(fluid, walls, config, fluid_neighbors, wall_neighbors,
 * temporary_velocity, temporary_wall_velocity, const MomentMatrixView& velocity_moment,
 * const MomentMatrixView& pressure_moment, const PpeWorkspaceView&) — ppe_matrix.cuh:120-129,
 * ppe_matrix.cu:517-551.  The views' explicit inverses are read (rows 0-2 of the
 * velocity kind, rows 3+4+5 of the pressure kind); a view that is not a prepared
 * Type-A matrix of the expected kind gives LSMPS_ERR_NOT_PREPARED with the
 * reference's message (validate_ppe_moment_matrix, ppe_matrix.cu:108-122). */
int lsmps_ppe_assemble_views(lsmps_context* ctx, const lsmps_config* config,
                             const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                             const lsmps_neighbor_list* fluid_neighbors,
                             const lsmps_neighbor_list* wall_neighbors,
                             const lsmps_fluid_soa* temporary_velocity,
                             const lsmps_wall_soa* temporary_wall_velocity,
                             const lsmps_moment_view* velocity_moment,
                             const lsmps_moment_view* pressure_moment,
                             const lsmps_ppe_workspace* workspace);
/* AmgxPpeSolver::solve(const CsrMatrixView&, const real* rhs, real* pressure) —
 * ppe_matrix.cuh:112.  Hand-written IDR(s) / BiCGStab / PCG from p = 0, then the
 * reference's clamp p = max(p, 0) (ppe_matrix.cu:662).  options may be NULL
 * (lsmps_default_solver_options: IDR(s), Schwarz preconditioner, relative true
 * residual 1e-10); report may be NULL.  Returns LSMPS_ERR_NOT_CONVERGED (with the
 * iterate written and the report filled) when max_iterations is reached. */
int lsmps_ppe_solve(lsmps_context* ctx, const lsmps_csr* matrix, const double* rhs,
                    double* pressure, const lsmps_solver_options* options,
                    lsmps_solve_report* report);
int lsmps_default_solver_options(lsmps_solver_options* out);

/* ---- correction (include/lsmps3d/correction/pressure_correction.cuh) ---- */
/* DevicePressureCorrection::apply(...) — pressure_correction.cuh:70-77.  The three
 * workspaces (3N each) are caller-owned here. */
int lsmps_correction_apply(lsmps_context* ctx, const lsmps_config* config,
                           const lsmps_fluid_soa* fluid, const lsmps_wall_soa* walls,
                           const lsmps_neighbor_list* fluid_neighbors,
                           const lsmps_neighbor_list* wall_neighbors,
                           const lsmps_fluid_soa* temporary_velocity, double* pressure,
                           lsmps_moment_matrix* moments, uint64_t geometry_generation,
                           const lsmps_vector3* pressure_gradient,
                           const lsmps_vector3* ps_displacement,
                           const lsmps_vector3* smoothed_velocity);
/* clamp_particle_pressure — pressure_correction.cuh:91 */
int lsmps_clamp_particle_pressure(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                  double* pressure);
/* compute_pressure_gradient — pressure_correction.cuh:93-100; needs prepared moments */
int lsmps_compute_pressure_gradient(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                    const lsmps_wall_soa* walls,
                                    const lsmps_neighbor_list* fluid_neighbors,
                                    const lsmps_neighbor_list* wall_neighbors,
                                    const double* pressure, const lsmps_config* config,
                                    lsmps_moment_matrix* moments,
                                    const lsmps_vector3* pressure_gradient);
/* compute_pressure_gradient(..., const MomentMatrixView& pressure_moment,
 * CorrectionVectorView) — pressure_correction.cuh:93-100, pressure_correction.cu:657-682:
 * rows 0-2 of the view's explicit inverses; validation as
 * validate_pressure_moment_matrix (pressure_correction.cu:96-111). */
int lsmps_compute_pressure_gradient_view(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                         const lsmps_wall_soa* walls,
                                         const lsmps_neighbor_list* fluid_neighbors,
                                         const lsmps_neighbor_list* wall_neighbors,
                                         const double* pressure, const lsmps_config* config,
                                         const lsmps_moment_view* pressure_moment,
                                         const lsmps_vector3* pressure_gradient);
/* compute_particle_shifting_displacement — pressure_correction.cuh:102-107 */
int lsmps_compute_particle_shifting(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                    const lsmps_wall_soa* walls,
                                    const lsmps_neighbor_list* fluid_neighbors,
                                    const lsmps_neighbor_list* wall_neighbors,
                                    const lsmps_config* config,
                                    const lsmps_vector3* ps_displacement);
/* update_velocity_and_position — pressure_correction.cuh:109-113 */
int lsmps_update_velocity_and_position(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                       const lsmps_fluid_soa* temporary_velocity,
                                       const lsmps_config* config,
                                       const lsmps_vector3* pressure_gradient,
                                       const lsmps_vector3* ps_displacement);
/* anti_penetration_kernel launch inside apply — pressure_correction.cu:404-444,639-643 */
int lsmps_apply_anti_penetration(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                 const lsmps_wall_soa* walls,
                                 const lsmps_neighbor_list* wall_neighbors,
                                 const lsmps_config* config);
/* smooth_particle_velocity — pressure_correction.cuh:115-118 */
int lsmps_smooth_particle_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid,
                                   const lsmps_neighbor_list* fluid_neighbors,
                                   const lsmps_config* config,
                                   const lsmps_vector3* smoothed_velocity);

/* ---- fused step (src/app/simulation_driver.cu:336-393 stage order) ------ */
/* A stepper owns every intermediate buffer of one time step: cell lists, the two
 * CSR tables (capacity `max_neighbors_per_particle` per fluid particle, 256 in the
 * reference driver, simulation_driver.cu:29-30), moment operators, u*, the PPE
 * system and the correction workspaces. */
int lsmps_stepper_create(lsmps_context* ctx, size_t fluid_capacity, size_t wall_capacity,
                         size_t max_neighbors_per_particle, const lsmps_config* config,
                         lsmps_stepper** out);
int lsmps_stepper_destroy(lsmps_stepper* stepper);
int lsmps_stepper_set_config(lsmps_stepper* stepper, const lsmps_config* config);
int lsmps_stepper_set_solver(lsmps_stepper* stepper, const lsmps_solver_options* options);
/* Batched independent tanks (the many-tanks workload; no reference counterpart:
 * the reference steps one system per process).  The SoAs given to
 * lsmps_stepper_step then hold the tanks back to back: tank k owns fluid
 * particles [fluid_offsets[k], fluid_offsets[k+1]) and wall particles
 * [wall_offsets[k], wall_offsets[k+1]), and its own cell grid grids[k]; the
 * other config fields are shared.  Every stage runs once over all tanks: the
 * neighbor search is segment-aware (each particle scans its tank's grid only,
 * with the arithmetic of a single-tank run), and the PPE is one block-diagonal
 * system solved to the tolerance relative to the whole right-hand side.  Host
 * arrays of segment_count + 1 offsets and segment_count grids; segment_count = 0
 * returns to one system on the config's grid. */
int lsmps_stepper_set_segments(lsmps_stepper* stepper, int32_t segment_count,
                               const int64_t* fluid_offsets, const int64_t* wall_offsets,
                               const lsmps_cell_grid* grids);
size_t lsmps_stepper_bytes(const lsmps_stepper* stepper);
/* One step: rebuild_geometry (neighbor CSR + surface labels, generation++),
 * provision, PPE assembly, PPE solve, pressure correction (with shifting,
 * anti-penetration, smoothing).  fluid x, v, pressure, surface_type are updated
 * in place; walls are read only. */
int lsmps_stepper_step(lsmps_stepper* stepper, const lsmps_fluid_soa* fluid,
                       const lsmps_wall_soa* walls, lsmps_solve_report* report);
/* Stage times of the last step in ms (hipEvents on the context stream):
 * [0] geometry, [1] moments, [2] provision, [3] assembly, [4] solve,
 * [5] correction, [6] total.  No reference counterpart (the reference has no timers). */
int lsmps_stepper_stage_times(const lsmps_stepper* stepper, double* ms, int count);
/* Views of the stepper's internal buffers (device pointers; valid until the next
 * step or destroy) for inspection and tests. */
int lsmps_stepper_neighbor_lists(lsmps_stepper* stepper, lsmps_neighbor_list* fluid_neighbors,
                                 lsmps_neighbor_list* wall_neighbors);
int lsmps_stepper_ppe_workspace(lsmps_stepper* stepper, lsmps_ppe_workspace* out);
int lsmps_stepper_temporary_velocity(lsmps_stepper* stepper, lsmps_fluid_soa* fluid_out,
                                     lsmps_wall_soa* wall_out);
int lsmps_stepper_correction_vectors(lsmps_stepper* stepper, lsmps_vector3* pressure_gradient,
                                     lsmps_vector3* ps_displacement);
lsmps_moment_matrix* lsmps_stepper_moments(lsmps_stepper* stepper);

/* ---- multi-step driver --------------------------------------------------- */
/* TimeStepLimits (include/lsmps3d/core/time_step.hpp:8-17). */
typedef struct lsmps_time_limits {
  double current_time_step;
  double min_time_step;
  double max_time_step;
  double growth_factor;
  double cfl;
  double particle_spacing;
  double final_time;
  double output_interval;
} lsmps_time_limits;

/* TimeStepStatus (time_step.hpp:19-29); the two flags are 0 / 1. */
typedef struct lsmps_time_status {
  size_t step_index;
  size_t output_index;
  double previous_time;
  double current_time;
  double time_step;
  double cfl_time_step;
  double max_velocity;
  int32_t should_output;
  int32_t reached_final_time;
} lsmps_time_status;

/* SimulationTimeManager (time_step.hpp:31-82, src/core/time_step.cpp:47-172):
 * a host-side clock, no device work. */
typedef struct lsmps_time_manager lsmps_time_manager;
/* make_time_step_limits (time_step.cpp:47-58) */
int lsmps_time_limits_from_config(const lsmps_config* config, lsmps_time_limits* out);
/* validate_time_step_limits (time_step.cpp:60-77): LSMPS_ERR_INVALID_ARGUMENT
 * with the reference's message where it would throw std::invalid_argument. */
int lsmps_time_limits_validate(const lsmps_time_limits* limits);
/* SimulationTimeManager(limits) / reset(limits) (time_step.cpp:79-100) */
int lsmps_time_manager_create(const lsmps_time_limits* limits, lsmps_time_manager** out);
int lsmps_time_manager_destroy(lsmps_time_manager* clock);
int lsmps_time_manager_reset(lsmps_time_manager* clock, const lsmps_time_limits* limits);
/* current_time / current_time_step / step_index / output_index / output_due /
 * finished (time_step.hpp:44-64, time_step.cpp:102-104); NULL outputs are skipped. */
int lsmps_time_manager_state(const lsmps_time_manager* clock, double* current_time,
                             double* current_time_step, size_t* step_index,
                             size_t* output_index, int32_t* output_due, int32_t* finished);
/* cfl_limited_time_step / next_time_step (time_step.cpp:106-124) */
int lsmps_time_manager_next_time_step(const lsmps_time_manager* clock, double max_velocity,
                                      double* time_step, double* cfl_time_step);
/* advance / mark_initial_output (time_step.cpp:126-170) */
int lsmps_time_manager_advance(lsmps_time_manager* clock, double max_velocity,
                               lsmps_time_status* status);
int lsmps_time_manager_mark_initial_output(lsmps_time_manager* clock, lsmps_time_status* status);

/* max_velocity (src/app/simulation_driver.cu:122-143): the largest |v| of the
 * fluid, non-finite magnitudes skipped, as a device reduction (one double read
 * back). */
int lsmps_max_velocity(lsmps_context* ctx, const lsmps_fluid_soa* fluid, double* out);

/* Output hook of lsmps_stepper_run: called for the initial state and after every
 * step whose status has should_output or reached_final_time, with the fluid
 * buffers holding that state (lists and labels rebuilt at its positions).
 * Returning non-zero stops the run. */
typedef int (*lsmps_output_callback)(const lsmps_time_status* status, void* user);

/* The driver loop (run_simulation, simulation_driver.cu:340-391) on the
 * stepper: while the clock has not reached final_time (and at most max_steps
 * steps, < 0 for no limit): max velocity, clock advance, config.time_step = dt,
 * one step, and on output steps a rebuild of the neighbor lists and labels at
 * the new positions (the reference's end-of-step rebuild_geometry) followed by
 * the callback.  The clock's initial output, if still pending, is reported
 * first.  `last` (may be NULL) receives the last status.  Returns
 * LSMPS_ERR_NOT_CONVERGED when some step's solve missed the tolerance (the run
 * continues with that pressure, as the reference does). */
int lsmps_stepper_run(lsmps_stepper* stepper, const lsmps_fluid_soa* fluid,
                      const lsmps_wall_soa* walls, lsmps_time_manager* clock, int64_t max_steps,
                      lsmps_output_callback callback, void* user, lsmps_time_status* last);
/* rebuild_geometry (simulation_driver.cu:252-264) at the fluid's current
 * positions: neighbor lists and labels (fluid.surface_type), generation + 1. */
int lsmps_stepper_refresh_geometry(lsmps_stepper* stepper, const lsmps_fluid_soa* fluid,
                                   const lsmps_wall_soa* walls);

/* ---- on-disk formats (host code) ----------------------------------------- */
/* The string and flag fields of SimulationConfig (config.hpp:73-80) that
 * lsmps_config leaves out, NUL-terminated. */
typedef struct lsmps_config_files {
  char fluid_particle_file[1024];
  char wall_particle_file[1024];
  char output_directory[1024];
  char vtk_file_prefix[256];
  char amgx_config_path[1024];
  int32_t vtk_write_point_fields;
  int32_t amgx_print_solve_stats;
} lsmps_config_files;
/* SimulationConfig{} file-field defaults (config.hpp:73-80). */
int lsmps_default_config_files(lsmps_config_files* out);
/* validate_simulation_config (src/core/config.cpp:278-335); `files` may be NULL
 * to check the numeric fields only. */
int lsmps_config_validate(const lsmps_config* config, const lsmps_config_files* files);
/* load_simulation_config (config.cpp:337-385): INI sections geometry,
 * simulation, surface, lsmps, correction, files; '#' comments; defaults for
 * absent keys; unknown keys and sections are errors.  weight_kernel (not an INI
 * key) is left at its default. */
int lsmps_config_load(const char* path, lsmps_config* config, lsmps_config_files* files);
/* save_simulation_config (config.cpp:387-431), same text as the reference. */
int lsmps_config_save(const char* path, const lsmps_config* config,
                      const lsmps_config_files* files);
/* FileManager::load_fluid_particles / load_wall_particles (src/io/file_manager.cpp:
 * 126-197): CSV rows of x,y,z,vx,vy,vz (walls: + nx,ny,nz); '#' comments,
 * blank lines and one header row before the first record are skipped.  HOST
 * arrays: with host == NULL (or host->x == NULL) only the record count is
 * returned; otherwise host->count is the capacity, and pressure / surface_type
 * (if given) are set to 0 / Inner. */
int lsmps_read_fluid_csv(const char* path, const lsmps_fluid_soa* host, size_t* count);
int lsmps_read_wall_csv(const char* path, const lsmps_wall_soa* host, size_t* count);
/* Point field of a VTK file (HostVtkPointFields, vtk_writer.hpp:17-46): REAL
 * scalars read x, INT scalars ivalues, VECTORs x, y, z; host arrays of count
 * values. */
enum { LSMPS_VTK_REAL_SCALAR = 0, LSMPS_VTK_INT_SCALAR = 1, LSMPS_VTK_VECTOR = 2 };
typedef struct lsmps_vtk_field {
  const char* name;
  int32_t kind;
  size_t count;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* ivalues;
} lsmps_vtk_field;
/* LegacyVtkWriter::make_path (vtk_writer.cpp:83-88): directory/prefix_NNNNNN.vtk */
int lsmps_vtk_path(const char* output_directory, const char* prefix, size_t step, char* out,
                   size_t capacity);
/* LegacyVtkWriter::write (vtk_writer.cpp:90-149): legacy ASCII POLYDATA with
 * one vertex per particle; point fields (vectors, then real scalars, then
 * integer scalars) when write_point_fields is non-zero. */
int lsmps_write_vtk(const char* output_directory, const char* prefix, int32_t write_point_fields,
                    size_t step, size_t count, const double* x, const double* y, const double* z,
                    const lsmps_vtk_field* fields, int32_t field_count);

/* ---- adjoint of one step (no reference counterpart) -------------------- */
/* Stepper modes (bit flags, default 0):
 *  LSMPS_STEP_RECORD           keep what the adjoint needs and the step
 *                              overwrites (start positions and velocities, the
 *                              PPE solution before the clamp): 7 doubles per
 *                              fluid particle.
 *  LSMPS_STEP_FROZEN_GEOMETRY  skip the neighbor search and the surface
 *                              classification and reuse the lists of the
 *                              previous step and the labels already in
 *                              fluid.surface_type (same particle counts).  The
 *                              step is then a smooth function of positions,
 *                              velocities and gravity, which is what central
 *                              finite differences of the adjoint need. */
enum { LSMPS_STEP_RECORD = 1u, LSMPS_STEP_FROZEN_GEOMETRY = 2u };
int lsmps_stepper_set_mode(lsmps_stepper* stepper, uint32_t mode);

/* Outputs of lsmps_adjoint_step.  Device pointers may be NULL to skip a
 * gradient; gravity and the report are host values. */
typedef struct lsmps_adjoint_grads {
  lsmps_vector3 position;   /* dL/dx at the start of the step (fluid) */
  lsmps_vector3 velocity;   /* dL/dv at the start of the step (fluid) */
  double* pressure;         /* dL/dp for the clamped pressure the correction read */
  double gravity[3];        /* dL/dg (body force) */
  lsmps_solve_report transpose_report; /* the A^T solve of the adjoint */
} lsmps_adjoint_grads;

/* Reverse mode of the stepper's last step, which must have been taken with
 * LSMPS_STEP_RECORD.  L is any scalar function of the step's outputs; the caller
 * passes dL/dv_new (3N) and optionally dL/dx_new (NULL = 0).  Neighbor lists and
 * surface labels are held fixed; every stage is differentiated: moment
 * operators (both kinds), provision, PPE assembly, PPE solve (one solve with
 * A^T, same solver and tolerance, no clamp), the pressure clamp, pressure
 * gradient, particle shifting, update, anti-penetration and smoothing.
 * `fluid` is the fluid after the step (its surface_type holds the step's
 * labels), `walls` the walls of the step.  Arithmetic is FP64 in both precision
 * modes except the A^T solve, which follows the context precision. */
int lsmps_adjoint_step(lsmps_stepper* stepper, const lsmps_fluid_soa* fluid,
                       const lsmps_wall_soa* walls, const lsmps_vector3* grad_velocity_out,
                       const lsmps_vector3* grad_position_out, lsmps_adjoint_grads* grads);

#ifdef __cplusplus
}
#endif

#endif /* LSMPS3D_AMD_H */
