"""ctypes mirror of include/lsmps3d_amd.h (POD structs and enums).

The layouts follow the C header field for field; tests assert the sizes the
library reports match these (see tests/test_abi.py).
"""
import ctypes as C

LSMPS_OK = 0
LSMPS_ERR_INVALID_ARGUMENT = 1
LSMPS_ERR_CAPACITY = 2
LSMPS_ERR_DEVICE = 3
LSMPS_ERR_NOT_PREPARED = 4
LSMPS_ERR_NOT_CONVERGED = 5
LSMPS_ERR_OUT_OF_MEMORY = 6
LSMPS_ERR_NUMERICAL = 7
LSMPS_ERR_IO = 8
LSMPS_ERR_INTERNAL = 9

INNER, NEAR_SURFACE, SURFACE, SPLASH = 0, 1, 2, 3

MOMENT_VELOCITY_WALL_DIRICHLET_A = 0
MOMENT_PRESSURE_WALL_NEUMANN_A = 1
MOMENT_PRESSURE_WALL_NEUMANN_B = 2
MOMENT_FLUID_ONLY_A = 3

PRECISION_FP64 = 0
PRECISION_MIXED = 1

WEIGHT_LINEAR = 0
WEIGHT_QUADRATIC = 1

SOLVER_PCG = 0
SOLVER_BICGSTAB = 1
SOLVER_IDRS = 2

PRECOND_JACOBI = 0
PRECOND_SCHWARZ = 1
PRECOND_SCHWARZ_COARSE = 2

STEP_RECORD = 1
STEP_FROZEN_GEOMETRY = 2

FLAG_ASYNC = 1
FLAG_KERNEL_TIMING = 2
FLAG_GRAPH = 4
FLAG_AUTOTUNE = 8
FLAG_SCHWARZ_FP32 = 16

# kernels the autotuner shapes (LSMPS_TUNE_*) and their lanes-per-particle candidates
TUNED_KERNELS = {"scan_count": 0, "scan_write": 1, "provision": 2, "assembly": 3,
                 "gradient": 4, "shifting": 5, "smoothing": 6}
TUNED_CANDIDATES = {"scan_count": (1, 2, 4), "scan_write": (4, 8, 16), "provision": (4, 8, 16),
                    "assembly": (4, 8), "gradient": (4, 8, 16), "shifting": (4, 8, 16),
                    "smoothing": (4, 8, 16)}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class FluidSoA(C.Structure):
    _fields_ = [("count", C.c_size_t), ("x", _dp), ("y", _dp), ("z", _dp), ("vx", _dp),
                ("vy", _dp), ("vz", _dp), ("pressure", _dp), ("surface_type", _ip)]


class WallSoA(C.Structure):
    _fields_ = [("count", C.c_size_t), ("x", _dp), ("y", _dp), ("z", _dp), ("vx", _dp),
                ("vy", _dp), ("vz", _dp), ("normal_x", _dp), ("normal_y", _dp),
                ("normal_z", _dp)]


class NeighborList(C.Structure):
    _fields_ = [("particle_count", C.c_size_t), ("neighbor_count", C.c_size_t),
                ("offsets", _ip), ("indices", _ip)]


class CellGrid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("cell_size", C.c_double),
                ("dims", C.c_int32 * 3)]


class CellList(C.Structure):
    _fields_ = [("particle_count", C.c_size_t), ("cell_count", C.c_size_t),
                ("cell_ids", _ip), ("sorted_indices", _ip), ("cell_begin", _ip),
                ("cell_end", _ip)]


class SurfaceDiagnostics(C.Structure):
    _fields_ = [("fluid_neighbor_count", _ip), ("wall_neighbor_count", _ip),
                ("number_density", _dp), ("number_density_ratio", _dp), ("anisotropy", _dp),
                ("air_open_ratio", _dp), ("air_anisotropy", _dp), ("surface_normal_x", _dp),
                ("surface_normal_y", _dp), ("surface_normal_z", _dp)]


class VirtualLightDiagnostics(C.Structure):
    _fields_ = [("open_direction_count", _ip), ("open_fraction", _dp)]


class Csr(C.Structure):
    _fields_ = [("rows", C.c_size_t), ("cols", C.c_size_t), ("nnz", C.c_size_t),
                ("row_offsets", _ip), ("col_indices", _ip), ("values", _dp)]


class PpeWorkspace(C.Structure):
    _fields_ = [("matrix", Csr), ("rhs", _dp), ("pressure", _dp), ("divergence", _dp),
                ("pressure_laplacian", _dp)]


class Vector3(C.Structure):
    _fields_ = [("count", C.c_size_t), ("x", _dp), ("y", _dp), ("z", _dp)]


class MomentView(C.Structure):
    _fields_ = [("particle_count", C.c_size_t), ("matrix_size", C.c_int32),
                ("basis_kind", C.c_int32), ("kind", C.c_int32), ("is_ready", C.c_int32),
                ("support_radius", C.c_double), ("wall_weight_scale", C.c_double),
                ("geometry_generation", C.c_uint64), ("inverse_matrices", _dp),
                ("moment_trace", _dp), ("regularization_added", _dp), ("info", _ip),
                ("inversion_count", _ip)]


class Config(C.Structure):
    _fields_ = [
        ("particle_spacing", C.c_double), ("support_radius", C.c_double),
        ("near_surface_radius", C.c_double), ("cell_origin", C.c_double * 3),
        ("cell_size", C.c_double), ("cell_dims", C.c_int32 * 3),
        ("time_step", C.c_double), ("min_time_step", C.c_double),
        ("max_time_step", C.c_double), ("time_step_growth_factor", C.c_double),
        ("final_time", C.c_double), ("output_interval", C.c_double), ("density", C.c_double),
        ("kinematic_viscosity", C.c_double), ("cfl", C.c_double), ("gravity", C.c_double * 3),
        ("splash_neighbor_threshold", C.c_int32), ("include_wall_neighbors", C.c_int32),
        ("number_density_ratio_threshold", C.c_double),
        ("air_open_ratio_threshold", C.c_double), ("air_anisotropy_threshold", C.c_double),
        ("wall_normal_independence_threshold", C.c_double),
        ("virtual_light_cone_cosine", C.c_double),
        ("lsmps_regularization", C.c_double), ("lsmps_wall_weight_scale", C.c_double),
        ("ps_displacement_scale", C.c_double), ("ps_min_distance_ratio", C.c_double),
        ("ps_max_displacement_ratio", C.c_double), ("wall_clearance_ratio", C.c_double),
        ("velocity_smoothing_strength", C.c_double),
        ("weight_kernel", C.c_int32), ("reserved0", C.c_int32),
    ]

    def copy(self):
        other = Config()
        C.memmove(C.byref(other), C.byref(self), C.sizeof(Config))
        return other

    def grid(self):
        g = CellGrid()
        for k in range(3):
            g.origin[k] = self.cell_origin[k]
            g.dims[k] = self.cell_dims[k]
        g.cell_size = self.cell_size
        return g

    def cell_count(self):
        return int(self.cell_dims[0]) * int(self.cell_dims[1]) * int(self.cell_dims[2])


class SolverOptions(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_iterations", C.c_int32),
                ("tolerance", C.c_double), ("inner_tolerance", C.c_double),
                ("check_interval", C.c_int32), ("clamp_negative", C.c_int32),
                ("shadow_dimension", C.c_int32), ("preconditioner", C.c_int32),
                ("coarse_cells", C.c_int32)]


class SolveReport(C.Structure):
    _fields_ = [("method", C.c_int32), ("iterations", C.c_int32), ("refinements", C.c_int32),
                ("converged", C.c_int32), ("relative_residual", C.c_double),
                ("rhs_norm", C.c_double), ("schwarz_blocks", C.c_int32),
                ("schwarz_rows", C.c_int32), ("coarse_unknowns", C.c_int32),
                ("graph_cycles", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ConfigFiles(C.Structure):
    _fields_ = [("fluid_particle_file", C.c_char * 1024), ("wall_particle_file", C.c_char * 1024),
                ("output_directory", C.c_char * 1024), ("vtk_file_prefix", C.c_char * 256),
                ("amgx_config_path", C.c_char * 1024), ("vtk_write_point_fields", C.c_int32),
                ("amgx_print_solve_stats", C.c_int32)]


VTK_REAL_SCALAR, VTK_INT_SCALAR, VTK_VECTOR = 0, 1, 2


class VtkField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("kind", C.c_int32), ("count", C.c_size_t),
                ("x", _dp), ("y", _dp), ("z", _dp), ("ivalues", _ip)]


class TimeLimits(C.Structure):
    _fields_ = [("current_time_step", C.c_double), ("min_time_step", C.c_double),
                ("max_time_step", C.c_double), ("growth_factor", C.c_double), ("cfl", C.c_double),
                ("particle_spacing", C.c_double), ("final_time", C.c_double),
                ("output_interval", C.c_double)]


class TimeStatus(C.Structure):
    _fields_ = [("step_index", C.c_size_t), ("output_index", C.c_size_t),
                ("previous_time", C.c_double), ("current_time", C.c_double),
                ("time_step", C.c_double), ("cfl_time_step", C.c_double),
                ("max_velocity", C.c_double), ("should_output", C.c_int32),
                ("reached_final_time", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


OutputCallback = C.CFUNCTYPE(C.c_int, C.POINTER(TimeStatus), C.c_void_p)


class AdjointGrads(C.Structure):
    _fields_ = [("position", Vector3), ("velocity", Vector3), ("pressure", _dp),
                ("gravity", C.c_double * 3), ("transpose_report", SolveReport)]


def default_config():
    """SimulationConfig{} defaults (include/lsmps3d/core/config.hpp:12-82)."""
    c = Config()
    c.particle_spacing = 0.01
    c.support_radius = 0.031
    c.near_surface_radius = 0.015
    c.cell_size = c.support_radius
    for k in range(3):
        c.cell_origin[k] = 0.0
        c.cell_dims[k] = 1
    c.time_step = 1.0e-4
    c.min_time_step = 1.0e-6
    c.max_time_step = 1.0e-3
    c.time_step_growth_factor = 1.05
    c.final_time = 1.0
    c.output_interval = 0.01
    c.density = 1000.0
    c.kinematic_viscosity = 1.0e-6
    c.cfl = 0.2
    c.gravity[0], c.gravity[1], c.gravity[2] = 0.0, 0.0, -9.81
    c.splash_neighbor_threshold = 4
    c.include_wall_neighbors = 1
    c.number_density_ratio_threshold = 0.85
    c.air_open_ratio_threshold = 0.33
    c.air_anisotropy_threshold = 0.033
    c.wall_normal_independence_threshold = 0.25
    c.virtual_light_cone_cosine = 0.8660254037844386
    c.lsmps_regularization = 1.0e-8
    c.lsmps_wall_weight_scale = 1.0
    c.ps_displacement_scale = 0.05
    c.ps_min_distance_ratio = 0.85
    c.ps_max_displacement_ratio = 0.2
    c.wall_clearance_ratio = 0.25
    c.velocity_smoothing_strength = 0.1
    c.weight_kernel = WEIGHT_LINEAR
    c.reserved0 = 0
    return c


def default_solver_options():
    o = SolverOptions()
    o.method = SOLVER_IDRS
    o.max_iterations = 50000
    o.tolerance = 1.0e-10
    o.inner_tolerance = 1.0e-3
    o.check_interval = 16
    o.clamp_negative = 1
    o.shadow_dimension = 0  # 14 with the coarse space, 16 without
    o.preconditioner = PRECOND_SCHWARZ_COARSE
    o.coarse_cells = 4
    return o


def dptr(addr):
    """Cast an integer address (numpy .ctypes.data or torch .data_ptr()) to double*."""
    return C.cast(C.c_void_p(addr), _dp) if addr else _dp()


def iptr(addr):
    return C.cast(C.c_void_p(addr), _ip) if addr else _ip()
