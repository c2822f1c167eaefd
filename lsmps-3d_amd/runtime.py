"""ctypes bindings of liblsmps3d_amd.so and device-buffer plumbing.

torch is used only to own device memory and streams; every computation runs in
the HIP library.  Loading fails loudly when the library is not built.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsmps3d_amd.so")

EXPORTS = [
    "lsmps_context_create", "lsmps_context_destroy", "lsmps_context_set_stream",
    "lsmps_context_synchronize", "lsmps_context_precision", "lsmps_last_error", "lsmps_version",
    "lsmps_default_config", "lsmps_derive_cell_grid", "lsmps_build_cell_list_fluid",
    "lsmps_build_cell_list_wall", "lsmps_build_neighbor_lists", "lsmps_reference_number_density",
    "lsmps_classify_surface", "lsmps_moment_create", "lsmps_moment_destroy", "lsmps_moment_resize",
    "lsmps_moment_set_config", "lsmps_moment_bytes", "lsmps_moment_prepare",
    "lsmps_moment_get_view", "lsmps_provision_compute", "lsmps_provision_temporary_velocity",
    "lsmps_provision_temporary_wall_velocity", "lsmps_ppe_assemble", "lsmps_ppe_solve",
    "lsmps_default_solver_options", "lsmps_correction_apply", "lsmps_clamp_particle_pressure",
    "lsmps_compute_pressure_gradient", "lsmps_compute_particle_shifting",
    "lsmps_update_velocity_and_position", "lsmps_apply_anti_penetration",
    "lsmps_smooth_particle_velocity", "lsmps_stepper_create", "lsmps_stepper_destroy",
    "lsmps_stepper_set_config", "lsmps_stepper_set_solver", "lsmps_stepper_bytes",
    "lsmps_stepper_step", "lsmps_stepper_stage_times", "lsmps_stepper_neighbor_lists",
    "lsmps_stepper_ppe_workspace", "lsmps_stepper_temporary_velocity",
    "lsmps_stepper_correction_vectors", "lsmps_stepper_moments", "lsmps_context_kernel_stats",
    "lsmps_stepper_set_segments", "lsmps_stepper_set_mode", "lsmps_adjoint_step",
    "lsmps_time_limits_from_config", "lsmps_time_limits_validate", "lsmps_time_manager_create",
    "lsmps_time_manager_destroy", "lsmps_time_manager_reset", "lsmps_time_manager_state",
    "lsmps_time_manager_next_time_step", "lsmps_time_manager_advance",
    "lsmps_time_manager_mark_initial_output", "lsmps_max_velocity", "lsmps_stepper_run",
    "lsmps_stepper_refresh_geometry", "lsmps_default_config_files", "lsmps_config_validate",
    "lsmps_config_load", "lsmps_config_save", "lsmps_read_fluid_csv", "lsmps_read_wall_csv",
    "lsmps_vtk_path", "lsmps_write_vtk", "lsmps_compute_virtual_light_diagnostics",
    "lsmps_ppe_assemble_views", "lsmps_compute_pressure_gradient_view",
    "lsmps_build_neighbor_lists_noalloc", "lsmps_context_lanes", "lsmps_context_force_lanes",
    "lsmps_dense_inverse", "lsmps_ppe_spmv",
]

_LIB = None


def load(path=LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build the HIP library first "
                           "(python __graft_entry__.py or lsmps-3d_amd/build.py)")
    lib = C.CDLL(path)
    lib.lsmps_last_error.restype = C.c_char_p
    lib.lsmps_version.restype = C.c_char_p
    lib.lsmps_moment_bytes.restype = C.c_size_t
    lib.lsmps_stepper_bytes.restype = C.c_size_t
    lib.lsmps_stepper_moments.restype = C.c_void_p
    lib.lsmps_stepper_moments.argtypes = [C.c_void_p]
    lib.lsmps_context_precision.restype = C.c_int
    lib.lsmps_moment_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_uint64]
    lib.lsmps_context_create.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_void_p]
    lib.lsmps_derive_cell_grid.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.c_void_p]
    lib.lsmps_context_kernel_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_int]
    lib.lsmps_reference_number_density.argtypes = [C.c_double, C.c_double, C.c_int32,
                                                   C.POINTER(C.c_double)]
    lib.lsmps_time_manager_advance.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    lib.lsmps_time_manager_next_time_step.argtypes = [C.c_void_p, C.c_double, C.c_void_p,
                                                      C.c_void_p]
    lib.lsmps_stepper_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int64, abi.OutputCallback, C.c_void_p, C.c_void_p]
    _LIB = lib
    return lib


class LsmpsError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc, what=""):
    if rc != abi.LSMPS_OK:
        msg = load().lsmps_last_error().decode()
        raise LsmpsError(rc, f"{what}: {msg}" if what else msg)
    return rc


# ---- host helpers (no GPU needed) ------------------------------------------

def reference_number_density(l0, re, kernel=abi.WEIGHT_LINEAR):
    out = C.c_double()
    check(load().lsmps_reference_number_density(l0, re, kernel, C.byref(out)))
    return out.value


def derive_cell_grid(config, fluid, walls):
    f = [np.ascontiguousarray(fluid[k], np.float64) for k in ("x", "y", "z")]
    w = [np.ascontiguousarray(walls[k], np.float64) for k in ("x", "y", "z")]
    cfg = config.copy()
    check(load().lsmps_derive_cell_grid(*(a.ctypes.data for a in f), f[0].shape[0],
                                        *(a.ctypes.data for a in w), w[0].shape[0],
                                        C.byref(cfg)))
    return cfg


# ---- on-disk formats (host code) ---------------------------------------------

def default_config_files():
    f = abi.ConfigFiles()
    check(load().lsmps_default_config_files(C.byref(f)))
    return f


def validate_config(config, files=None):
    check(load().lsmps_config_validate(C.byref(config), C.byref(files) if files else None),
          "config_validate")


def load_config(path):
    """load_simulation_config: (lsmps_config, ConfigFiles)."""
    cfg, files = abi.Config(), abi.ConfigFiles()
    check(load().lsmps_config_load(os.fsencode(path), C.byref(cfg), C.byref(files)),
          "config_load")
    return cfg, files


def save_config(path, config, files=None):
    check(load().lsmps_config_save(os.fsencode(path), C.byref(config),
                                   C.byref(files) if files else None), "config_save")


def read_fluid_csv(path):
    """FileManager::load_fluid_particles into numpy arrays (x..vz, pressure, surface_type)."""
    lib = load()
    n = C.c_size_t()
    check(lib.lsmps_read_fluid_csv(os.fsencode(path), None, C.byref(n)), "read_fluid_csv")
    out = {k: np.zeros(n.value) for k in ("x", "y", "z", "vx", "vy", "vz", "pressure")}
    out["surface_type"] = np.full(n.value, -1, np.int32)
    soa = abi.FluidSoA(n.value, *(abi.dptr(out[k].ctypes.data) for k in
                                  ("x", "y", "z", "vx", "vy", "vz", "pressure")),
                       abi.iptr(out["surface_type"].ctypes.data))
    check(lib.lsmps_read_fluid_csv(os.fsencode(path), C.byref(soa), C.byref(n)), "read_fluid_csv")
    return out


def read_wall_csv(path):
    lib = load()
    n = C.c_size_t()
    check(lib.lsmps_read_wall_csv(os.fsencode(path), None, C.byref(n)), "read_wall_csv")
    keys = ("x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z")
    out = {k: np.zeros(n.value) for k in keys}
    soa = abi.WallSoA(n.value, *(abi.dptr(out[k].ctypes.data) for k in keys))
    check(lib.lsmps_read_wall_csv(os.fsencode(path), C.byref(soa), C.byref(n)), "read_wall_csv")
    return out


def vtk_path(directory, prefix, step):
    buf = C.create_string_buffer(4096)
    check(load().lsmps_vtk_path(os.fsencode(directory), prefix.encode(), C.c_size_t(step), buf,
                                C.c_size_t(4096)), "vtk_path")
    return buf.value.decode()


def write_vtk(directory, prefix, step, x, y, z, vectors=(), real_scalars=(), int_scalars=(),
              point_fields=True):
    """LegacyVtkWriter::write of host arrays; fields as (name, arrays) pairs."""
    keep = []
    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a
    xs, ys, zs = (arr(a, np.float64) for a in (x, y, z))
    fields = []
    for name, (fx, fy, fz) in vectors:
        a, b, c = (arr(v, np.float64) for v in (fx, fy, fz))
        # the library checks the count against the particles; a short component fails it
        fields.append(abi.VtkField(name.encode(), abi.VTK_VECTOR, min(len(a), len(b), len(c)),
                                   abi.dptr(a.ctypes.data),
                                   abi.dptr(b.ctypes.data), abi.dptr(c.ctypes.data), abi.iptr(0)))
    for name, v in real_scalars:
        a = arr(v, np.float64)
        fields.append(abi.VtkField(name.encode(), abi.VTK_REAL_SCALAR, len(a),
                                   abi.dptr(a.ctypes.data), abi.dptr(0), abi.dptr(0), abi.iptr(0)))
    for name, v in int_scalars:
        a = arr(v, np.int32)
        fields.append(abi.VtkField(name.encode(), abi.VTK_INT_SCALAR, len(a), abi.dptr(0),
                                   abi.dptr(0), abi.dptr(0), abi.iptr(a.ctypes.data)))
    farr = (abi.VtkField * max(len(fields), 1))(*fields)
    check(load().lsmps_write_vtk(os.fsencode(directory), prefix.encode(), 1 if point_fields else 0,
                                 C.c_size_t(step), C.c_size_t(len(xs)),
                                 abi.dptr(xs.ctypes.data), abi.dptr(ys.ctypes.data),
                                 abi.dptr(zs.ctypes.data), farr, len(fields)), "write_vtk")
    return vtk_path(directory, prefix, step)


class TimeManager:
    """SimulationTimeManager over the C-ABI (host-side clock, no GPU needed)."""

    def __init__(self, config=None, limits=None):
        self.lib = load()
        if limits is None:
            limits = abi.TimeLimits()
            check(self.lib.lsmps_time_limits_from_config(C.byref(config), C.byref(limits)),
                  "time_limits_from_config")
        self.limits = limits
        self.handle = C.c_void_p()
        check(self.lib.lsmps_time_manager_create(C.byref(limits), C.byref(self.handle)),
              "time_manager_create")

    @staticmethod
    def validate(limits):
        check(load().lsmps_time_limits_validate(C.byref(limits)), "time_limits_validate")

    def advance(self, max_velocity):
        st = abi.TimeStatus()
        check(self.lib.lsmps_time_manager_advance(self.handle, float(max_velocity), C.byref(st)))
        return st.as_dict()

    def mark_initial_output(self):
        st = abi.TimeStatus()
        check(self.lib.lsmps_time_manager_mark_initial_output(self.handle, C.byref(st)))
        return st.as_dict()

    def next_time_step(self, max_velocity):
        dt, cdt = C.c_double(), C.c_double()
        check(self.lib.lsmps_time_manager_next_time_step(self.handle, float(max_velocity),
                                                         C.byref(dt), C.byref(cdt)))
        return dt.value, cdt.value

    def state(self):
        t, dt = C.c_double(), C.c_double()
        si, oi = C.c_size_t(), C.c_size_t()
        due, fin = C.c_int32(), C.c_int32()
        check(self.lib.lsmps_time_manager_state(self.handle, C.byref(t), C.byref(dt), C.byref(si),
                                                C.byref(oi), C.byref(due), C.byref(fin)))
        return {"current_time": t.value, "current_time_step": dt.value, "step_index": si.value,
                "output_index": oi.value, "output_due": bool(due.value),
                "finished": bool(fin.value)}

    def close(self):
        if self.handle:
            self.lib.lsmps_time_manager_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- device plumbing --------------------------------------------------------

def _torch():
    import torch
    return torch


def to_device(a, device="cuda:0"):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def ptr(t):
    return t.data_ptr() if t is not None else 0


class DeviceState:
    """Fluid and wall SoA buffers of a case on the GPU."""

    FLUID_KEYS = ("x", "y", "z", "vx", "vy", "vz", "pressure")
    WALL_KEYS = ("x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z")

    def __init__(self, case, device="cuda:0"):
        self.device = device
        self.n_fluid = case.n_fluid
        self.n_wall = case.n_wall
        self.fluid = {k: to_device(case.fluid[k].astype(np.float64), device) for k in self.FLUID_KEYS}
        self.fluid["surface_type"] = to_device(case.fluid["surface_type"].astype(np.int32), device)
        self.walls = {k: to_device(case.walls[k].astype(np.float64), device) for k in self.WALL_KEYS}

    def fluid_soa(self):
        f = self.fluid
        return abi.FluidSoA(self.n_fluid, *(abi.dptr(ptr(f[k])) for k in self.FLUID_KEYS),
                            abi.iptr(ptr(f["surface_type"])))

    def wall_soa(self):
        w = self.walls
        return abi.WallSoA(self.n_wall, *(abi.dptr(ptr(w[k])) for k in self.WALL_KEYS))

    def snapshot(self):
        return {k: v.clone() for k, v in self.fluid.items()}

    def restore(self, snap):
        for k, v in snap.items():
            self.fluid[k].copy_(v)

    def fluid_host(self):
        return {k: v.cpu().numpy() for k, v in self.fluid.items()}


class Context:
    def __init__(self, precision=abi.PRECISION_FP64, device=0, flags=0, stream=None):
        torch = _torch()
        self.lib = load()
        self.handle = C.c_void_p()
        self.precision = precision
        check(self.lib.lsmps_context_create(device, precision, flags, C.byref(self.handle)),
              "context_create")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(self.lib.lsmps_context_set_stream(self.handle, C.c_void_p(s)), "set_stream")

    def close(self):
        if self.handle:
            self.lib.lsmps_context_destroy(self.handle)
            self.handle = C.c_void_p()

    def lanes(self):
        """{kernel: lanes of its latest launch}, and the number of tuned keys."""
        out = {}
        keys = C.c_int64()
        for name, k in abi.TUNED_KERNELS.items():
            v = C.c_int32()
            check(self.lib.lsmps_context_lanes(self.handle, k, C.byref(v), C.byref(keys)),
                  "context_lanes")
            out[name] = v.value
        return out, keys.value

    def force_lanes(self, kernel, lanes):
        check(self.lib.lsmps_context_force_lanes(self.handle, abi.TUNED_KERNELS[kernel], lanes),
              "context_force_lanes")

    def ppe_spmv(self, row_offsets, col_indices, values, x, precision=None):
        """y = A x on the GPU through the PPE solve's SpMV layout and kernel;
        precision LSMPS_PRECISION_MIXED rounds the stored values to FP32.  The
        CSR arrays and x are torch tensors (int32, int32, float64, float64)."""
        import torch
        rows = row_offsets.to(torch.int32).contiguous()
        cols = col_indices.to(torch.int32).contiguous()
        vals = values.to(torch.float64).contiguous()
        xv = x.to(torch.float64).contiguous()
        n = rows.shape[0] - 1
        assert rows.is_cuda and cols.is_cuda and vals.is_cuda and xv.is_cuda and xv.shape[0] == n
        y = torch.empty(n, dtype=torch.float64, device=xv.device)
        csr = abi.Csr(n, n, vals.shape[0], abi.iptr(rows.data_ptr()), abi.iptr(cols.data_ptr()),
                      abi.dptr(vals.data_ptr()))
        prec = self.precision if precision is None else precision
        check(self.lib.lsmps_ppe_spmv(self.handle, C.byref(csr), abi.dptr(xv.data_ptr()),
                                      abi.dptr(y.data_ptr()), prec), "ppe_spmv")
        return y

    def dense_inverse(self, matrix):
        """Inverse of a square float64 torch tensor on the GPU (row-major), by
        the coarse-space Gauss-Jordan kernels."""
        import torch
        a = matrix.to(torch.float64).contiguous()
        n = a.shape[0]
        assert a.dim() == 2 and a.shape[1] == n and a.is_cuda
        out = torch.empty_like(a)
        check(self.lib.lsmps_dense_inverse(self.handle, n, abi.dptr(a.data_ptr()),
                                           abi.dptr(out.data_ptr())), "dense_inverse")
        return out  # the entry point waits for its stream

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stepper:
    def __init__(self, ctx, state, config, max_neighbors=256, solver=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.state = state
        self.config = config.copy()
        self.max_neighbors = max_neighbors
        self.handle = C.c_void_p()
        check(self.lib.lsmps_stepper_create(ctx.handle, C.c_size_t(state.n_fluid),
                                            C.c_size_t(state.n_wall), C.c_size_t(max_neighbors),
                                            C.byref(self.config), C.byref(self.handle)),
              "stepper_create")
        if solver is not None:
            self.set_solver(solver)

    def set_segments(self, fluid_offsets, wall_offsets, grids):
        """Batched tanks: offsets (len T+1) into the fluid / wall SoAs and T cell grids."""
        fo = np.ascontiguousarray(fluid_offsets, np.int64)
        wo = np.ascontiguousarray(wall_offsets, np.int64)
        t = fo.shape[0] - 1
        arr = (abi.CellGrid * max(t, 1))(*grids) if t > 0 else None
        check(self.lib.lsmps_stepper_set_segments(
            self.handle, C.c_int32(t), fo.ctypes.data_as(C.c_void_p),
            wo.ctypes.data_as(C.c_void_p), arr), "set_segments")

    def set_solver(self, options):
        check(self.lib.lsmps_stepper_set_solver(self.handle, C.byref(options)), "set_solver")

    def run(self, clock, max_steps=-1, on_output=None):
        """lsmps_stepper_run: the driver loop until the clock's final time (or
        max_steps); on_output(status_dict) is called for every output state and
        may return True to stop.  Returns (last status, return code)."""
        def cb(status, user):
            if on_output is None:
                return 0
            return 1 if on_output(status.contents.as_dict()) else 0
        fn = abi.OutputCallback(cb)
        last = abi.TimeStatus()
        fs = self.state.fluid_soa()
        ws = self.state.wall_soa()
        rc = self.lib.lsmps_stepper_run(self.handle, C.byref(fs), C.byref(ws), clock.handle,
                                        C.c_int64(max_steps), fn, None, C.byref(last))
        if rc not in (abi.LSMPS_OK, abi.LSMPS_ERR_NOT_CONVERGED):
            check(rc, "stepper_run")
        return last.as_dict(), rc

    def refresh_geometry(self):
        fs = self.state.fluid_soa()
        ws = self.state.wall_soa()
        check(self.lib.lsmps_stepper_refresh_geometry(self.handle, C.byref(fs), C.byref(ws)),
              "refresh_geometry")

    def set_config(self, config):
        self.config = config.copy()
        check(self.lib.lsmps_stepper_set_config(self.handle, C.byref(self.config)), "set_config")

    def set_mode(self, mode):
        """abi.STEP_RECORD (keep what the adjoint needs) | abi.STEP_FROZEN_GEOMETRY."""
        check(self.lib.lsmps_stepper_set_mode(self.handle, C.c_uint32(mode)), "set_mode")

    def adjoint(self, grad_v, grad_x=None, want=("position", "velocity", "pressure"),
                raise_on_nonconvergence=True):
        """Reverse mode of the last recorded step.  grad_v / grad_x: (3, N) float64
        device tensors (dL/dv_new, dL/dx_new).  Returns device tensors for the
        requested gradients, the host dL/dg and the A^T solve report."""
        torch = _torch()
        n = self.state.n_fluid
        dev = self.state.device
        gv = grad_v.contiguous()
        assert gv.shape == (3, n) and gv.dtype == torch.float64 and gv.is_cuda
        gvs = abi.Vector3(n, *(abi.dptr(ptr(gv[k])) for k in range(3)))
        gxs = None
        if grad_x is not None:
            gx = grad_x.contiguous()
            assert gx.shape == (3, n) and gx.dtype == torch.float64 and gx.is_cuda
            gxs = abi.Vector3(n, *(abi.dptr(ptr(gx[k])) for k in range(3)))
        out = {}
        g = abi.AdjointGrads()
        if "position" in want:
            out["position"] = torch.empty((3, n), dtype=torch.float64, device=dev)
            g.position = abi.Vector3(n, *(abi.dptr(ptr(out["position"][k])) for k in range(3)))
        if "velocity" in want:
            out["velocity"] = torch.empty((3, n), dtype=torch.float64, device=dev)
            g.velocity = abi.Vector3(n, *(abi.dptr(ptr(out["velocity"][k])) for k in range(3)))
        if "pressure" in want:
            out["pressure"] = torch.empty(n, dtype=torch.float64, device=dev)
            g.pressure = abi.dptr(ptr(out["pressure"]))
        fs = self.state.fluid_soa()
        ws = self.state.wall_soa()
        rc = self.lib.lsmps_adjoint_step(self.handle, C.byref(fs), C.byref(ws), C.byref(gvs),
                                         C.byref(gxs) if gxs is not None else None, C.byref(g))
        if rc != abi.LSMPS_OK and (rc != abi.LSMPS_ERR_NOT_CONVERGED or raise_on_nonconvergence):
            check(rc, "adjoint_step")
        out["gravity"] = np.array(list(g.gravity))
        out["report"] = g.transpose_report.as_dict()
        return out

    def step(self, raise_on_nonconvergence=True):
        rep = abi.SolveReport()
        fs = self.state.fluid_soa()
        ws = self.state.wall_soa()
        rc = self.lib.lsmps_stepper_step(self.handle, C.byref(fs), C.byref(ws), C.byref(rep))
        if rc != abi.LSMPS_OK and (rc != abi.LSMPS_ERR_NOT_CONVERGED or raise_on_nonconvergence):
            check(rc, "stepper_step")
        return rep.as_dict()

    def stage_times(self):
        ms = (C.c_double * 7)()
        check(self.lib.lsmps_stepper_stage_times(self.handle, ms, 7))
        return list(ms)

    def bytes(self):
        return int(self.lib.lsmps_stepper_bytes(self.handle))

    def neighbor_lists(self):
        """(fluid_offsets, fluid_indices, wall_offsets, wall_indices) as numpy arrays."""
        fn, wn = abi.NeighborList(), abi.NeighborList()
        check(self.lib.lsmps_stepper_neighbor_lists(self.handle, C.byref(fn), C.byref(wn)))
        n = self.state.n_fluid
        f_off = device_to_numpy(fn.offsets, n + 1, np.int32)
        w_off = device_to_numpy(wn.offsets, n + 1, np.int32)
        f_idx = device_to_numpy(fn.indices, int(f_off[-1]), np.int32)
        w_idx = device_to_numpy(wn.indices, int(w_off[-1]), np.int32)
        return f_off, f_idx, w_off, w_idx

    def ppe(self):
        ws = abi.PpeWorkspace()
        check(self.lib.lsmps_stepper_ppe_workspace(self.handle, C.byref(ws)))
        n = self.state.n_fluid
        rows = device_to_numpy(ws.matrix.row_offsets, n + 1, np.int32)
        used = int(rows[-1])
        return {
            "row_offsets": rows,
            "col_indices": device_to_numpy(ws.matrix.col_indices, used, np.int32),
            "values": device_to_numpy(ws.matrix.values, used, np.float64),
            "nnz": ws.matrix.nnz,
            "rhs": device_to_numpy(ws.rhs, n, np.float64),
            "pressure": device_to_numpy(ws.pressure, n, np.float64),
            "divergence": device_to_numpy(ws.divergence, n, np.float64),
            "pressure_laplacian": device_to_numpy(ws.pressure_laplacian, n, np.float64),
        }

    def temporary_velocity(self):
        f, w = abi.FluidSoA(), abi.WallSoA()
        check(self.lib.lsmps_stepper_temporary_velocity(self.handle, C.byref(f), C.byref(w)))
        n, nw = self.state.n_fluid, self.state.n_wall
        return {"vx": device_to_numpy(f.vx, n, np.float64), "vy": device_to_numpy(f.vy, n, np.float64),
                "vz": device_to_numpy(f.vz, n, np.float64),
                "wall_vx": device_to_numpy(w.vx, nw, np.float64),
                "wall_vy": device_to_numpy(w.vy, nw, np.float64),
                "wall_vz": device_to_numpy(w.vz, nw, np.float64)}

    def correction_vectors(self):
        g, p = abi.Vector3(), abi.Vector3()
        check(self.lib.lsmps_stepper_correction_vectors(self.handle, C.byref(g), C.byref(p)))
        n = self.state.n_fluid
        return ([device_to_numpy(a, n, np.float64) for a in (g.x, g.y, g.z)],
                [device_to_numpy(a, n, np.float64) for a in (p.x, p.y, p.z)])

    def moments_handle(self):
        return C.c_void_p(self.lib.lsmps_stepper_moments(self.handle))

    def close(self):
        if self.handle:
            self.lib.lsmps_stepper_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def max_velocity(ctx, state):
    """lsmps_max_velocity: largest |v| of the fluid on the device."""
    out = C.c_double()
    fs = state.fluid_soa()
    check(ctx.lib.lsmps_max_velocity(ctx.handle, C.byref(fs), C.byref(out)), "max_velocity")
    return out.value


def surface_on_device(case, config=None, max_neighbors=256, virtual_light=True,
                      device="cuda:0"):
    """Neighbor lists, surface labels with every diagnostic and (optionally) the
    virtual-light counts of `case` on the GPU through the module entries
    (build_neighbor_lists, classify_surface_particles,
    compute_virtual_light_diagnostics); numpy outputs."""
    torch = _torch()
    cfg = config or case.config
    ctx = Context(abi.PRECISION_FP64)
    state = DeviceState(case, device)
    n, nw = case.n_fluid, case.n_wall
    cells = int(cfg.cell_dims[0]) * int(cfg.cell_dims[1]) * int(cfg.cell_dims[2])

    def i32(k):
        return torch.zeros(max(k, 1), dtype=torch.int32, device=device)

    def f64(k):
        return torch.zeros(max(k, 1), dtype=torch.float64, device=device)

    fc = [i32(n), i32(n), i32(cells), i32(cells)]
    wc = [i32(nw), i32(nw), i32(cells), i32(cells)]
    cap = n * max_neighbors
    fo, fi, wo, wi = i32(n + 1), i32(cap), i32(n + 1), i32(cap)
    fcl = abi.CellList(n, cells, *(abi.iptr(ptr(t)) for t in fc))
    wcl = abi.CellList(nw, cells, *(abi.iptr(ptr(t)) for t in wc))
    fnl = abi.NeighborList(n, cap, abi.iptr(ptr(fo)), abi.iptr(ptr(fi)))
    wnl = abi.NeighborList(n, cap, abi.iptr(ptr(wo)), abi.iptr(ptr(wi)))
    fs, ws = state.fluid_soa(), state.wall_soa()
    lib = ctx.lib
    check(lib.lsmps_build_neighbor_lists(ctx.handle, C.byref(fs), C.byref(ws), C.byref(cfg),
                                         C.byref(fcl), C.byref(wcl), C.byref(fnl), C.byref(wnl)),
          "build_neighbor_lists")
    ints = {k: i32(n) for k in ("fluid_neighbor_count", "wall_neighbor_count")}
    reals = {k: f64(n) for k in ("number_density", "number_density_ratio", "anisotropy",
                                 "air_open_ratio", "air_anisotropy", "surface_normal_x",
                                 "surface_normal_y", "surface_normal_z")}
    diag = abi.SurfaceDiagnostics(abi.iptr(ptr(ints["fluid_neighbor_count"])),
                                  abi.iptr(ptr(ints["wall_neighbor_count"])),
                                  *(abi.dptr(ptr(reals[k])) for k in (
                                      "number_density", "number_density_ratio", "anisotropy",
                                      "air_open_ratio", "air_anisotropy", "surface_normal_x",
                                      "surface_normal_y", "surface_normal_z")))
    check(lib.lsmps_classify_surface(ctx.handle, C.byref(fs), C.byref(ws), C.byref(fnl),
                                     C.byref(wnl), C.byref(cfg), C.byref(diag)),
          "classify_surface")
    out = {}
    if virtual_light:
        cnt, frac = i32(n), f64(n)
        vl = abi.VirtualLightDiagnostics(abi.iptr(ptr(cnt)), abi.dptr(ptr(frac)))
        check(lib.lsmps_compute_virtual_light_diagnostics(
            ctx.handle, C.byref(fs), C.byref(ws), C.byref(fnl), C.byref(wnl), C.byref(cfg),
            C.byref(vl)), "compute_virtual_light_diagnostics")
        out["open_direction_count"] = cnt[:n].cpu().numpy()
        out["open_fraction"] = frac[:n].cpu().numpy()
    torch.cuda.synchronize()
    f_off = fo[:n + 1].cpu().numpy()
    w_off = wo[:n + 1].cpu().numpy()
    out["fluid_offsets"] = f_off
    out["fluid_indices"] = fi[:int(f_off[-1])].cpu().numpy()
    out["wall_offsets"] = w_off
    out["wall_indices"] = wi[:int(w_off[-1])].cpu().numpy()
    out["surface_type"] = state.fluid["surface_type"][:n].cpu().numpy()
    for k, t in list(ints.items()) + list(reals.items()):
        out[k] = t[:n].cpu().numpy()
    ctx.close()
    return out


def device_to_numpy(p, n, dtype):
    """Copy n elements from a device pointer (ctypes pointer or int) to a numpy array."""
    torch = _torch()
    n = int(n)
    addr = C.cast(p, C.c_void_p).value if not isinstance(p, int) else p
    out = np.empty(n, dtype=dtype)
    if n == 0:
        return out
    hip = _hip_runtime()
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(addr), C.c_size_t(out.nbytes), 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
    return out


_HIP = None


def _hip_runtime():
    global _HIP
    if _HIP is None:
        # by soname, so the runtime torch already loaded is reused
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7"):
            try:
                _HIP = C.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64.so not found")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _HIP


def default_solver(precision=abi.PRECISION_FP64):
    o = abi.default_solver_options()
    return o


def run_step_on_device(case, precision=abi.PRECISION_FP64, solver=None, max_neighbors=256,
                       device="cuda:0", flags=0, lanes=None):
    """One fused step of `case` on the GPU; returns every inspectable output as numpy.
    `lanes` ({tuned kernel name: lanes per particle}) pins the kernels' shapes."""
    ctx = Context(precision, flags=flags)
    for k, v in (lanes or {}).items():
        ctx.force_lanes(k, v)
    state = DeviceState(case, device)
    st = Stepper(ctx, state, case.config, max_neighbors, solver or abi.default_solver_options())
    if "fluid_offsets" in case.extra:  # batched tanks
        st.set_segments(case.extra["fluid_offsets"], case.extra["wall_offsets"],
                        case.extra["grids"])
    report = st.step(raise_on_nonconvergence=False)
    f_off, f_idx, w_off, w_idx = st.neighbor_lists()
    ppe = st.ppe()
    grad, ps = st.correction_vectors()
    out = {
        "fluid_offsets": f_off, "fluid_indices": f_idx, "wall_offsets": w_off,
        "wall_indices": w_idx, "ppe": ppe, "pressure": state.fluid["pressure"].cpu().numpy(),
        "surface_type": state.fluid["surface_type"].cpu().numpy(), "fluid": state.fluid_host(),
        "provision": st.temporary_velocity(), "grad": grad, "ps": ps, "report": report,
        "stage_ms": st.stage_times(),
    }
    out["lanes"], out["tuned_keys"] = ctx.lanes()
    st.close()
    ctx.close()
    return out
