"""The C-ABI without a GPU: the library loads and exports every function the
header declares, the ctypes mirror matches the C struct layouts (checked by a
C program compiled here with gcc), and the host-side entries agree with the
reference defaults and with independent computations."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lsmps3d_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s*(lsmps_\w+)\s*\(", text, re.M)))


def test_header_functions_exported(rt):
    lib = rt.load()
    names = _declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert len(rt.EXPORTS) == len(set(rt.EXPORTS))
    assert set(rt.EXPORTS) == set(names)


STRUCTS = ["lsmps_fluid_soa", "lsmps_wall_soa", "lsmps_neighbor_list", "lsmps_cell_grid",
           "lsmps_cell_list", "lsmps_surface_diagnostics", "lsmps_csr", "lsmps_ppe_workspace",
           "lsmps_vector3", "lsmps_moment_view", "lsmps_config", "lsmps_solver_options",
           "lsmps_solve_report", "lsmps_adjoint_grads", "lsmps_time_limits", "lsmps_time_status",
           "lsmps_config_files", "lsmps_vtk_field"]
MIRRORS = {"lsmps_fluid_soa": "FluidSoA", "lsmps_wall_soa": "WallSoA",
           "lsmps_neighbor_list": "NeighborList", "lsmps_cell_grid": "CellGrid",
           "lsmps_cell_list": "CellList", "lsmps_surface_diagnostics": "SurfaceDiagnostics",
           "lsmps_csr": "Csr", "lsmps_ppe_workspace": "PpeWorkspace", "lsmps_vector3": "Vector3",
           "lsmps_moment_view": "MomentView", "lsmps_config": "Config",
           "lsmps_solver_options": "SolverOptions", "lsmps_solve_report": "SolveReport",
           "lsmps_adjoint_grads": "AdjointGrads", "lsmps_time_limits": "TimeLimits",
           "lsmps_time_status": "TimeStatus", "lsmps_config_files": "ConfigFiles",
           "lsmps_vtk_field": "VtkField"}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """sizeof and every field offset of the header's structs, from gcc."""
    d = tmp_path_factory.mktemp("layout")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lsmps3d_amd.h"',
             'int main(void) {']
    from importlib import import_module
    abi = import_module("lsmps3d_amd.abi")
    for s in STRUCTS:
        lines.append(f'  printf("{s} size %zu\\n", sizeof({s}));')
        for fname, _ in getattr(abi, MIRRORS[s])._fields_:
            lines.append(f'  printf("{s} {fname} %zu\\n", offsetof({s}, {fname}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        s, f, v = line.split()
        table[(s, f)] = int(v)
    return table


@pytest.mark.parametrize("struct", STRUCTS)
def test_ctypes_mirror_matches_c_layout(abi, c_layout, struct):
    mirror = getattr(abi, MIRRORS[struct])
    assert C.sizeof(mirror) == c_layout[(struct, "size")]
    for fname, _ in mirror._fields_:
        assert getattr(mirror, fname).offset == c_layout[(struct, fname)], fname


def test_default_config_matches_reference(rt, abi):
    """SimulationConfig{} defaults (include/lsmps3d/core/config.hpp:12-82)."""
    lib = rt.load()
    c = abi.Config()
    assert lib.lsmps_default_config(C.byref(c)) == 0
    ref = abi.default_config()
    for fname, ftype in abi.Config._fields_:
        got, want = getattr(c, fname), getattr(ref, fname)
        if hasattr(got, "__len__"):
            assert list(got) == list(want), fname
        else:
            assert got == want, fname
    assert c.particle_spacing == 0.01 and c.support_radius == 0.031
    assert c.near_surface_radius == 0.015 and c.density == 1000.0
    assert list(c.gravity) == [0.0, 0.0, -9.81]
    assert c.splash_neighbor_threshold == 4 and c.include_wall_neighbors == 1
    assert c.number_density_ratio_threshold == 0.85 and c.air_open_ratio_threshold == 0.33
    assert c.air_anisotropy_threshold == 0.033 and c.lsmps_regularization == 1e-8


def test_default_solver_options(rt, abi):
    lib = rt.load()
    o = abi.SolverOptions()
    assert lib.lsmps_default_solver_options(C.byref(o)) == 0
    ref = abi.default_solver_options()
    for fname, _ in abi.SolverOptions._fields_:
        assert getattr(o, fname) == getattr(ref, fname), fname
    assert o.method == abi.SOLVER_IDRS and o.tolerance == 1e-10
    assert o.preconditioner == abi.PRECOND_SCHWARZ_COARSE and o.coarse_cells == 4
    assert o.clamp_negative == 1


def _lattice_density(l0, re, quadratic):
    """Sum of w(|k| l0) over the cubic lattice points k != 0 within re."""
    m = int(math.ceil(re / l0))
    k = np.arange(-m, m + 1)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    r = l0 * np.sqrt(kx ** 2 + ky ** 2 + kz ** 2).ravel()
    r = r[(r > 0) & (r < re)]
    w = 1.0 - r / re
    return float(np.sum(w * w if quadratic else w))


@pytest.mark.parametrize("kernel", [0, 1])
def test_reference_number_density(rt, abi, oracle, kernel):
    """n0 of a uniform lattice: library == oracle == brute-force lattice sum;
    29.993 for the linear weight at re = 3.1 l0."""
    out = C.c_double()
    assert rt.load().lsmps_reference_number_density(0.01, 0.031, kernel, C.byref(out)) == 0
    brute = _lattice_density(0.01, 0.031, kernel == 1)
    assert abs(out.value - brute) <= 1e-12 * brute
    assert abs(out.value - oracle.reference_number_density(0.01, 0.031, kernel)) <= 1e-12 * brute
    if kernel == 0:
        assert round(out.value, 3) == 29.993


def test_derive_cell_grid_matches_case_builder(rt, cases):
    """derive_cell_grid_from_particles (simulation_driver.cu:83-108) against the
    case builder's restatement, on the dam-break and a jittered splash box."""
    for case in (cases.dam_break(block=(20, 10, 10), tank=(40, 10, 20), l0=0.01, seed=1),
                 cases.splash(n=16, l0=0.01, seed=3, top_layers=2)):
        cfg = case.config.copy()
        rt.derive_cell_grid(cfg, case.fluid, case.walls)
        assert list(cfg.cell_dims) == list(case.config.cell_dims)
        assert list(cfg.cell_origin) == list(case.config.cell_origin)
        assert cfg.cell_size == case.config.cell_size


def test_errors_are_status_codes(rt, abi):
    """No entry exits the process: bad arguments return a status and a message."""
    lib = rt.load()
    h = C.c_void_p()
    assert lib.lsmps_context_create(0, 7, 0, C.byref(h)) == abi.LSMPS_ERR_INVALID_ARGUMENT
    assert b"precision" in lib.lsmps_last_error()
    assert lib.lsmps_default_config(None) == abi.LSMPS_ERR_INVALID_ARGUMENT
    assert lib.lsmps_version().startswith(b"lsmps3d_amd")


def test_cpp_wrapper_compiles():
    """The header-only C++ front end compiles on its own (and builds the test program)."""
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "module_step")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    src = "#include \"lsmps3d_amd.hpp\"\nint main() { lsmps3d::SimulationConfig c; return c.cell_capacity() == 1 ? 0 : 1; }\n"
    res = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-fsyntax-only", "-x", "hip",
                          f"-I{os.path.join(ROOT, 'include')}", "--offload-arch=gfx950", "-"],
                         input=src, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_cpp_wrapper_clock_known_answers(golden, tmp_path):
    """lsmps3d::SimulationTimeManager (header-only layer over the C-ABI clock)
    replays the sequence of the reference's tests/unit/smoke_time_step.cpp with
    its known answers; host code only."""
    fx = golden["time_step"]
    c = fx["config"]
    a = fx["advances"]
    src = f"""
#include <cmath>
#include <cstdio>
#include "lsmps3d_amd.hpp"
static bool near(double x, double y) {{ return std::abs(x - y) <= {fx['tolerance']!r}; }}
int main() {{
  lsmps3d::SimulationConfig config;
  config.particle_spacing = {c['particle_spacing']!r};
  config.time_step = {c['time_step']!r};
  config.min_time_step = {c['min_time_step']!r};
  config.max_time_step = {c['max_time_step']!r};
  config.time_step_growth_factor = {c['time_step_growth_factor']!r};
  config.cfl = {c['cfl']!r};
  config.final_time = {c['final_time']!r};
  config.output_interval = {c['output_interval']!r};
  lsmps3d::SimulationTimeManager clock(config);
  if (!clock.output_due()) return 1;
  auto init = clock.mark_initial_output();
  if (!init.should_output || init.output_index != {fx['initial_output_index']} || clock.output_due()) return 2;
  auto s = clock.advance({a[0]['max_velocity']!r});
  if (!near(s.time_step, {a[0]['expected_dt']!r}) || s.should_output || s.reached_final_time) return 3;
  s = clock.advance({a[1]['max_velocity']!r});
  if (!near(s.time_step, {a[1]['expected_dt']!r}) || !s.should_output || s.output_index != {a[1]['output_index']}) return 4;
  s = clock.advance({a[2]['max_velocity']!r});
  if (!near(s.time_step, {a[2]['expected_dt']!r})) return 5;
  int guard = 0;
  while (!s.reached_final_time && guard < {fx['guard']}) {{ s = clock.advance({fx['then_max_velocity']!r}); ++guard; }}
  if (!s.reached_final_time || !near(clock.current_time(), {fx['expected_final_time']!r})) return 6;
  auto bad = lsmps3d::make_time_step_limits(config);
  bad.growth_factor = {fx['invalid_growth_factor']!r};
  try {{ lsmps3d::validate_time_step_limits(bad); return 7; }} catch (const std::invalid_argument&) {{}}
  std::puts("ok");
  return 0;
}}
"""
    cpp = tmp_path / "clock.cpp"
    cpp.write_text(src)
    exe = tmp_path / "clock"
    lib_dir = os.path.join(ROOT, "lsmps-3d_amd")
    res = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", str(cpp),
                          f"-I{os.path.join(ROOT, 'include')}", f"-L{lib_dir}", "-llsmps3d_amd",
                          f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True)
    assert run.returncode == 0 and run.stdout.strip() == "ok", (run.returncode, run.stdout, run.stderr)
