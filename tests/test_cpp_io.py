"""The header-only C++ layer's config / FileManager / LegacyVtkWriter replay the
sequences of the reference's tests/unit/smoke_config.cpp, smoke_file_manager.cpp
and smoke_vtk_writer.cpp with their inputs and known answers (host code only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cpp_vec(v):
    return "{" + ", ".join(repr(float(x)) for x in v) + "}"


def _cpp_ivec(v):
    return "{" + ", ".join(str(int(x)) for x in v) + "}"


def test_cpp_io_known_answers(golden, tmp_path):
    cf, fm, vw = golden["config"], golden["file_manager"], golden["vtk_writer"]
    (tmp_path / "smoke.ini").write_text(cf["ini"])
    (tmp_path / "bad.ini").write_text(cf["bad_ini"])
    (tmp_path / "fluid.csv").write_text(fm["fluid_csv"])
    (tmp_path / "wall.csv").write_text(fm["wall_csv"])
    (tmp_path / "bad.csv").write_text(fm["bad_csv"])
    (tmp_path / "extra.csv").write_text(fm["extra_column_csv"])
    ex = cf["expected"]
    checks = " ||\n      ".join(f"!near(c.{k}, {v!r})" for k, v in ex.items())
    p = vw["points"]
    vec_adds = "\n  ".join(
        f"pf.add_vector(\"{k}\", {_cpp_vec(v[0])}, {_cpp_vec(v[1])}, {_cpp_vec(v[2])});"
        for k, v in vw["vectors"].items())
    real_adds = "\n  ".join(f"pf.add_real_scalar(\"{k}\", {_cpp_vec(v)});"
                            for k, v in vw["real_scalars"].items())
    int_adds = "\n  ".join(f"pf.add_int_scalar(\"{k}\", {_cpp_ivec(v)});"
                           for k, v in vw["int_scalars"].items())
    tokens = " && ".join(f"contains(vtk, \"{t}\")" for t in vw["expected_tokens"])
    src = f"""
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include "lsmps3d_amd.hpp"
namespace fs = std::filesystem;
static bool near(double a, double b) {{ return std::abs(a - b) <= {cf['tolerance']!r}; }}
static bool contains(const fs::path& p, const char* t) {{
  std::ifstream in(p); std::stringstream s; s << in.rdbuf(); return s.str().find(t) != std::string::npos;
}}
int main(int argc, char** argv) {{
  const fs::path root = argv[1];
  // smoke_config.cpp
  const lsmps3d::SimulationConfig d = lsmps3d::default_simulation_config();
  if (!near(d.support_radius_ratio(), {cf['default_ratios']['support_radius_ratio']!r})) return 1;
  const lsmps3d::SimulationConfig c = lsmps3d::load_simulation_config(root / "smoke.ini");
  if ({checks}) return 2;
  if (c.vtk_write_point_fields != {str(bool(cf['expected_files']['vtk_write_point_fields'])).lower()} ||
      c.amgx_print_solve_stats != {str(bool(cf['expected_files']['amgx_print_solve_stats'])).lower()} ||
      c.vtk_file_prefix != "{cf['expected_files']['vtk_file_prefix']}" ||
      c.amgx_config_path != "{cf['expected_files']['amgx_config_path']}") return 3;
  lsmps3d::save_simulation_config(c, root / "saved.ini");
  const auto r = lsmps3d::load_simulation_config(root / "saved.ini");
  if (!near(r.support_radius, c.support_radius) || r.output_directory != c.output_directory ||
      r.vtk_file_prefix != c.vtk_file_prefix) return 4;
  auto bad = d; bad.support_radius = -1.0;
  try {{ lsmps3d::validate_simulation_config(bad); return 5; }} catch (const std::invalid_argument&) {{}}
  bad = d; bad.min_time_step = 0.01; bad.max_time_step = 0.001;
  try {{ lsmps3d::validate_simulation_config(bad); return 6; }} catch (const std::invalid_argument&) {{}}
  try {{ (void)lsmps3d::load_simulation_config(root / "bad.ini"); return 7; }} catch (const std::invalid_argument&) {{}}
  // smoke_file_manager.cpp
  lsmps3d::SimulationConfig fc;
  fc.fluid_particle_file = root / "fluid.csv";
  fc.wall_particle_file = root / "wall.csv";
  fc.output_directory = root / "vtk";
  fc.vtk_file_prefix = "{fm['vtk_prefix']}";
  const lsmps3d::FileManager manager(fc);
  const auto input = manager.load_particle_input();
  if (input.fluid.count() != {fm['expected']['fluid_count']} || input.walls.count() != {fm['expected']['wall_count']} ||
      !near(input.fluid.vx[1], {fm['expected']['fluid_vx'][1]!r}) || !near(input.fluid.pressure[1], 0.0) ||
      input.fluid.surface_type[0] != 0 || !near(input.walls.normal_y[0], {fm['expected']['wall_normal_y0']!r}) ||
      !near(input.walls.vz[1], {fm['expected']['wall_vz1']!r})) return 8;
  manager.save_config(fc, root / "fm.ini");
  const auto fr = manager.load_config(root / "fm.ini");
  if (fr.fluid_particle_file != fc.fluid_particle_file || fr.output_directory != fc.output_directory) return 9;
  lsmps3d::HostVtkPointFields ff;
  ff.add_vector("velocity", input.fluid.vx, input.fluid.vy, input.fluid.vz);
  ff.add_real_scalar("pressure", input.fluid.pressure);
  ff.add_int_scalar("surface_type", input.fluid.surface_type);
  manager.write_fluid_result({fm['vtk_step']}, input.fluid, ff);
  lsmps3d::HostVtkPointFields wf;
  wf.add_vector("normal", input.walls.normal_x, input.walls.normal_y, input.walls.normal_z);
  manager.write_wall_result({fm['vtk_step']}, input.walls, wf);
  if (!contains(fc.output_directory / "{fm['expected_files'][0]}", "VECTORS velocity float") ||
      !contains(fc.output_directory / "{fm['expected_files'][1]}", "VECTORS normal float")) return 10;
  try {{ (void)manager.load_fluid_particles(root / "bad.csv"); return 11; }} catch (const std::invalid_argument&) {{}}
  try {{ (void)manager.load_fluid_particles(root / "extra.csv"); return 12; }} catch (const std::invalid_argument&) {{}}
  try {{ (void)manager.load_fluid_particles(root / "missing.csv"); return 13; }} catch (const std::runtime_error&) {{}}
  // smoke_vtk_writer.cpp
  lsmps3d::SimulationConfig vc;
  vc.output_directory = root / "vtk2";
  vc.vtk_file_prefix = "{vw['prefix']}";
  const lsmps3d::LegacyVtkWriter writer(vc);
  const lsmps3d::HostParticleSnapshot pts{{{_cpp_vec(p['x'])}, {_cpp_vec(p['y'])}, {_cpp_vec(p['z'])}}};
  lsmps3d::HostVtkPointFields pf;
  {vec_adds}
  {real_adds}
  {int_adds}
  writer.write({vw['step']}, pts, pf);
  const fs::path vtk = writer.make_path({vw['step']});
  if (vtk.filename() != "{vw['expected_file']}" || !({tokens})) return 14;
  auto badf = pf;
  badf.vectors.front().z.pop_back();
  try {{ writer.write({vw['step']} + 1, pts, badf); return 15; }} catch (const std::invalid_argument&) {{}}
  std::puts("ok");
  return 0;
}}
"""
    cpp = tmp_path / "io.cpp"
    cpp.write_text(src)
    exe = tmp_path / "io"
    lib_dir = os.path.join(ROOT, "lsmps-3d_amd")
    res = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", str(cpp),
                          f"-I{os.path.join(ROOT, 'include')}", f"-L{lib_dir}", "-llsmps3d_amd",
                          f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    run = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert run.returncode == 0 and run.stdout.strip() == "ok", (run.returncode, run.stdout, run.stderr)
