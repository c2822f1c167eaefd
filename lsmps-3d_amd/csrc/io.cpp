// On-disk formats of the reference, host code: the INI simulation config
// (src/core/config.cpp:122-431), particle CSV input (src/io/file_manager.cpp:
// 126-197) and legacy ASCII VTK output (src/io/vtk_writer.cpp:83-149).  The
// parsers accept what the reference accepts ('#' comments, blank lines,
// whitespace around tokens, [section] key=value, a header row before the first
// CSV record) and reject what it rejects (unknown keys, trailing characters,
// wrong column counts), with its messages; the writers produce the same text.
#include <algorithm>
#include <cctype>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "objects.hpp"

namespace lsmps {
namespace {

std::string trim(std::string_view v) {
  size_t b = 0, e = v.size();
  while (b < e && std::isspace(static_cast<unsigned char>(v[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(v[e - 1]))) --e;
  return std::string(v.substr(b, e - b));
}

[[noreturn]] void bad(const std::string& m) { raise(LSMPS_ERR_INVALID_ARGUMENT, m); }

double parse_double(const std::string& text, bool* ok) {
  *ok = false;
  if (text.empty()) return 0.0;
  try {
    size_t used = 0;
    const double v = std::stod(text, &used);
    *ok = used == text.size();
    return v;
  } catch (const std::exception&) {
    return 0.0;
  }
}

double real_value(std::string_view key, std::string_view value) {
  const std::string t = trim(value);
  bool ok = false;
  const double v = parse_double(t, &ok);
  if (!ok) bad("Invalid real value for '" + std::string(key) + "': '" + std::string(value) + "'");
  return v;
}

int32_t index_value(std::string_view key, std::string_view value) {
  const std::string t = trim(value);
  try {
    size_t used = 0;
    const long v = std::stol(t, &used, 10);
    if (used == t.size() && !t.empty()) return static_cast<int32_t>(v);
  } catch (const std::exception&) {
  }
  bad("Invalid integer value for '" + std::string(key) + "': '" + std::string(value) + "'");
}

bool bool_value(std::string_view key, std::string_view value) {
  std::string t = trim(value);
  std::transform(t.begin(), t.end(), t.begin(),
                 [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  if (t == "true" || t == "1" || t == "yes" || t == "on") return true;
  if (t == "false" || t == "0" || t == "no" || t == "off") return false;
  bad("Invalid boolean value for '" + std::string(key) + "': '" + std::string(value) + "'");
}

void set_text(char* dst, size_t cap, const std::string& v, const char* what) {
  if (v.size() + 1 > cap) bad(std::string("SimulationConfig field '") + what + "' is too long");
  std::memcpy(dst, v.c_str(), v.size() + 1);
}

// One key of one section (config.cpp:122-270): unknown keys and sections are errors.
void apply(lsmps_config& c, lsmps_config_files& f, const std::string& section,
           const std::string& key, const std::string& value, size_t line) {
  const std::string full = section.empty() ? key : section + "." + key;
  auto unknown = [&]() {
    bad("Unknown simulation config key '" + full + "' on line " + std::to_string(line));
  };
  struct RealKey {
    const char* name;
    double* dst;
  };
  auto real_keys = [&](std::initializer_list<RealKey> keys) {
    for (const RealKey& k : keys)
      if (key == k.name) {
        *k.dst = real_value(full, value);
        return true;
      }
    return false;
  };
  if (section == "geometry") {
    if (real_keys({{"particle_spacing", &c.particle_spacing},
                   {"support_radius", &c.support_radius},
                   {"near_surface_radius", &c.near_surface_radius},
                   {"cell_origin_x", &c.cell_origin[0]},
                   {"cell_origin_y", &c.cell_origin[1]},
                   {"cell_origin_z", &c.cell_origin[2]},
                   {"cell_size", &c.cell_size}}))
      return;
    if (key == "cell_dim_x") c.cell_dims[0] = index_value(full, value);
    else if (key == "cell_dim_y") c.cell_dims[1] = index_value(full, value);
    else if (key == "cell_dim_z") c.cell_dims[2] = index_value(full, value);
    else unknown();
  } else if (section == "simulation") {
    if (!real_keys({{"time_step", &c.time_step},
                    {"min_time_step", &c.min_time_step},
                    {"max_time_step", &c.max_time_step},
                    {"time_step_growth_factor", &c.time_step_growth_factor},
                    {"final_time", &c.final_time},
                    {"output_interval", &c.output_interval},
                    {"density", &c.density},
                    {"kinematic_viscosity", &c.kinematic_viscosity},
                    {"cfl", &c.cfl},
                    {"gravity_x", &c.gravity[0]},
                    {"gravity_y", &c.gravity[1]},
                    {"gravity_z", &c.gravity[2]}}))
      unknown();
  } else if (section == "surface") {
    if (real_keys({{"number_density_ratio_threshold", &c.number_density_ratio_threshold},
                   {"air_open_ratio_threshold", &c.air_open_ratio_threshold},
                   {"air_anisotropy_threshold", &c.air_anisotropy_threshold},
                   {"wall_normal_independence_threshold", &c.wall_normal_independence_threshold},
                   {"virtual_light_cone_cosine", &c.virtual_light_cone_cosine}}))
      return;
    if (key == "splash_neighbor_threshold") c.splash_neighbor_threshold = index_value(full, value);
    else if (key == "include_wall_neighbors") c.include_wall_neighbors = bool_value(full, value);
    else unknown();
  } else if (section == "lsmps") {
    if (!real_keys({{"regularization", &c.lsmps_regularization},
                    {"wall_weight_scale", &c.lsmps_wall_weight_scale}}))
      unknown();
  } else if (section == "correction") {
    if (!real_keys({{"ps_displacement_scale", &c.ps_displacement_scale},
                    {"ps_min_distance_ratio", &c.ps_min_distance_ratio},
                    {"ps_max_displacement_ratio", &c.ps_max_displacement_ratio},
                    {"wall_clearance_ratio", &c.wall_clearance_ratio},
                    {"velocity_smoothing_strength", &c.velocity_smoothing_strength}}))
      unknown();
  } else if (section == "files") {
    const std::string t = trim(value);
    if (key == "fluid_particle_file")
      set_text(f.fluid_particle_file, sizeof(f.fluid_particle_file), t, "fluid_particle_file");
    else if (key == "wall_particle_file")
      set_text(f.wall_particle_file, sizeof(f.wall_particle_file), t, "wall_particle_file");
    else if (key == "output_directory")
      set_text(f.output_directory, sizeof(f.output_directory), t, "output_directory");
    else if (key == "vtk_file_prefix")
      set_text(f.vtk_file_prefix, sizeof(f.vtk_file_prefix), t, "vtk_file_prefix");
    else if (key == "vtk_write_point_fields") f.vtk_write_point_fields = bool_value(full, value);
    else if (key == "amgx_config_path")
      set_text(f.amgx_config_path, sizeof(f.amgx_config_path), t, "amgx_config_path");
    else if (key == "amgx_print_solve_stats") f.amgx_print_solve_stats = bool_value(full, value);
    else unknown();
  } else {
    bad("Unknown simulation config section '" + section + "' on line " + std::to_string(line));
  }
}

void write_real(std::ostream& o, const char* k, double v) {
  o << k << '=' << std::setprecision(17) << v << '\n';
}
void write_bool(std::ostream& o, const char* k, bool v) {
  o << k << '=' << (v ? "true" : "false") << '\n';
}

std::vector<std::string> split_csv(std::string_view line) {
  std::vector<std::string> out;
  size_t start = 0;
  while (start <= line.size()) {
    const size_t end = line.find(',', start);
    const size_t stop = end == std::string_view::npos ? line.size() : end;
    out.push_back(trim(line.substr(start, stop - start)));
    if (end == std::string_view::npos) break;
    start = end + 1;
  }
  return out;
}

bool header_row(const std::vector<std::string>& v) {
  if (v.empty()) return false;
  bool ok = false;
  (void)parse_double(v.front(), &ok);
  return !ok;
}

}  // namespace

void default_config_files(lsmps_config_files& f) {
  std::memset(&f, 0, sizeof(f));
  std::strcpy(f.fluid_particle_file, "input/fluid_particles.csv");
  std::strcpy(f.wall_particle_file, "input/wall_particles.csv");
  std::strcpy(f.output_directory, "output");
  std::strcpy(f.vtk_file_prefix, "lsmps3d");
  std::strcpy(f.amgx_config_path, "configs/amgx_ppe.json");
  f.vtk_write_point_fields = 1;
  f.amgx_print_solve_stats = 0;
}

// validate_simulation_config (config.cpp:278-335)
void validate_config(const lsmps_config& c, const lsmps_config_files* f) {
  auto pos = [](double v, const char* k) {
    if (v <= 0.0) bad(std::string("SimulationConfig field '") + k + "' must be positive");
  };
  auto nonneg = [](double v, const char* k) {
    if (v < 0.0) bad(std::string("SimulationConfig field '") + k + "' must be non-negative");
  };
  pos(c.particle_spacing, "particle_spacing");
  pos(c.support_radius, "support_radius");
  nonneg(c.near_surface_radius, "near_surface_radius");
  pos(c.cell_size, "cell_size");
  pos(c.time_step, "time_step");
  pos(c.min_time_step, "min_time_step");
  pos(c.max_time_step, "max_time_step");
  pos(c.time_step_growth_factor, "time_step_growth_factor");
  nonneg(c.final_time, "final_time");
  pos(c.output_interval, "output_interval");
  pos(c.density, "density");
  pos(c.kinematic_viscosity, "kinematic_viscosity");
  pos(c.cfl, "cfl");
  nonneg(c.lsmps_regularization, "lsmps_regularization");
  pos(c.lsmps_wall_weight_scale, "lsmps_wall_weight_scale");
  nonneg(c.ps_displacement_scale, "ps_displacement_scale");
  pos(c.ps_min_distance_ratio, "ps_min_distance_ratio");
  nonneg(c.ps_max_displacement_ratio, "ps_max_displacement_ratio");
  nonneg(c.wall_clearance_ratio, "wall_clearance_ratio");
  nonneg(c.velocity_smoothing_strength, "velocity_smoothing_strength");
  if (c.support_radius > c.cell_size) bad("SimulationConfig requires support_radius <= cell_size");
  if (c.min_time_step > c.max_time_step)
    bad("SimulationConfig requires min_time_step <= max_time_step");
  if (c.time_step_growth_factor < 1.0)
    bad("SimulationConfig field 'time_step_growth_factor' must be >= 1");
  if (c.number_density_ratio_threshold <= 0.0)
    bad("SimulationConfig field 'number_density_ratio_threshold' must be positive");
  if (c.virtual_light_cone_cosine < -1.0 || c.virtual_light_cone_cosine > 1.0)
    bad("SimulationConfig field 'virtual_light_cone_cosine' must be in [-1, 1]");
  if (c.velocity_smoothing_strength > 1.0)
    bad("SimulationConfig field 'velocity_smoothing_strength' must be in [0, 1]");
  if (f) {
    if (!f->amgx_config_path[0]) bad("SimulationConfig field 'amgx_config_path' must not be empty");
    if (!f->fluid_particle_file[0])
      bad("SimulationConfig field 'fluid_particle_file' must not be empty");
    if (!f->wall_particle_file[0])
      bad("SimulationConfig field 'wall_particle_file' must not be empty");
    if (!f->vtk_file_prefix[0]) bad("SimulationConfig field 'vtk_file_prefix' must not be empty");
  }
}

// load_simulation_config (config.cpp:337-385): defaults, then the file's keys.
void load_config(const char* path, lsmps_config& c, lsmps_config_files& f) {
  std::ifstream in(path);
  if (!in) raise(LSMPS_ERR_IO, std::string("Unable to open simulation config: ") + path);
  lsmps_default_config(&c);
  default_config_files(f);
  std::string line, section;
  size_t n = 0;
  while (std::getline(in, line)) {
    ++n;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    const std::string st = trim(line);
    if (st.empty()) continue;
    if (st.front() == '[' && st.back() == ']') {
      section = trim(std::string_view(st).substr(1, st.size() - 2));
      if (section.empty()) bad("Invalid config line " + std::to_string(n) + ": empty section");
      continue;
    }
    const size_t eq = st.find('=');
    if (eq == std::string::npos) bad("Invalid config line " + std::to_string(n) + ": expected key=value");
    const std::string key = trim(std::string_view(st).substr(0, eq));
    const std::string value = trim(std::string_view(st).substr(eq + 1));
    if (key.empty()) bad("Invalid config line " + std::to_string(n) + ": empty key");
    apply(c, f, section, key, value, n);
  }
  validate_config(c, &f);
}

// save_simulation_config (config.cpp:387-431)
void save_config(const char* path, const lsmps_config& c, const lsmps_config_files& f) {
  validate_config(c, &f);
  std::ofstream o(path);
  if (!o) raise(LSMPS_ERR_IO, std::string("Unable to write simulation config: ") + path);
  o << "# lsmps3d simulation config\n";
  o << "\n[geometry]\n";
  write_real(o, "particle_spacing", c.particle_spacing);
  write_real(o, "support_radius", c.support_radius);
  write_real(o, "near_surface_radius", c.near_surface_radius);
  write_real(o, "cell_size", c.cell_size);
  o << "\n[simulation]\n";
  write_real(o, "time_step", c.time_step);
  write_real(o, "min_time_step", c.min_time_step);
  write_real(o, "max_time_step", c.max_time_step);
  write_real(o, "time_step_growth_factor", c.time_step_growth_factor);
  write_real(o, "final_time", c.final_time);
  write_real(o, "output_interval", c.output_interval);
  write_real(o, "density", c.density);
  write_real(o, "kinematic_viscosity", c.kinematic_viscosity);
  write_real(o, "cfl", c.cfl);
  write_real(o, "gravity_x", c.gravity[0]);
  write_real(o, "gravity_y", c.gravity[1]);
  write_real(o, "gravity_z", c.gravity[2]);
  o << "\n[surface]\n";
  o << "splash_neighbor_threshold=" << c.splash_neighbor_threshold << '\n';
  write_real(o, "number_density_ratio_threshold", c.number_density_ratio_threshold);
  write_real(o, "air_open_ratio_threshold", c.air_open_ratio_threshold);
  write_real(o, "air_anisotropy_threshold", c.air_anisotropy_threshold);
  write_bool(o, "include_wall_neighbors", c.include_wall_neighbors != 0);
  write_real(o, "wall_normal_independence_threshold", c.wall_normal_independence_threshold);
  write_real(o, "virtual_light_cone_cosine", c.virtual_light_cone_cosine);
  o << "\n[lsmps]\n";
  write_real(o, "regularization", c.lsmps_regularization);
  write_real(o, "wall_weight_scale", c.lsmps_wall_weight_scale);
  o << "\n[correction]\n";
  write_real(o, "ps_displacement_scale", c.ps_displacement_scale);
  write_real(o, "ps_min_distance_ratio", c.ps_min_distance_ratio);
  write_real(o, "ps_max_displacement_ratio", c.ps_max_displacement_ratio);
  write_real(o, "wall_clearance_ratio", c.wall_clearance_ratio);
  write_real(o, "velocity_smoothing_strength", c.velocity_smoothing_strength);
  o << "\n[files]\n";
  o << "fluid_particle_file=" << f.fluid_particle_file << '\n';
  o << "wall_particle_file=" << f.wall_particle_file << '\n';
  o << "output_directory=" << f.output_directory << '\n';
  o << "vtk_file_prefix=" << f.vtk_file_prefix << '\n';
  write_bool(o, "vtk_write_point_fields", f.vtk_write_point_fields != 0);
  o << "amgx_config_path=" << f.amgx_config_path << '\n';
  write_bool(o, "amgx_print_solve_stats", f.amgx_print_solve_stats != 0);
  if (!o) raise(LSMPS_ERR_IO, std::string("Unable to write simulation config: ") + path);
}

// FileManager::load_fluid_particles / load_wall_particles (file_manager.cpp:126-197):
// rows of `columns` reals; writes into `out` (column-major by field, host
// arrays) when given, and returns the number of records.
size_t read_particle_csv(const char* path, int columns, double* const* out, size_t capacity) {
  const char* what = columns == 6 ? "fluid particle" : "wall particle";
  std::ifstream in(path);
  if (!in) raise(LSMPS_ERR_IO, std::string("Unable to open ") + what + " CSV: " + path);
  static const char* const kNames[9] = {"x", "y", "z", "vx", "vy", "vz",
                                        "normal_x", "normal_y", "normal_z"};
  std::string line;
  size_t n = 0, rows = 0;
  while (std::getline(in, line)) {
    ++n;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    const std::string st = trim(line);
    if (st.empty()) continue;
    const std::vector<std::string> v = split_csv(st);
    if (rows == 0 && header_row(v)) continue;
    if (static_cast<int>(v.size()) != columns)
      bad(std::string("Invalid ") + what + " CSV row in " + path + " line " + std::to_string(n) +
          ": expected " + std::to_string(columns) + " columns, got " + std::to_string(v.size()));
    for (int k = 0; k < columns; ++k) {
      bool ok = false;
      const double x = parse_double(v[k], &ok);
      if (!ok)
        bad(std::string("Invalid real value in ") + path + " line " + std::to_string(n) +
            " column " + kNames[k] + ": '" + v[k] + "'");
      if (out && rows < capacity && out[k]) out[k][rows] = x;
    }
    ++rows;
  }
  if (out && rows > capacity)
    raise(LSMPS_ERR_CAPACITY, std::string(what) + " CSV has more records than the arrays hold");
  return rows;
}

// LegacyVtkWriter::make_path (vtk_writer.cpp:83-88)
std::string vtk_path(const char* directory, const char* prefix, size_t step) {
  std::ostringstream name;
  name << prefix << '_' << std::setw(6) << std::setfill('0') << step << ".vtk";
  return (std::filesystem::path(directory) / name.str()).string();
}

// LegacyVtkWriter::write (vtk_writer.cpp:90-149): POLYDATA points and vertices;
// point data in the reference's order (vectors, then real scalars, then
// integer scalars, each in the order given).
void write_vtk(const char* directory, const char* prefix, bool point_fields, size_t step,
               size_t count, const double* x, const double* y, const double* z,
               const lsmps_vtk_field* fields, int nfields) {
  require(count == 0 || (x && y && z), "VTK output needs x, y and z");
  for (int k = 0; k < nfields; ++k) {
    const lsmps_vtk_field& fd = fields[k];
    if (!fd.name || !fd.name[0]) bad("VTK point field name must not be empty.");
    const bool ok = fd.kind == LSMPS_VTK_VECTOR ? (fd.x && fd.y && fd.z)
                    : fd.kind == LSMPS_VTK_INT_SCALAR ? fd.ivalues != nullptr
                                                       : fd.x != nullptr;
    if (!ok || fd.count != count)
      bad(std::string(fd.kind == LSMPS_VTK_VECTOR ? "Vector field '" : "Scalar field '") + fd.name +
          "' length must match particle count.");
  }
  std::filesystem::create_directories(directory);
  const std::string path = vtk_path(directory, prefix, step);
  std::ofstream o(path);
  if (!o) raise(LSMPS_ERR_IO, "Failed to open VTK output file: " + path);
  o << std::setprecision(9);
  o << "# vtk DataFile Version 3.0\n";
  o << "lsmps3d particle diagnostics\n";
  o << "ASCII\n";
  o << "DATASET POLYDATA\n";
  o << "POINTS " << count << " float\n";
  for (size_t i = 0; i < count; ++i) o << x[i] << ' ' << y[i] << ' ' << z[i] << '\n';
  o << "VERTICES " << count << ' ' << count * 2 << '\n';
  for (size_t i = 0; i < count; ++i) o << "1 " << i << '\n';
  o << "POINT_DATA " << count << '\n';
  if (point_fields) {
    for (int k = 0; k < nfields; ++k) {
      const lsmps_vtk_field& fd = fields[k];
      if (fd.kind != LSMPS_VTK_VECTOR) continue;
      o << "VECTORS " << fd.name << " float\n";
      for (size_t i = 0; i < count; ++i) o << fd.x[i] << ' ' << fd.y[i] << ' ' << fd.z[i] << '\n';
    }
    for (int k = 0; k < nfields; ++k) {
      const lsmps_vtk_field& fd = fields[k];
      if (fd.kind != LSMPS_VTK_REAL_SCALAR) continue;
      o << "SCALARS " << fd.name << " float 1\n" << "LOOKUP_TABLE default\n";
      for (size_t i = 0; i < count; ++i) o << fd.x[i] << '\n';
    }
    for (int k = 0; k < nfields; ++k) {
      const lsmps_vtk_field& fd = fields[k];
      if (fd.kind != LSMPS_VTK_INT_SCALAR) continue;
      o << "SCALARS " << fd.name << " int 1\n" << "LOOKUP_TABLE default\n";
      for (size_t i = 0; i < count; ++i) o << fd.ivalues[i] << '\n';
    }
  }
  if (!o) raise(LSMPS_ERR_IO, "Failed to write VTK output file: " + path);
}

}  // namespace lsmps
