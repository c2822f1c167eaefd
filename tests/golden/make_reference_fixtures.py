"""Writes reference_unit_tests.json: the inputs and known answers of lsmps-3D's
own unit tests for the per-step path, transcribed from the test sources
(paths relative to the lsmps-3D tree).  Nothing is executed from the reference;
the values below are the literal fixtures and expectations those tests state.

Run:  python tests/golden/make_reference_fixtures.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def lattice27():
    """27-point unit lattice, z-major then y then x (smoke_moment_matrix.cu:18-28)."""
    pts = []
    for iz in (-1, 0, 1):
        for iy in (-1, 0, 1):
            for ix in (-1, 0, 1):
                pts.append((float(ix), float(iy), float(iz)))
    return pts


def main():
    lat = lattice27()
    fixtures = {
        "neighbor_search": {
            "source": "tests/unit/smoke_neighbor_search.cu:59-190",
            "fluid": {"x": [0.1, 0.4, 1.1, 1.4, 2.6], "y": [0.1, 0.1, 0.1, 0.7, 2.6],
                      "z": [0.1, 0.1, 0.1, 0.1, 2.6]},
            "walls": {"x": [0.2, 1.2, 2.5], "y": [0.5, 0.1, 2.6], "z": [0.1, 0.1, 2.6]},
            "cell_origin": [0.0, 0.0, 0.0], "cell_size": 1.0, "cell_dims": [3, 3, 3],
            "support_radius": 0.75,
            "fluid_neighbors_per_particle": 4, "wall_neighbors_per_particle": 3,
            "expected_fluid_offsets": [0, 1, 3, 5, 6, 6],
            "expected_wall_offsets": [0, 1, 2, 3, 4, 5],
            "expected_fluid_rows_sorted": [[1], [0, 2], [1, 3], [2], []],
            "expected_wall_rows_sorted": [[0], [0], [1], [1], [2]],
        },
        "surface_detection": {
            "source": "tests/unit/smoke_surface_detection.cu:38-236",
            "note": "wall normals are never written by the test (cudaMalloc'd buffers); "
                    "they are taken as zero here",
            "fluid": {"x": [1.0, 0.4, 1.6, 1.0, 1.0, 1.0, 3.5, 2.5],
                      "y": [1.0, 1.0, 1.0, 0.4, 1.6, 1.0, 3.5, 2.5],
                      "z": [1.0, 1.0, 1.0, 1.0, 1.0, 1.6, 3.5, 1.0]},
            "walls": {"x": [1.0, 1.0, 1.0, 0.4, 1.9, 3.1, 2.5, 2.5, 2.5, 2.5],
                      "y": [1.0, 1.0, 0.4, 0.4, 2.5, 2.5, 1.9, 3.1, 2.5, 2.5],
                      "z": [0.4, 1.6, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.4, 1.6]},
            "config": {"cell_origin": [0.0, 0.0, 0.0], "cell_size": 1.0, "cell_dims": [4, 4, 4],
                       "support_radius": 0.75, "near_surface_radius": 0.75,
                       "particle_spacing": 0.6, "splash_neighbor_threshold": 2,
                       "number_density_ratio_threshold": 0.75, "air_open_ratio_threshold": 0.10,
                       "air_anisotropy_threshold": 0.12, "include_wall_neighbors": 1,
                       "wall_normal_independence_threshold": 0.25,
                       "virtual_light_cone_cosine": 0.8660254},
            "fluid_neighbors_per_particle": 6, "wall_neighbors_per_particle": 4,
            "expected_labels": {"0": 1, "1": 2, "6": 3, "7": 0},
            "ratio_identity_tolerance": 1e-5,
        },
        "moment_matrix": {
            "source": "tests/unit/smoke_moment_matrix.cu:11-132",
            "points": lat, "center": 13, "all_pairs": True,
            "config": {"support_radius": 4.0, "cell_size": 4.0, "lsmps_regularization": 1e-10,
                       "lsmps_wall_weight_scale": 1.0, "density": 1.0, "gravity": [0.0, 0.0, 0.0]},
            "expected_sizes": {"0": 9, "1": 9, "2": 10, "3": 9},
        },
        "provision": {
            "source": "tests/unit/smoke_provision_explicit_update.cu:39-219",
            "points": lat, "center": 13, "all_pairs": True,
            "velocity": "vx = 1.5 + 0.5 x + 2 x^2, vy = -2, vz = 0.25 z",
            "splash_index": 0,
            "walls": {"x": [-2.0, 2.0], "y": [0.0, 0.0], "z": [0.0, 0.0], "vx": [0.4, -0.1],
                      "vy": [0.2, 0.3], "vz": [-0.5, 0.7]},
            "wall_neighbors": "none (all wall offsets 0)",
            "config": {"support_radius": 4.0, "cell_size": 4.0, "lsmps_regularization": 1e-10,
                       "lsmps_wall_weight_scale": 1.0, "kinematic_viscosity": 0.25,
                       "time_step": 0.2, "gravity": [0.1, -0.2, 0.3]},
            "expected_center": [1.72, -2.04, 0.06],
            "expected_laplacian_x_center": 4.0,
            "expected_splash_vx": 3.02,
            "expected_wall0": [0.42, 0.16, -0.44],
            "tolerance": 2e-3,
        },
        "ppe_matrix": {
            "source": "tests/unit/smoke_ppe_matrix.cu:119-286",
            "points": lat, "center": 13, "all_pairs": True,
            "velocity": "v = (x, 2 y, 3 z) (fluid and temporary)",
            "surface_index": 0,
            "config": {"support_radius": 4.0, "cell_size": 4.0, "lsmps_regularization": 1e-10,
                       "time_step": 0.5, "density": 2.0, "gravity": [0.0, 0.0, 0.0]},
            "expected_nnz": 729, "expected_divergence_center": 6.0,
            "expected_rhs_center": 12.0, "tolerance": 2e-3,
        },
        "ppe_solve_2x2": {
            "source": "tests/unit/smoke_ppe_matrix.cu:55-107",
            "row_offsets": [0, 2, 4], "col_indices": [0, 1, 0, 1], "values": [4.0, -1.0, -1.0, 4.0],
            "cases": [{"rhs": [3.0, 3.0], "expected": [1.0, 1.0]},
                      {"rhs": [-3.0, -3.0], "expected": [0.0, 0.0]}],
            "tolerance": 2e-5,
        },
        "pressure_correction": {
            "source": "tests/unit/smoke_pressure_correction.cu:33-203",
            "points": lat, "center": 13, "all_pairs": True,
            "velocity": [0.5, -0.25, 0.75],
            "pressure": "p = 20 + 2 x - 3 y + 4 z",
            "splash_index": 0,
            "config": {"support_radius": 4.0, "cell_size": 4.0, "particle_spacing": 1.0,
                       "lsmps_regularization": 1e-10, "time_step": 0.5, "density": 2.0,
                       "ps_displacement_scale": 0.0, "wall_clearance_ratio": 0.0,
                       "velocity_smoothing_strength": 0.0},
            "expected_gradient_center": [2.0, -3.0, 4.0],
            "expected_velocity_center": [0.0, 0.5, -0.25],
            "expected_x_center": 0.125,
            "tolerance": 2e-3,
        },
        "config": {
            "source": "tests/unit/smoke_config.cpp:17-135",
            "default_ratios": {"support_radius_ratio": 3.1, "particle_volume": 1.0e-6},
            "ini": ("# comments and whitespace are allowed\n[geometry]\nparticle_spacing = 0.02\n"
                    "support_radius = 0.062\nnear_surface_radius = 0.04\ncell_size = 0.062\n\n"
                    "[simulation]\ntime_step = 0.0005\nmin_time_step = 0.0001\n"
                    "max_time_step = 0.002\ntime_step_growth_factor = 1.2\nfinal_time = 0.5\n"
                    "output_interval = 0.05\ndensity = 998\nkinematic_viscosity = 0.0000011\n"
                    "cfl = 0.25\ngravity_z = -9.8\n\n[surface]\nsplash_neighbor_threshold = 12\n"
                    "number_density_ratio_threshold = 0.82\nair_open_ratio_threshold = 0.31\n"
                    "air_anisotropy_threshold = 0.04\ninclude_wall_neighbors = true\n"
                    "virtual_light_cone_cosine = 0.8\n\n[lsmps]\nregularization = 0.000000001\n"
                    "wall_weight_scale = 0.75\n\n[correction]\nps_displacement_scale = 0.02\n"
                    "ps_min_distance_ratio = 0.8\nps_max_displacement_ratio = 0.15\n"
                    "wall_clearance_ratio = 0.3\nvelocity_smoothing_strength = 0.05\n\n[files]\n"
                    "output_directory = /tmp/lsmps3d_output\nvtk_file_prefix = smoke\n"
                    "vtk_write_point_fields = false\namgx_config_path = configs/custom.json\n"
                    "amgx_print_solve_stats = true\n"),
            "expected": {"particle_spacing": 0.02, "support_radius": 0.062, "cell_size": 0.062,
                         "min_time_step": 0.0001, "max_time_step": 0.002,
                         "time_step_growth_factor": 1.2, "final_time": 0.5,
                         "output_interval": 0.05, "ps_displacement_scale": 0.02,
                         "velocity_smoothing_strength": 0.05},
            "expected_files": {"vtk_write_point_fields": 0, "amgx_print_solve_stats": 1,
                               "vtk_file_prefix": "smoke",
                               "amgx_config_path": "configs/custom.json"},
            "invalid": [{"support_radius": -1.0},
                        {"min_time_step": 0.01, "max_time_step": 0.001}],
            "bad_ini": "[geometry]\nunknown_key=1\n",
            "tolerance": 1e-10,
        },
        "file_manager": {
            "source": "tests/unit/smoke_file_manager.cpp:34-137",
            "fluid_csv": "# x,y,z,vx,vy,vz\nx,y,z,vx,vy,vz\n0,0,0,0,0,0\n1,0,0,0.5,0.25,0.125\n",
            "wall_csv": ("x,y,z,vx,vy,vz,nx,ny,nz\n0,-1,0,0,0,0,0,1,0\n"
                         "1,-1,0,0.1,0.2,0.3,0,1,0\n"),
            "expected": {"fluid_count": 2, "wall_count": 2, "fluid_vx": [0.0, 0.5],
                         "fluid_pressure": [0.0, 0.0], "fluid_surface_type": [0, 0],
                         "wall_normal_y0": 1.0, "wall_vz1": 0.3},
            "vtk_prefix": "case", "vtk_step": 3,
            "expected_files": ["case_fluid_000003.vtk", "case_wall_000003.vtk"],
            "expected_fluid_tokens": ["VECTORS velocity float", "SCALARS pressure float 1"],
            "expected_wall_tokens": ["VECTORS normal float"],
            "bad_csv": "x,y,z,vx,vy,vz\n0,not-a-number,0,0,0,0\n",
            "extra_column_csv": "x,y,z,vx,vy,vz,pressure\n0,0,0,0,0,0,100\n",
            "tolerance": 1e-10,
        },
        "vtk_writer": {
            "source": "tests/unit/smoke_vtk_writer.cpp:29-88",
            "prefix": "particles", "step": 7,
            "points": {"x": [0.0, 1.0, 0.0], "y": [0.0, 0.0, 1.0], "z": [0.0, 0.0, 0.5]},
            "vectors": {"velocity": [[1.0, 0.0, -1.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0]],
                        "pressure_gradient": [[0.5, 0.0, -0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.5]]},
            "real_scalars": {"pressure": [0.0, 10.0, 20.0], "number_density": [1.0, 0.55, 0.0],
                             "anisotropy": [0.1, 0.8, 0.0]},
            "int_scalars": {"surface_type": [0, 2, 3], "fluid_neighbor_count": [6, 4, 0],
                            "wall_neighbor_count": [2, 1, 0]},
            "expected_file": "particles_000007.vtk",
            "expected_tokens": ["DATASET POLYDATA", "POINTS 3 float", "VECTORS velocity float",
                                "VECTORS pressure_gradient float", "SCALARS pressure float 1",
                                "SCALARS surface_type int 1", "SCALARS fluid_neighbor_count int 1",
                                "SCALARS wall_neighbor_count int 1",
                                "SCALARS number_density float 1", "SCALARS anisotropy float 1"],
            "mismatched_vector": "velocity",
        },
        "time_step": {
            "source": "tests/unit/smoke_time_step.cpp:24-98",
            "config": {"particle_spacing": 0.02, "time_step": 0.001, "min_time_step": 0.0001,
                       "max_time_step": 0.01, "time_step_growth_factor": 2.0, "cfl": 0.25,
                       "final_time": 0.012, "output_interval": 0.004},
            "initial_output_due": True,
            "initial_output_index": 1,
            "advances": [
                {"max_velocity": 2.0, "expected_dt": 0.002, "should_output": False,
                 "reached_final_time": False, "label": "growth-limited dt"},
                {"max_velocity": 2.0, "expected_dt": 0.0025, "should_output": True,
                 "output_index": 2, "label": "CFL-limited dt"},
                {"max_velocity": 1000.0, "expected_dt": 0.0001, "label": "minimum-clamped dt"},
            ],
            "then_max_velocity": 0.0, "guard": 20, "expected_final_time": 0.012,
            "tolerance": 1e-12,
            "invalid_growth_factor": 0.5,
        },
    }
    with open(os.path.join(HERE, "reference_unit_tests.json"), "w") as fh:
        json.dump(fixtures, fh, indent=1)


if __name__ == "__main__":
    main()
