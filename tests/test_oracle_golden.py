"""The FP64 oracle against the known answers of lsmps-3D's own unit tests
(tests/golden/reference_unit_tests.json, transcribed from tests/unit/*.cu)."""
import numpy as np

from helpers import all_pairs_neighbors, config_from, lattice, make_case
from conftest import rel_l2


def _neighbor_case(cases, abi, g):
    cfg = abi.default_config()
    for k in range(3):
        cfg.cell_origin[k] = g["cell_origin"][k]
        cfg.cell_dims[k] = g["cell_dims"][k]
    cfg.cell_size = g["cell_size"]
    cfg.support_radius = g["support_radius"]
    f = g["fluid"]
    return make_case(cases, abi, "golden_neighbor", (f["x"], f["y"], f["z"]),
                     walls=g["walls"], config=cfg)


def test_neighbor_search_golden(oracle, cases, abi, golden):
    """smoke_neighbor_search.cu:152-183 (offsets exact, rows as sets)."""
    g = golden["neighbor_search"]
    case = _neighbor_case(cases, abi, g)
    nb = oracle.neighbors(case, max_neighbors=g["fluid_neighbors_per_particle"])
    assert nb["fluid_offsets"].tolist() == g["expected_fluid_offsets"]
    assert nb["wall_offsets"].tolist() == g["expected_wall_offsets"]
    for i, row in enumerate(g["expected_fluid_rows_sorted"]):
        got = nb["fluid_indices"][nb["fluid_offsets"][i]:nb["fluid_offsets"][i + 1]]
        assert sorted(got.tolist()) == row
    for i, row in enumerate(g["expected_wall_rows_sorted"]):
        got = nb["wall_indices"][nb["wall_offsets"][i]:nb["wall_offsets"][i + 1]]
        assert sorted(got.tolist()) == row


def test_neighbor_capacity_error(oracle, cases, abi, golden):
    """Capacity overflow is an error (require_neighbor_capacity, neighbor_search.cu:22-30)."""
    g = golden["neighbor_search"]
    case = _neighbor_case(cases, abi, g)
    import pytest
    with pytest.raises(RuntimeError, match="capacity"):
        # 5 fluid x 0 per particle -> capacity 1 (max(...,1)) < 6 required
        oracle.neighbors(case, max_neighbors=0)


def _surface_case(cases, abi, g):
    cfg = config_from(abi, g["config"])
    f = g["fluid"]
    return make_case(cases, abi, "golden_surface", (f["x"], f["y"], f["z"]), walls=g["walls"],
                     config=cfg)


def test_surface_detection_golden(oracle, cases, abi, golden):
    """smoke_surface_detection.cu:187-236: four labels and the diagnostic contrasts."""
    g = golden["surface_detection"]
    case = _surface_case(cases, abi, g)
    nb = oracle.neighbors(case, max_neighbors=g["fluid_neighbors_per_particle"])
    out = oracle.classify(case, nb)
    st = out["surface_type"]
    for idx, label in g["expected_labels"].items():
        assert st[int(idx)] == label, (idx, st)
    n0 = oracle.reference_number_density(case.config.particle_spacing, case.config.support_radius)
    nd, ratio = out["number_density"], out["number_density_ratio"]
    assert abs(ratio[1] - nd[1] / n0) <= g["ratio_identity_tolerance"]
    assert ratio[0] > ratio[1]
    assert out["anisotropy"][1] > out["anisotropy"][0]
    assert out["air_open_ratio"][1] > 0.0
    assert out["air_anisotropy"][1] > 0.0


def _lattice_case(cases, abi, g, velocity=None, surface_type=None, pressure=None, walls=None):
    cfg = config_from(abi, g["config"])
    return make_case(cases, abi, "golden_lattice", lattice(g["points"]), walls=walls, config=cfg,
                     velocity=velocity, surface_type=surface_type, pressure=pressure)


def test_moment_matrix_golden(oracle, cases, abi, golden):
    """smoke_moment_matrix.cu:92-129: sizes per kind and a finite centre inverse."""
    g = golden["moment_matrix"]
    case = _lattice_case(cases, abi, g)
    nb = all_pairs_neighbors(case.n_fluid)
    for kind, size in g["expected_sizes"].items():
        m = oracle.moments(case, nb, int(kind))
        assert m["size"] == size
        assert np.all(np.isfinite(m["inverse"][g["center"]]))
        assert np.all(m["info"] == 0)


def test_provision_golden(oracle, cases, abi, golden):
    """smoke_provision_explicit_update.cu:185-211."""
    g = golden["provision"]
    x, y, z = lattice(g["points"])
    vx = 1.5 + 0.5 * x + 2.0 * x * x
    st = np.zeros(x.shape[0], np.int32)
    st[g["splash_index"]] = abi.SPLASH
    case = _lattice_case(cases, abi, g, velocity=(vx, np.full_like(x, -2.0), 0.25 * z),
                         surface_type=st, walls=g["walls"])
    nb = all_pairs_neighbors(case.n_fluid)
    vel = oracle.moments(case, nb, abi.MOMENT_VELOCITY_WALL_DIRICHLET_A)
    out = oracle.provision(case, nb, vel["inverse"])
    c, tol = g["center"], g["tolerance"]
    assert abs(out["laplacian"][0][c] - g["expected_laplacian_x_center"]) <= tol
    for k, key in enumerate(("vx", "vy", "vz")):
        assert abs(out[key][c] - g["expected_center"][k]) <= tol
    assert abs(out["vx"][g["splash_index"]] - g["expected_splash_vx"]) <= tol
    for k, key in enumerate(("wall_vx", "wall_vy", "wall_vz")):
        assert abs(out[key][0] - g["expected_wall0"][k]) <= tol


def _ppe_setup(oracle, cases, abi, g):
    x, y, z = lattice(g["points"])
    st = np.zeros(x.shape[0], np.int32)
    st[g["surface_index"]] = abi.SURFACE
    case = _lattice_case(cases, abi, g, velocity=(x, 2.0 * y, 3.0 * z), surface_type=st)
    nb = all_pairs_neighbors(case.n_fluid)
    vel = oracle.moments(case, nb, abi.MOMENT_VELOCITY_WALL_DIRICHLET_A)
    pre = oracle.moments(case, nb, abi.MOMENT_PRESSURE_WALL_NEUMANN_A)
    tmp = {"vx": case.fluid["vx"], "vy": case.fluid["vy"], "vz": case.fluid["vz"],
           "wall_vx": np.zeros(0), "wall_vy": np.zeros(0), "wall_vz": np.zeros(0)}
    ppe = oracle.assemble(case, nb, tmp, vel["inverse"], pre["inverse"])
    return case, nb, ppe


def test_ppe_assembly_golden(oracle, cases, abi, golden):
    """smoke_ppe_matrix.cu:243-261: nnz, Dirichlet row, divergence and rhs."""
    g = golden["ppe_matrix"]
    case, nb, ppe = _ppe_setup(oracle, cases, abi, g)
    assert ppe["nnz"] == g["expected_nnz"]
    assert ppe["row_offsets"][0] == 0 and ppe["row_offsets"][-1] == g["expected_nnz"]
    r0 = ppe["row_offsets"][0]
    assert ppe["col_indices"][r0] == 0 and ppe["values"][r0] == 1.0 and ppe["rhs"][0] == 0.0
    c, tol = g["center"], g["tolerance"]
    assert abs(ppe["divergence"][c] - g["expected_divergence_center"]) <= tol
    assert abs(ppe["rhs"][c] - g["expected_rhs_center"]) <= tol
    rb, re_ = ppe["row_offsets"][c], ppe["row_offsets"][c + 1]
    assert ppe["col_indices"][rb] == c and np.isfinite(ppe["values"][rb]) and ppe["values"][rb] != 0


def test_ppe_solve_golden(oracle, abi, golden):
    """smoke_ppe_matrix.cu:60-107: [[4,-1],[-1,4]] solve and the p >= 0 clamp."""
    g = golden["ppe_solve_2x2"]
    for method in (abi.SOLVER_IDRS, abi.SOLVER_BICGSTAB, abi.SOLVER_PCG):
        for case in g["cases"]:
            opt = abi.default_solver_options()
            opt.method = method
            x, rep = oracle.solve(np.array(g["row_offsets"]), np.array(g["col_indices"]),
                                  np.array(g["values"]), np.array(case["rhs"]), opt)
            assert np.max(np.abs(x - np.array(case["expected"]))) <= g["tolerance"], (method, x)


def test_pressure_correction_golden(oracle, cases, abi, golden):
    """smoke_pressure_correction.cu:172-195."""
    g = golden["pressure_correction"]
    x, y, z = lattice(g["points"])
    n = x.shape[0]
    p = 20.0 + 2.0 * x - 3.0 * y + 4.0 * z
    st = np.zeros(n, np.int32)
    st[g["splash_index"]] = abi.SPLASH
    v = [np.full(n, c) for c in g["velocity"]]
    case = _lattice_case(cases, abi, g, velocity=v, surface_type=st, pressure=p)
    nb = all_pairs_neighbors(n)
    pre = oracle.moments(case, nb, abi.MOMENT_PRESSURE_WALL_NEUMANN_A)
    tmp = {"vx": v[0].copy(), "vy": v[1].copy(), "vz": v[2].copy()}
    out = oracle.correction(case, nb, tmp, p, pre["inverse"])
    c, tol = g["center"], g["tolerance"]
    for k in range(3):
        assert abs(out["grad"][k][c] - g["expected_gradient_center"][k]) <= tol
    for k, key in enumerate(("vx", "vy", "vz")):
        assert abs(out["fluid"][key][c] - g["expected_velocity_center"][k]) <= tol
    assert abs(out["fluid"]["x"][c] - g["expected_x_center"]) <= tol
