"""HIP path vs the FP64 oracle, through the C-ABI (liblsmps3d_amd.so).

Bars: neighbor CSR and surface labels bit-exact; FP64-mode floating results
within 1e-9 relative L2 (the oracle and the GPU both solve the PPE to a 1e-10
relative true residual); mixed-mode pressure within the 1e-5 relative L2 the
target sets.  Velocities after the correction are compared relative to the
magnitude of the predictor u*, because in near-hydrostatic states
u = u* - dt/rho grad p cancels to a tiny remainder.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import brute_force_neighbors

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_PRESSURE_TOL = 1e-5


def _small_cases(cases):
    return [
        cases.hydrostatic(n=8, l0=0.01),
        cases.hydrostatic(n=20, l0=0.01),
        cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0),
        cases.splash(n=24, l0=0.01, seed=3, top_layers=3),
    ]


@pytest.fixture(scope="module")
def small(cases):
    return _small_cases(cases)


@pytest.fixture(scope="module")
def oracle_runs(oracle, small):
    return {c.name: oracle.pipeline(c) for c in small}


def _vec_err(got, ref):
    """Relative L2 error of a 3-component field, normalised by the whole field
    (a single component can be ~1e-9 of the others, e.g. grad_x in a hydrostatic box)."""
    num = sum(float(np.sum((np.asarray(a) - np.asarray(b)) ** 2)) for a, b in zip(got, ref))
    den = sum(float(np.sum(np.asarray(b) ** 2)) for b in ref)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def _check_geometry(g, ref):
    nb = ref["neighbors"]
    assert np.array_equal(g["fluid_offsets"], nb["fluid_offsets"])
    assert np.array_equal(g["fluid_indices"], nb["fluid_indices"])
    assert np.array_equal(g["wall_offsets"], nb["wall_offsets"])
    assert np.array_equal(g["wall_indices"], nb["wall_indices"])
    assert np.array_equal(g["surface_type"], ref["surface"]["surface_type"])


@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_step_fp64_parity(rt, abi, small, oracle_runs, idx):
    case = small[idx]
    ref = oracle_runs[case.name]
    g = rt.run_step_on_device(case, precision=abi.PRECISION_FP64)
    print(case.name, case.n_fluid, g["report"], flush=True)
    _check_geometry(g, ref)
    pv = ref["provision"]
    for k in ("vx", "vy", "vz", "wall_vx", "wall_vy", "wall_vz"):
        assert rel_l2(g["provision"][k], pv[k]) <= FP64_TOL, k
    pp = ref["ppe"]
    assert np.array_equal(g["ppe"]["row_offsets"], pp["row_offsets"])
    assert np.array_equal(g["ppe"]["col_indices"], pp["col_indices"])
    assert g["ppe"]["nnz"] == pp["nnz"]
    assert rel_l2(g["ppe"]["values"], pp["values"]) <= FP64_TOL
    assert rel_l2(g["ppe"]["rhs"], pp["rhs"]) <= FP64_TOL
    assert rel_l2(g["ppe"]["divergence"], pp["divergence"]) <= FP64_TOL
    assert g["report"]["converged"] == 1
    assert rel_l2(g["pressure"], ref["pressure"]) <= FP64_TOL
    cg = ref["correction"]
    assert _vec_err(g["grad"], cg["grad"]) <= 1e-8
    assert _vec_err(g["ps"], cg["ps"]) <= 1e-8
    for k in ("x", "y", "z"):
        assert rel_l2(g["fluid"][k], cg["fluid"][k]) <= 1e-12
    scale = max(np.linalg.norm(pv["vz"]), 1e-300)
    for k in ("vx", "vy", "vz"):
        assert np.linalg.norm(g["fluid"][k] - cg["fluid"][k]) / scale <= 1e-8, k


@pytest.mark.parametrize("idx", [1, 2, 3])
def test_step_mixed_parity(rt, abi, small, oracle_runs, idx):
    case = small[idx]
    ref = oracle_runs[case.name]
    g = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED)
    print(case.name, case.n_fluid, g["report"], flush=True)
    _check_geometry(g, ref)
    assert g["report"]["converged"] == 1
    assert g["report"]["relative_residual"] <= 1e-10
    assert rel_l2(g["pressure"], ref["pressure"]) <= MIXED_PRESSURE_TOL
    cg = ref["correction"]
    assert _vec_err(g["grad"], cg["grad"]) <= 1e-5
    scale = max(np.linalg.norm(ref["provision"]["vz"]), 1e-300)
    for k in ("vx", "vy", "vz"):
        assert np.linalg.norm(g["fluid"][k] - cg["fluid"][k]) / scale <= 1e-5, k


def test_neighbor_brute_force_jittered(rt, abi, cases):
    """Neighbor sets against O(N^2) distance tests; order = oracle canonical order."""
    case = cases.dam_break(block=(14, 9, 9), tank=(20, 9, 14), l0=0.01, seed=5, amplitude=0.3)
    g = rt.run_step_on_device(case, precision=abi.PRECISION_FP64)
    f = case.fluid
    w = case.walls
    re = case.config.support_radius
    rows = brute_force_neighbors(f["x"], f["y"], f["z"], f["x"], f["y"], f["z"], re, True)
    wrows = brute_force_neighbors(f["x"], f["y"], f["z"], w["x"], w["y"], w["z"], re, False)
    for i in range(case.n_fluid):
        got = g["fluid_indices"][g["fluid_offsets"][i]:g["fluid_offsets"][i + 1]]
        assert sorted(got.tolist()) == rows[i].tolist()
        got = g["wall_indices"][g["wall_offsets"][i]:g["wall_offsets"][i + 1]]
        assert sorted(got.tolist()) == wrows[i].tolist()


def field_errors(g, ref):
    """Max / RMS absolute and relative-L2 errors of pressure, pressure gradient
    and corrected velocity against the oracle's step (velocity relative to the
    predictor u*, whose magnitude the correction nearly cancels)."""
    cg = ref["correction"]
    out = {}
    p, pr = np.asarray(g["pressure"]), np.asarray(ref["pressure"])
    out["pressure"] = (float(np.max(np.abs(p - pr))), float(np.sqrt(np.mean((p - pr) ** 2))),
                       rel_l2(p, pr))
    gg = np.stack(g["grad"]).ravel()
    gr = np.stack(cg["grad"]).ravel()
    out["grad_p"] = (float(np.max(np.abs(gg - gr))), float(np.sqrt(np.mean((gg - gr) ** 2))),
                     _vec_err(g["grad"], cg["grad"]))
    v = np.concatenate([g["fluid"][k] for k in ("vx", "vy", "vz")])
    vr = np.concatenate([cg["fluid"][k] for k in ("vx", "vy", "vz")])
    scale = max(float(np.linalg.norm(np.concatenate(
        [ref["provision"][k] for k in ("vx", "vy", "vz")]))), 1e-300)
    out["velocity"] = (float(np.max(np.abs(v - vr))), float(np.sqrt(np.mean((v - vr) ** 2))),
                       float(np.linalg.norm(v - vr) / scale))
    return out


def print_field_errors(label, errs):
    for k, (mx, rms, rel) in errs.items():
        print(f"{label} {k:9s} max {mx:.3e} rms {rms:.3e} rel-L2 {rel:.3e}", flush=True)


def test_c2_per_field_errors(rt, abi, cases, oracle):
    """Dam-break C2 (250,000 fluid) in both precision modes against the FP64
    oracle's full step: CSR and labels bit-exact, and the max / RMS / relative
    errors of pressure, grad p and velocity (printed) within each mode's bar."""
    case = cases.dam_break()
    ref = oracle.pipeline(case)
    nb = ref["neighbors"]
    bars = {"FP64": {"pressure": 1e-9, "grad_p": 1e-8, "velocity": 1e-8},
            "MIXED": {"pressure": 1e-5, "grad_p": 1e-5, "velocity": 1e-5}}
    pressures = {}
    for mode, prec in (("FP64", abi.PRECISION_FP64), ("MIXED", abi.PRECISION_MIXED)):
        print(f"C2 {mode} n={case.n_fluid}", flush=True)
        g = rt.run_step_on_device(case, precision=prec)
        print(mode, g["report"], g["stage_ms"], flush=True)
        assert np.array_equal(g["fluid_offsets"], nb["fluid_offsets"])
        assert np.array_equal(g["fluid_indices"], nb["fluid_indices"])
        assert np.array_equal(g["wall_offsets"], nb["wall_offsets"])
        assert np.array_equal(g["wall_indices"], nb["wall_indices"])
        assert np.array_equal(g["surface_type"], ref["surface"]["surface_type"])
        assert g["report"]["converged"] == 1
        errs = field_errors(g, ref)
        print_field_errors(f"C2 {mode}", errs)
        for k, bar in bars[mode].items():
            assert errs[k][2] <= bar, (mode, k, errs[k])
        pressures[mode] = g["pressure"]
    assert rel_l2(pressures["MIXED"], pressures["FP64"]) <= MIXED_PRESSURE_TOL


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_c1_quadratic_weight(rt, abi, cases, oracle, precision):
    """C1 as stated for the workloads: the 20^3 hydrostatic box with the weight
    w = (1 - r/re)^2 in every stage (surface labels, moments, provision, PPE,
    correction), against the oracle run with the same weight."""
    case = cases.hydrostatic(n=20, l0=0.01, weight_kernel=abi.WEIGHT_QUADRATIC)
    assert case.config.weight_kernel == abi.WEIGHT_QUADRATIC
    ref = oracle.pipeline(case)
    prec = abi.PRECISION_FP64 if precision == "FP64" else abi.PRECISION_MIXED
    g = rt.run_step_on_device(case, precision=prec)
    print("C1 quadratic", precision, g["report"], flush=True)
    _check_geometry(g, ref)
    assert g["report"]["converged"] == 1
    errs = field_errors(g, ref)
    print_field_errors(f"C1-quadratic {precision}", errs)
    if precision == "FP64":
        pp = ref["ppe"]
        assert np.array_equal(g["ppe"]["row_offsets"], pp["row_offsets"])
        assert np.array_equal(g["ppe"]["col_indices"], pp["col_indices"])
        assert rel_l2(g["ppe"]["values"], pp["values"]) <= FP64_TOL
        assert rel_l2(g["ppe"]["rhs"], pp["rhs"]) <= FP64_TOL
        for k, kk in (("vx", "vx"), ("vy", "vy"), ("vz", "vz")):
            assert rel_l2(g["provision"][k], ref["provision"][kk]) <= FP64_TOL, k
        assert errs["pressure"][2] <= FP64_TOL
        assert errs["grad_p"][2] <= 1e-8
        assert errs["velocity"][2] <= 1e-8
    else:
        assert errs["pressure"][2] <= MIXED_PRESSURE_TOL
        assert errs["grad_p"][2] <= 1e-5
        assert errs["velocity"][2] <= 1e-5
    # the quadratic weight really changed the operators
    lin = oracle.pipeline(cases.hydrostatic(n=20, l0=0.01))
    assert rel_l2(lin["ppe"]["values"], ref["ppe"]["values"]) > 1e-3


def test_metric_case_against_oracle(rt, abi, cases, oracle):
    """M1, the 1,000,000-particle dam-break the throughput is quoted on, in
    mixed precision against the FP64 oracle's full step: neighbour CSR (fluid
    and wall) and labels bit-exact, pressure within 1e-5 relative L2, and the
    per-field errors printed."""
    case = cases.metric_case()
    print("M1 oracle step", case.n_fluid, flush=True)
    ref = oracle.pipeline(case)
    nb = ref["neighbors"]
    print("M1 GPU mixed step", flush=True)
    g = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED)
    print("M1", g["report"], g["stage_ms"], "oracle", ref["report"], flush=True)
    assert np.array_equal(g["fluid_offsets"], nb["fluid_offsets"])
    assert np.array_equal(g["fluid_indices"], nb["fluid_indices"])
    assert np.array_equal(g["wall_offsets"], nb["wall_offsets"])
    assert np.array_equal(g["wall_indices"], nb["wall_indices"])
    assert np.array_equal(g["surface_type"], ref["surface"]["surface_type"])
    assert g["report"]["converged"] == 1
    errs = field_errors(g, ref)
    print_field_errors("M1 MIXED", errs)
    assert errs["pressure"][2] <= MIXED_PRESSURE_TOL
    assert errs["grad_p"][2] <= 1e-5
    assert errs["velocity"][2] <= 1e-5


@pytest.mark.parametrize("case_name", ["C2", "M1"])
def test_mixed_error_target(rt, abi, cases, oracle, case_name):
    """The throughput configuration: the mixed step with the PPE solved to the
    1e-6 relative true residual of the reference's own AMGX solve
    (ppe_matrix.cu:607) against the oracle's FP64 step solved to 1e-10.
    Pressure, grad p and the corrected velocity (relative to the velocity
    itself and to u*) stay within the 1e-5 relative-L2 bar; CSR and labels are
    bit-exact as in every step."""
    case = cases.dam_break() if case_name == "C2" else cases.metric_case()
    ref = oracle.pipeline(case)
    opts = abi.default_solver_options()
    opts.tolerance = 1e-6
    g = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED, solver=opts)
    print(case_name, "mixed at 1e-6", g["report"], "oracle", ref["report"], flush=True)
    _check_geometry(g, ref)
    assert g["report"]["converged"] == 1
    assert g["report"]["relative_residual"] <= 1e-6
    errs = field_errors(g, ref)
    print_field_errors(f"{case_name} MIXED 1e-6", errs)
    v = np.concatenate([g["fluid"][k] for k in ("vx", "vy", "vz")])
    vr = np.concatenate([ref["correction"]["fluid"][k] for k in ("vx", "vy", "vz")])
    v_rel = rel_l2(v, vr)
    print(case_name, "velocity rel L2 of the velocity itself", v_rel, flush=True)
    assert errs["pressure"][2] <= 1e-5
    assert errs["grad_p"][2] <= 1e-5
    assert errs["velocity"][2] <= 1e-5
    assert v_rel <= 1e-5


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_preconditioners_agree(rt, abi, cases, precision):
    """The preconditioners only change the Krylov path: Jacobi, Schwarz blocks,
    and Schwarz plus the coarse-space correction all solve the same FP64 system
    to the 1e-10 residual, so the pressures agree to well within the parity
    bar; each addition needs fewer SpMVs on a dam-break with wrong-sign
    wall-corner rows; and every solve is bit-reproducible."""
    case = cases.dam_break(block=(40, 24, 24), tank=(80, 24, 48), l0=0.01, seed=1)
    prec = getattr(abi, "PRECISION_" + precision)
    jac = abi.default_solver_options()
    jac.preconditioner = abi.PRECOND_JACOBI
    sch = abi.default_solver_options()
    sch.preconditioner = abi.PRECOND_SCHWARZ
    crs = abi.default_solver_options()
    assert crs.preconditioner == abi.PRECOND_SCHWARZ_COARSE
    gj = rt.run_step_on_device(case, precision=prec, solver=jac)
    gs = rt.run_step_on_device(case, precision=prec, solver=sch)
    gc = rt.run_step_on_device(case, precision=prec, solver=crs)
    print(precision, "jacobi", gj["report"], "schwarz", gs["report"], "coarse", gc["report"],
          flush=True)
    for g in (gj, gs, gc):
        assert g["report"]["converged"] == 1
    assert gj["report"]["schwarz_blocks"] == 0 and gj["report"]["coarse_unknowns"] == 0
    assert gs["report"]["schwarz_blocks"] > 0 and gs["report"]["coarse_unknowns"] == 0
    assert gs["report"]["schwarz_rows"] >= gs["report"]["schwarz_blocks"]
    assert gc["report"]["schwarz_blocks"] == gs["report"]["schwarz_blocks"]
    assert gc["report"]["coarse_unknowns"] > 1
    assert rel_l2(gs["pressure"], gj["pressure"]) <= 1e-8
    assert rel_l2(gc["pressure"], gj["pressure"]) <= 1e-8
    assert gs["report"]["iterations"] < gj["report"]["iterations"]
    assert gc["report"]["iterations"] < gs["report"]["iterations"]
    # deterministic: second runs reproduce the pressure bit for bit
    assert np.array_equal(rt.run_step_on_device(case, precision=prec, solver=sch)["pressure"],
                          gs["pressure"])
    assert np.array_equal(rt.run_step_on_device(case, precision=prec, solver=crs)["pressure"],
                          gc["pressure"])


@pytest.mark.parametrize("block,tank,cells,lo,hi", [
    ((60, 40, 40), (120, 40, 80), 1, 2048, 4096),  # 4 rows per panel thread, 8-column panels
    ((44, 28, 28), (88, 28, 56), 1, 1024, 2048),   # 2 rows per panel thread, 16-column panels
    ((60, 40, 40), (120, 40, 80), 2, 64, 1024),    # 1 row per panel thread, 16-column panels
])
def test_coarse_space_sizes(rt, abi, cases, block, tank, cells, lo, hi):
    """The coarse inverse runs in Gauss-Jordan panels whose width and rows per
    thread follow the number of coarse unknowns.  At each size the coarse
    space must leave the solution unchanged (same FP64 system, same 1e-10
    residual) and cut the iteration count against the Schwarz blocks alone,
    which a wrong inverse would not do."""
    case = cases.dam_break(block=block, tank=tank, l0=0.01, seed=2)
    prec = abi.PRECISION_FP64
    sch = abi.default_solver_options()
    sch.preconditioner = abi.PRECOND_SCHWARZ
    crs = abi.default_solver_options()
    crs.coarse_cells = cells
    gs = rt.run_step_on_device(case, precision=prec, solver=sch)
    gc = rt.run_step_on_device(case, precision=prec, solver=crs)
    print(block, cells, "schwarz", gs["report"], "coarse", gc["report"], flush=True)
    assert gs["report"]["converged"] == 1 and gc["report"]["converged"] == 1
    assert lo < gc["report"]["coarse_unknowns"] <= hi
    assert rel_l2(gc["pressure"], gs["pressure"]) <= 1e-8
    assert gc["report"]["iterations"] < gs["report"]["iterations"]


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_batched_tanks_match_single_tanks(rt, abi, cases, oracle, precision):
    """Four independent tanks (hydrostatic and dam-break, uneven sizes) stepped
    as one batched system: every tank's neighbor rows and labels equal its own
    single-tank oracle run bit for bit (indices offset by the tank's start), and
    its pressure agrees to the tolerance of the union solve."""
    case = cases.batched_tanks(count=4, min_fluid=2000, max_fluid=8000)
    g = rt.run_step_on_device(case, precision=getattr(abi, "PRECISION_" + precision))
    print(precision, case.n_fluid, g["report"], flush=True)
    assert g["report"]["converged"] == 1
    fo, wo = case.extra["fluid_offsets"], case.extra["wall_offsets"]
    tol = 1e-7 if precision == "FP64" else 1e-5
    for k, tank in enumerate(case.extra["tanks"]):
        ref = oracle.pipeline(tank)
        nb = ref["neighbors"]
        a, b = fo[k], fo[k + 1]
        f_off = g["fluid_offsets"][a:b + 1]
        w_off = g["wall_offsets"][a:b + 1]
        assert np.array_equal(f_off - f_off[0], nb["fluid_offsets"]), k
        assert np.array_equal(w_off - w_off[0], nb["wall_offsets"]), k
        assert np.array_equal(g["fluid_indices"][f_off[0]:f_off[-1]] - a, nb["fluid_indices"]), k
        assert np.array_equal(g["wall_indices"][w_off[0]:w_off[-1]] - wo[k], nb["wall_indices"]), k
        assert np.array_equal(g["surface_type"][a:b], ref["surface"]["surface_type"]), k
        assert rel_l2(g["pressure"][a:b], ref["pressure"]) <= tol, k


@pytest.mark.gpu
def test_batched_tanks_coarse_space_per_tank(rt, abi, cases):
    """The batched solve's coarse space is the union of the tanks' own: each
    tank aggregated over its own grid gives the aggregates a single-tank solve
    of that tank builds (sum of the single-tank counts = the batched count), the
    block-diagonal coarse term leaves the batched solution unchanged against
    the Schwarz blocks alone, and it cuts the iteration count."""
    case = cases.batched_tanks(count=4, min_fluid=6000, max_fluid=30000)
    prec = abi.PRECISION_FP64
    sch = abi.default_solver_options()
    sch.preconditioner = abi.PRECOND_SCHWARZ
    gs = rt.run_step_on_device(case, precision=prec, solver=sch)
    gc = rt.run_step_on_device(case, precision=prec)
    print("schwarz", gs["report"], "coarse", gc["report"], flush=True)
    assert gs["report"]["converged"] == 1 and gc["report"]["converged"] == 1
    singles = [rt.run_step_on_device(t, precision=prec)["report"]["coarse_unknowns"]
               for t in case.extra["tanks"]]
    print("single-tank coarse unknowns", singles, flush=True)
    assert all(c > 1 for c in singles)
    assert gc["report"]["coarse_unknowns"] == sum(singles)
    fo = case.extra["fluid_offsets"]
    for k in range(len(case.extra["tanks"])):
        a, b = fo[k], fo[k + 1]
        assert rel_l2(gc["pressure"][a:b], gs["pressure"][a:b]) <= 1e-8, k
    assert gc["report"]["iterations"] < gs["report"]["iterations"]


@pytest.mark.gpu
@pytest.mark.parametrize("batched", [False, True])
def test_schwarz_fp16_application_matches_fp32(rt, abi, cases, batched):
    """Mixed precision applies the Schwarz inverses from FP16 copies scaled per
    column by a power of two.  It is a slightly different preconditioner of the
    same FP64 system: the solve must reach the same 1e-10 residual, the pressure
    must agree with the FP32 application's to the solve's accuracy, and the
    iteration count must stay within a few percent."""
    case = (cases.batched_tanks(count=4, min_fluid=4000, max_fluid=20000) if batched
            else cases.dam_break(block=(40, 40, 40), tank=(80, 40, 80), l0=0.01, seed=2))
    prec = abi.PRECISION_MIXED
    g16 = rt.run_step_on_device(case, precision=prec)
    g32 = rt.run_step_on_device(case, precision=prec, flags=abi.FLAG_SCHWARZ_FP32)
    print("fp16", g16["report"], "fp32", g32["report"], flush=True)
    assert g16["report"]["converged"] == 1 and g32["report"]["converged"] == 1
    assert g16["report"]["schwarz_blocks"] == g32["report"]["schwarz_blocks"] > 0
    assert rel_l2(g16["pressure"], g32["pressure"]) <= 1e-8
    assert g16["report"]["iterations"] <= 1.1 * g32["report"]["iterations"] + 16


def _batch_of(cases, tanks):
    """Tanks concatenated into one batched case (as cases.batched_tanks does)."""
    fluid = {key: np.concatenate([t.fluid[key] for t in tanks]) for key in tanks[0].fluid}
    walls = {key: np.concatenate([t.walls[key] for t in tanks]) for key in tanks[0].walls}
    fo = np.concatenate([[0], np.cumsum([t.n_fluid for t in tanks])]).astype(np.int64)
    wo = np.concatenate([[0], np.cumsum([t.n_wall for t in tanks])]).astype(np.int64)
    return cases.Case("batch", fluid, walls, tanks[0].config.copy(),
                      {"fluid_offsets": fo, "wall_offsets": wo,
                       "grids": [t.config.grid() for t in tanks], "tanks": tanks})


@pytest.mark.gpu
def test_batched_coarse_space_widens_and_skips_empty_tanks(rt, abi, cases, oracle):
    """Per-tank coarse space at its edges: 2-cell aggregates give the two large
    tanks more than the 256 aggregates a tank's dense block may hold, so the
    factor widens for the whole batch; a 2 x 2 x 2 droplet in mid-air between
    them has only Surface rows and no aggregate at all.  The batch must
    converge to the pressure of the Schwarz-only solve, with a coarse space no
    larger than 256 unknowns per tank, and match the oracle tank by tank.
    (A 2 x 2 x 2 block filling a 2 x 2 x 2 tank would not do: its rows are all
    Inner, so its PPE has no Dirichlet row and is singular.)"""
    l0 = 0.005
    x, y, z = cases.lattice_block(2, 2, 2, l0, origin=(6 * l0, 6 * l0, 6 * l0))
    drop_fluid, drop_walls = cases._fluid_dict(x, y, z), cases.tank_walls(16, 16, 16, l0)
    drop_cfg = cases.case_config(l0)
    cases.derive_cell_grid(drop_cfg, drop_fluid, drop_walls)
    tanks = [cases.block_in_tank("big0", (40, 40, 30), (80, 40, 60), l0, 0.05, 3),
             cases.Case("droplet", drop_fluid, drop_walls, drop_cfg, {"l0": l0}),
             cases.block_in_tank("big1", (36, 36, 36), (72, 36, 72), l0, 0.05, 4)]
    case = _batch_of(cases, tanks)
    prec = abi.PRECISION_FP64
    sch = abi.default_solver_options()
    sch.preconditioner = abi.PRECOND_SCHWARZ
    crs = abi.default_solver_options()
    crs.coarse_cells = 2
    gs = rt.run_step_on_device(case, precision=prec, solver=sch)
    gc = rt.run_step_on_device(case, precision=prec, solver=crs)
    print("schwarz", gs["report"], "coarse", gc["report"], flush=True)
    assert gs["report"]["converged"] == 1 and gc["report"]["converged"] == 1
    # the oracle's aggregation of each tank: the smallest factor >= 2 at which
    # no tank exceeds 256 aggregates is the batch's factor
    refs = [oracle.pipeline(t) for t in tanks]
    counts = {f: [oracle.coarse_aggregation(t, r["neighbors"], r["surface"]["surface_type"],
                                            factor=f)[1] for t, r in zip(tanks, refs)]
              for f in (2, 3, 4)}
    f_used = min(f for f in counts if max(counts[f]) <= 256)
    print("aggregates per factor", counts, "factor used", f_used, flush=True)
    assert f_used > 2 and counts[2][1] == 0
    assert gc["report"]["coarse_unknowns"] == sum(counts[f_used])
    fo = case.extra["fluid_offsets"]
    tiny = gc["surface_type"][fo[1]:fo[2]]
    assert np.all((tiny == abi.SURFACE) | (tiny == abi.SPLASH)), tiny
    assert rel_l2(gc["pressure"], gs["pressure"]) <= 1e-8
    assert gc["report"]["iterations"] < gs["report"]["iterations"]
    for k, (tank, ref) in enumerate(zip(tanks, refs)):
        a, b = fo[k], fo[k + 1]
        assert np.array_equal(gc["surface_type"][a:b], ref["surface"]["surface_type"]), k
        if np.any(ref["pressure"] != 0.0):
            assert rel_l2(gc["pressure"][a:b], ref["pressure"]) <= 1e-7, k
        else:
            assert np.all(gc["pressure"][a:b] == 0.0), k
