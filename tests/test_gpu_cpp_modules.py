"""The C++ module API (include/lsmps3d_amd.hpp, the reference's class and
function names) driven by a compiled program, tests/cpp/module_step.cpp, that
steps a case module by module as the reference driver does.  Its outputs are
compared with the FP64 oracle (same bars as the fused-step parity tests), and
the correction replayed through the free functions must equal
DevicePressureCorrection::apply bit for bit.  The reference-signature free
functions that read MomentMatrixViews, and the explicit inverses of all four
moment kinds (velocity/pressure Type-A, pressure Type-B, fluid-only Type-A),
are checked against the oracle as well."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "_build", "module_step")
FLUID = ("x", "y", "z", "vx", "vy", "vz", "pressure")
WALL = ("x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z")


def _run(case, abi, tmp_path, precision, max_neighbors=256):
    if not os.path.exists(BINARY):
        pytest.fail(f"{BINARY} is not built (run __graft_entry__.build())")
    d = str(tmp_path)
    with open(os.path.join(d, "sizes.txt"), "w") as f:
        f.write(f"{case.n_fluid} {case.n_wall} {max_neighbors} {precision}\n")
    with open(os.path.join(d, "config.bin"), "wb") as f:
        f.write(bytes(case.config))
    for k in FLUID:
        np.ascontiguousarray(case.fluid[k], np.float64).tofile(os.path.join(d, f"fluid_{k}.bin"))
    np.ascontiguousarray(case.fluid["surface_type"], np.int32).tofile(
        os.path.join(d, "fluid_surface_type.bin"))
    for k in WALL:
        np.ascontiguousarray(case.walls[k], np.float64).tofile(os.path.join(d, f"wall_{k}.bin"))
    print(f"module_step {case.name} n={case.n_fluid} precision={precision}", flush=True)
    res = subprocess.run([BINARY, d], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr

    def out(name, dtype=np.float64):
        return np.fromfile(os.path.join(d, f"out_{name}.bin"), dtype=dtype)
    return out


def _cases(cases):
    return [cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0),
            cases.splash(n=24, l0=0.01, seed=3, top_layers=3)]


@pytest.mark.parametrize("idx", [0, 1])
def test_cpp_module_api_fp64(abi, cases, oracle, tmp_path, idx):
    case = _cases(cases)[idx]
    n, m = case.n_fluid, case.n_wall
    out = _run(case, abi, tmp_path, abi.PRECISION_FP64)
    ref = oracle.pipeline(case)
    nb = ref["neighbors"]
    i32 = np.int32
    for k in ("cell_ids", "sorted_indices", "cell_begin", "cell_end"):
        assert np.array_equal(out(f"fluid_{k}", i32), nb["fluid_cells"][k]), k
    assert np.array_equal(out("fluid_offsets", i32), nb["fluid_offsets"])
    assert np.array_equal(out("fluid_indices", i32), nb["fluid_indices"])
    assert np.array_equal(out("wall_offsets", i32), nb["wall_offsets"])
    assert np.array_equal(out("wall_indices", i32), nb["wall_indices"])
    assert np.array_equal(out("surface_type", i32), ref["surface"]["surface_type"])
    tv = out("temporary_velocity").reshape(3, n)
    tw = out("temporary_wall_velocity").reshape(3, m)
    pv = ref["provision"]
    for k, key in enumerate(("vx", "vy", "vz")):
        assert rel_l2(tv[k], pv[key]) <= 1e-9, key
        assert rel_l2(tw[k], pv["wall_" + key]) <= 1e-9, key
    pp = ref["ppe"]
    assert np.array_equal(out("ppe_row_offsets", i32), pp["row_offsets"])
    assert np.array_equal(out("ppe_col_indices", i32), pp["col_indices"])
    assert rel_l2(out("ppe_values"), pp["values"]) <= 1e-9
    assert rel_l2(out("ppe_rhs"), pp["rhs"]) <= 1e-9
    assert rel_l2(out("pressure"), ref["pressure"]) <= 1e-9
    cg = ref["correction"]
    grad = out("grad").reshape(3, n)
    num = sum(np.sum((grad[k] - cg["grad"][k]) ** 2) for k in range(3))
    den = sum(np.sum(cg["grad"][k] ** 2) for k in range(3))
    assert np.sqrt(num / den) <= 1e-8
    for k in ("x", "y", "z"):
        assert rel_l2(out(f"fluid_{k}"), cg["fluid"][k]) <= 1e-12
    scale = max(np.linalg.norm(pv["vz"]), 1e-300)
    for k in ("vx", "vy", "vz"):
        assert np.linalg.norm(out(f"fluid_{k}") - cg["fluid"][k]) / scale <= 1e-8, k
    # the reference-signature free functions reading MomentMatrixViews
    # (ppe_matrix.cuh:120-129, pressure_correction.cuh:93-100)
    assert np.array_equal(out("view_ppe_row_offsets", i32), pp["row_offsets"])
    assert np.array_equal(out("view_ppe_col_indices", i32), pp["col_indices"])
    assert rel_l2(out("view_ppe_values"), pp["values"]) <= 1e-9
    assert rel_l2(out("view_ppe_rhs"), pp["rhs"]) <= 1e-9
    vgrad = out("view_grad").reshape(3, n)
    num = sum(np.sum((vgrad[k] - cg["grad"][k]) ** 2) for k in range(3))
    assert np.sqrt(num / den) <= 1e-8
    # all four moment views against the oracle's explicit inverses
    c2 = case.copy()
    c2.fluid["surface_type"] = ref["surface"]["surface_type"]
    kinds = {"velocity_a": abi.MOMENT_VELOCITY_WALL_DIRICHLET_A,
             "pressure_a": abi.MOMENT_PRESSURE_WALL_NEUMANN_A,
             "pressure_b": abi.MOMENT_PRESSURE_WALL_NEUMANN_B,
             "fluid_only_a": abi.MOMENT_FLUID_ONLY_A}
    for name, kind in kinds.items():
        mo = oracle.moments(c2, nb, kind)
        size, basis = out(f"view_{name}_shape", i32)
        assert size == mo["size"] and basis == (1 if name == "pressure_b" else 0), name
        inv = out(f"view_{name}_inverse").reshape(n, size, size)
        err = rel_l2(inv, mo["inverse"])
        print(name, "inverse rel L2", err, flush=True)
        assert err <= 1e-7, name
        assert rel_l2(out(f"view_{name}_trace"), mo["trace"]) <= 1e-12, name
        assert rel_l2(out(f"view_{name}_regularization"), mo["regularization"]) <= 1e-12, name
        assert np.array_equal(out(f"view_{name}_info", i32), mo["info"]), name
    # the free functions in apply's order reproduce apply exactly
    for k in FLUID:
        assert np.array_equal(out(f"replay_fluid_{k}"), out(f"fluid_{k}")), k
    assert np.array_equal(out("replay_grad"), out("grad"))
    assert np.array_equal(out("replay_ps"), out("ps"))


def test_cpp_module_api_mixed(abi, cases, oracle, tmp_path):
    case = _cases(cases)[0]
    out = _run(case, abi, tmp_path, abi.PRECISION_MIXED)
    ref = oracle.pipeline(case)
    assert np.array_equal(out("fluid_indices", np.int32), ref["neighbors"]["fluid_indices"])
    assert np.array_equal(out("surface_type", np.int32), ref["surface"]["surface_type"])
    assert rel_l2(out("pressure"), ref["pressure"]) <= 1e-5
