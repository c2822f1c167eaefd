"""Surface module on the GPU against the oracle, with every diagnostic output:
neighbor CSR, labels, neighbor counts and the virtual-light open-direction
counts bit-exact; the real diagnostics (number density, its ratio, anisotropy,
air-side open ratio and anisotropy, surface normal) within 1e-12 of the oracle
(both evaluate the same FP64 expressions in the same order).  Cases: a
hydrostatic box, a jittered dam-break, a splash, the five free-surface
geometries of the diagnostic driver, and variants with wall neighbors excluded
and a wider light cone."""
import numpy as np
import pytest

import diagnostic_geometries as dg
from helpers import make_case

pytestmark = pytest.mark.gpu

INT_KEYS = ("fluid_neighbor_count", "wall_neighbor_count", "surface_type")
REAL_KEYS = ("number_density", "number_density_ratio", "anisotropy", "air_open_ratio",
             "air_anisotropy", "surface_normal_x", "surface_normal_y", "surface_normal_z")


def _compare(rt, oracle, case, config=None):
    cfg = config or case.config
    g = rt.surface_on_device(case, config=cfg)
    nb = oracle.neighbors(case, config=cfg)
    ref = oracle.classify(case, nb, config=cfg, diagnostics=True)
    vl = oracle.virtual_light(case, nb, config=cfg)
    for k in ("fluid_offsets", "fluid_indices", "wall_offsets", "wall_indices"):
        assert np.array_equal(g[k], nb[k]), k
    for k in INT_KEYS:
        assert np.array_equal(g[k], ref[k]), k
    for k in REAL_KEYS:
        scale = max(float(np.max(np.abs(ref[k]))), 1e-300)
        assert float(np.max(np.abs(g[k] - ref[k]))) <= 1e-12 * scale, k
    assert np.array_equal(g["open_direction_count"], vl["open_direction_count"])
    assert np.array_equal(g["open_fraction"], vl["open_fraction"])
    return g


def _small_cases(cases):
    return [cases.hydrostatic(n=20, l0=0.01),
            cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0),
            cases.splash(n=24, l0=0.01, seed=3, top_layers=3)]


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_surface_diagnostics_and_virtual_light(rt, oracle, cases, idx):
    case = _small_cases(cases)[idx]
    print(f"\nsurface {case.name} n={case.n_fluid}", flush=True)
    g = _compare(rt, oracle, case)
    counts = np.bincount(g["surface_type"], minlength=4)
    print("labels", counts.tolist(), "open mean", float(g["open_fraction"].mean()), flush=True)


@pytest.mark.parametrize("idx", [1, 2])
def test_variants_without_walls_and_wide_cone(rt, oracle, cases, idx):
    case = _small_cases(cases)[idx]
    cfg = case.config.copy()
    cfg.include_wall_neighbors = 0
    cfg.virtual_light_cone_cosine = 0.5
    print(f"\nsurface variant {case.name} n={case.n_fluid}", flush=True)
    _compare(rt, oracle, case, config=cfg)


@pytest.mark.parametrize("name", dg.CASE_NAMES)
def test_free_surface_geometries(rt, oracle, cases, abi, name):
    f, _, w = dg.geometry(name)
    walls = {"x": w["x"], "y": w["y"], "z": w["z"], "normal_x": w["nx"], "normal_y": w["ny"],
             "normal_z": w["nz"]}
    case = make_case(cases, abi, name, (f["x"], f["y"], f["z"]), walls=walls,
                     config=dg.case_config(cases))
    print(f"\ngeometry {name} n={case.n_fluid} walls={case.n_wall}", flush=True)
    _compare(rt, oracle, case)


def test_virtual_light_at_250k(rt, oracle, cases):
    """Dam-break C2 (250,000 fluid): open-direction counts bit-exact."""
    case = cases.dam_break()
    print(f"\nvirtual light C2 n={case.n_fluid}", flush=True)
    _compare(rt, oracle, case)
