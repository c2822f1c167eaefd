"""Lanes-per-particle shapes of the neighbour-list kernels and the autotuner
(LSMPS_FLAG_AUTOTUNE), through the C-ABI.  Every shape computes the same step:
the neighbour tables are written in list order whatever the lane count, and the
floating-point sums follow a fixed slot tree, so the whole step -- CSR, labels,
PPE, pressure, correction -- is bit-identical across forced shapes and the
autotuner's own choice, which then only changes speed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shapes(abi, pick):
    """One lane count per tuned kernel: pick(candidates) chooses it."""
    return {k: pick(c) for k, c in abi.TUNED_CANDIDATES.items()}


def _same(a, b):
    for k in ("fluid_offsets", "fluid_indices", "wall_offsets", "wall_indices", "surface_type",
              "pressure"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("row_offsets", "col_indices", "values", "rhs"):
        assert np.array_equal(a["ppe"][k], b["ppe"][k]), k
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(a["fluid"][k], b["fluid"][k]), k
    for u, v in zip(a["grad"] + a["ps"], b["grad"] + b["ps"]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("name", ["dam_break", "splash"])
def test_shapes_give_identical_steps(rt, abi, cases, name):
    case = (cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0)
            if name == "dam_break" else cases.splash(n=24, l0=0.01, seed=3, top_layers=3))
    runs = {}
    for label, pick in (("narrow", min), ("wide", max), ("middle", lambda c: c[len(c) // 2])):
        shapes = _shapes(abi, pick)
        print(name, label, shapes, flush=True)
        g = rt.run_step_on_device(case, precision=abi.PRECISION_FP64, lanes=shapes)
        assert g["lanes"] == shapes, g["lanes"]
        runs[label] = g
    default = rt.run_step_on_device(case, precision=abi.PRECISION_FP64)
    for label, g in runs.items():
        _same(g, default)


def test_autotuner_choice(rt, abi, cases):
    case = cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0)
    default = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED)
    tuned = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED, flags=abi.FLAG_AUTOTUNE)
    print("tuned lanes", tuned["lanes"], "keys", tuned["tuned_keys"], flush=True)
    assert tuned["tuned_keys"] >= 5
    for k, c in abi.TUNED_CANDIDATES.items():
        assert tuned["lanes"][k] in c, k
    _same(tuned, default)
