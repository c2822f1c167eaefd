"""Multi-step driver on the GPU (lsmps_stepper_run, lsmps_max_velocity) against
the oracle's driver loop (oracle/driver.py over orc_step).

Bars: the device speed reduction equals numpy's max exactly (it is a max of
correctly rounded magnitudes); the dt sequence of an adaptive run agrees to
1e-8 relative in FP64 mode and 1e-7 in mixed mode (dt follows the CFL limit of the step-start speed, whose value
differs between the two runs only by the PPE tolerance); positions after 11
steps (growth-limited, then CFL-limited dt) within 1e-9 relative L2; every step is reported to the output callback
with the refreshed labels equal to the oracle's classification of the same
positions except where the two runs' positions straddle a threshold (none
expected; at most 0.1 % allowed).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_max_velocity_matches_numpy(rt, abi, pkg):
    case = pkg.cases.dam_break(block=(20, 10, 10), tank=(40, 10, 20), l0=0.01)
    n = case.n_fluid
    rng = np.random.default_rng(4)
    for k in ("vx", "vy", "vz"):
        case.fluid[k] = rng.standard_normal(n)
    case.fluid["vx"][7] = np.inf
    case.fluid["vy"][11] = np.nan
    ctx = rt.Context(abi.PRECISION_FP64)
    state = rt.DeviceState(case)
    got = rt.max_velocity(ctx, state)
    from oracle.driver import max_velocity
    assert got == max_velocity(case.fluid)
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_driver_run_matches_oracle_loop(oracle, rt, abi, pkg, precision):
    from oracle import driver
    case = pkg.cases.dam_break(block=(16, 10, 10), tank=(32, 10, 20), l0=0.01, seed=0)
    # a gentle shear flow vx = 0.05 z / H; with CFL 0.02 its limit (4 ms) takes
    # over from growth at step 10
    z = case.fluid["z"]
    case.fluid["vx"] = 0.05 * z / z.max()
    cfg = case.config
    cfg.time_step = 1.0e-4
    cfg.min_time_step = 1.0e-5
    cfg.max_time_step = 2.0e-2
    cfg.time_step_growth_factor = 1.5
    cfg.cfl = 0.02
    cfg.final_time = 1.0
    cfg.output_interval = 1.0e-9  # every step is an output step
    steps = 11
    print(f"\ndriver run n={case.n_fluid} precision={precision} steps={steps}", flush=True)
    ref_fluid, ref_status = driver.run(oracle, case, steps)
    ctx = rt.Context(precision)
    state = rt.DeviceState(case)
    st = rt.Stepper(ctx, state, cfg, 256, abi.default_solver_options())
    clock = rt.TimeManager(cfg)
    seen = []

    def on_output(status):
        seen.append(status)
        return False

    last, rc = st.run(clock, max_steps=steps, on_output=on_output)
    assert rc == abi.LSMPS_OK
    assert seen[0]["step_index"] == 0 and seen[0]["output_index"] == 1  # initial output
    got = seen[1:]
    assert [s["step_index"] for s in got] == list(range(1, steps + 1))
    assert last["step_index"] == steps
    dts = np.array([s["time_step"] for s in got])
    ref_dts = np.array([s["time_step"] for s in ref_status])
    print("dt", dts.tolist(), "oracle", ref_dts.tolist(), flush=True)
    assert np.allclose(dts, ref_dts, rtol=1e-8 if precision == 0 else 1e-7, atol=0.0)
    assert dts[0] > cfg.time_step  # the clock grew dt
    cfl_limited = [s for s in got if s["time_step"] == s["cfl_time_step"]]
    assert cfl_limited, "no step reached the CFL limit"
    f = state.fluid_host()
    pos = np.stack([f["x"], f["y"], f["z"]])
    ref_pos = np.stack([ref_fluid["x"], ref_fluid["y"], ref_fluid["z"]])
    err = rel_l2(pos, ref_pos)
    disp = rel_l2(pos, np.stack([case.fluid["x"], case.fluid["y"], case.fluid["z"]]))
    print("positions rel L2", err, "displacement", disp, flush=True)
    assert err <= (1e-9 if precision == 0 else 1e-7)
    # labels refreshed at the final positions vs the oracle's classification there
    c2 = case.copy()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        c2.fluid[k] = f[k].copy()
    nb = oracle.neighbors(c2)
    labels = oracle.classify(c2, nb, diagnostics=False)["surface_type"]
    mism = int((labels != f["surface_type"]).sum())
    print("label mismatches", mism, flush=True)
    assert mism <= max(1, case.n_fluid // 1000)
    st.close()
    clock.close()
    ctx.close()
