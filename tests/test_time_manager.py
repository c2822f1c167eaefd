"""The simulation clock (lsmps_time_manager_*, host code behind the C-ABI) and
its oracle restatement against the known answers of the reference's
tests/unit/smoke_time_step.cpp."""
import pytest

from oracle.driver import Clock


def _config(abi, fx):
    cfg = abi.default_config()
    for k, v in fx["config"].items():
        setattr(cfg, k, v)
    return cfg




def test_library_clock_matches_reference_known_answers(rt, abi, golden):
    fx = golden["time_step"]
    clock = rt.TimeManager(_config(abi, fx))
    assert clock.state()["output_due"] == fx["initial_output_due"]
    init = clock.mark_initial_output()
    assert init["should_output"] == 1 and init["output_index"] == fx["initial_output_index"]
    assert not clock.state()["output_due"]
    tol = fx["tolerance"]
    st = None
    for step in fx["advances"]:
        st = clock.advance(step["max_velocity"])
        assert abs(st["time_step"] - step["expected_dt"]) <= tol, step["label"]
        if "should_output" in step:
            assert bool(st["should_output"]) == step["should_output"], step["label"]
        if "reached_final_time" in step:
            assert bool(st["reached_final_time"]) == step["reached_final_time"], step["label"]
        if "output_index" in step:
            assert st["output_index"] == step["output_index"], step["label"]
    guard = 0
    while not st["reached_final_time"] and guard < fx["guard"]:
        st = clock.advance(fx["then_max_velocity"])
        guard += 1
    assert st["reached_final_time"]
    assert abs(clock.state()["current_time"] - fx["expected_final_time"]) <= tol


def test_library_clock_rejects_invalid_limits(rt, abi, golden):
    fx = golden["time_step"]
    lim = abi.TimeLimits()
    rt.check(rt.load().lsmps_time_limits_from_config(rt.C.byref(_config(abi, fx)), rt.C.byref(lim)))
    rt.TimeManager.validate(lim)
    lim.growth_factor = fx["invalid_growth_factor"]
    with pytest.raises(rt.LsmpsError) as e:
        rt.TimeManager.validate(lim)
    assert e.value.code == abi.LSMPS_ERR_INVALID_ARGUMENT
    assert "growth_factor" in str(e.value)
    lim.growth_factor = 2.0
    lim.min_time_step, lim.max_time_step = 1.0, 0.5
    with pytest.raises(rt.LsmpsError):
        rt.TimeManager(limits=lim)


def test_oracle_clock_matches_reference_known_answers(abi, golden):
    fx = golden["time_step"]
    clock = Clock.from_config(_config(abi, fx))
    assert clock.output_due()
    init = clock.mark_initial_output()
    assert init["should_output"] == 1 and init["output_index"] == fx["initial_output_index"]
    tol = fx["tolerance"]
    st = None
    for step in fx["advances"]:
        st = clock.advance(step["max_velocity"])
        assert abs(st["time_step"] - step["expected_dt"]) <= tol, step["label"]
        if "output_index" in step:
            assert st["output_index"] == step["output_index"]
    guard = 0
    while not st["reached_final_time"] and guard < fx["guard"]:
        st = clock.advance(fx["then_max_velocity"])
        guard += 1
    assert abs(clock.time - fx["expected_final_time"]) <= tol
    with pytest.raises(ValueError):
        Clock(0.001, 1e-4, 1e-2, fx["invalid_growth_factor"], 0.25, 0.02, 0.012, 0.004)


def test_library_and_oracle_clocks_agree_on_a_speed_series(rt, abi, golden):
    """Random speeds: the two restatements give the same statuses bit for bit."""
    import numpy as np
    fx = golden["time_step"]
    cfg = _config(abi, fx)
    cfg.final_time = 0.5
    lib, orc = rt.TimeManager(cfg), Clock.from_config(cfg)
    assert lib.mark_initial_output()["output_index"] == orc.mark_initial_output()["output_index"]
    speeds = np.random.default_rng(3).lognormal(0.0, 1.5, 400)
    speeds[::37] = 0.0
    speeds[::53] = np.inf
    for v in speeds:
        a, b = lib.advance(v), orc.advance(v)
        for k in ("step_index", "output_index", "should_output", "reached_final_time"):
            assert a[k] == b[k]
        for k in ("time_step", "current_time", "cfl_time_step"):
            assert a[k] == b[k], (k, a[k], b[k])
        if a["reached_final_time"]:
            break
    assert a["reached_final_time"]
