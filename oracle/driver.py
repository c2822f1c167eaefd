"""Multi-step oracle: the simulation clock and the driver loop.

TEST INFRASTRUCTURE: only tests/ may import this module.

`Clock` restates SimulationTimeManager (src/core/time_step.cpp:79-170) in plain
Python; it is pinned by the known answers of tests/unit/smoke_time_step.cpp
(tests/golden/reference_unit_tests.json, "time_step").  `run` follows the
driver loop of src/app/simulation_driver.cu:340-391 with the FP64 oracle's one
step (orc_step, which rebuilds the geometry at the start of each step):
speed -> clock -> dt -> step.
"""
import math

import numpy as np


class Clock:
    def __init__(self, dt, dt_min, dt_max, growth, cfl, spacing, final_time, output_interval):
        for name, v in (("current_time_step", dt), ("min_time_step", dt_min),
                        ("max_time_step", dt_max), ("growth_factor", growth), ("cfl", cfl),
                        ("particle_spacing", spacing), ("output_interval", output_interval)):
            if not (math.isfinite(v) and v > 0.0):
                raise ValueError(f"TimeStepLimits field '{name}' must be positive and finite")
        if not (math.isfinite(final_time) and final_time >= 0.0):
            raise ValueError("TimeStepLimits field 'final_time' must be non-negative and finite")
        if growth < 1.0:
            raise ValueError("TimeStepLimits field 'growth_factor' must be >= 1")
        if dt_min > dt_max:
            raise ValueError("TimeStepLimits requires min_time_step <= max_time_step")
        self.dt_min, self.dt_max, self.growth, self.cfl = dt_min, dt_max, growth, cfl
        self.spacing, self.final_time, self.output_interval = spacing, final_time, output_interval
        self.time = 0.0
        self.dt = min(max(dt, dt_min), dt_max)
        self.next_output = 0.0
        self.step_index = 0
        self.output_index = 0
        self.initial_pending = True

    @classmethod
    def from_config(cls, c):
        return cls(c.time_step, c.min_time_step, c.max_time_step, c.time_step_growth_factor,
                   c.cfl, c.particle_spacing, c.final_time, c.output_interval)

    def finished(self):
        return self.time >= self.final_time

    def output_due(self):
        return self.initial_pending or self.time >= self.next_output

    def cfl_dt(self, vmax):
        if not (vmax > 0.0) or not math.isfinite(vmax):
            return self.dt_max
        return self.cfl * self.spacing / vmax

    def next_dt(self, vmax):
        dt = min(max(min(self.dt * self.growth, self.cfl_dt(vmax)), self.dt_min), self.dt_max)
        remaining = self.final_time - self.time
        if remaining > 0.0 and dt > remaining:
            dt = remaining
        return dt

    def advance(self, vmax):
        prev = self.time
        cdt = self.cfl_dt(vmax)
        dt = self.next_dt(vmax)
        self.time += dt
        self.dt = dt
        self.step_index += 1
        out = False
        if self.initial_pending or self.time >= self.next_output:
            out = True
            self.initial_pending = False
            self.output_index += 1
            while self.next_output <= self.time:
                self.next_output += self.output_interval
        return {"step_index": self.step_index, "output_index": self.output_index,
                "previous_time": prev, "current_time": self.time, "time_step": dt,
                "cfl_time_step": cdt, "max_velocity": vmax, "should_output": int(out),
                "reached_final_time": int(self.finished())}

    def mark_initial_output(self):
        out = False
        if self.initial_pending:
            out = True
            self.initial_pending = False
            self.output_index += 1
            self.next_output = self.output_interval
        return {"step_index": self.step_index, "output_index": self.output_index,
                "should_output": int(out), "reached_final_time": int(self.finished())}


def max_velocity(fluid):
    """Largest finite |v| (simulation_driver.cu:122-143)."""
    s = np.sqrt(fluid["vx"] ** 2 + fluid["vy"] ** 2 + fluid["vz"] ** 2)
    s = s[np.isfinite(s)]
    return float(s.max()) if s.size else 0.0


def run(orc, case, max_steps, options=None):
    """Oracle steps of `case` under its clock; returns (fluid, statuses)."""
    c = case.copy()
    clock = Clock.from_config(c.config)
    clock.mark_initial_output()
    statuses = []
    while not clock.finished() and len(statuses) < max_steps:
        st = clock.advance(max_velocity(c.fluid))
        c.config.time_step = st["time_step"]
        fluid, _, _ = orc.step(c, options)
        c.fluid = fluid
        statuses.append(st)
    return c.fluid, statuses
