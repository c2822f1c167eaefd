"""End to end: the application program (tests/cpp/run_simulation.cpp, i.e.
lsmps3d::run_simulation over the library) reads an INI config and particle CSV
files written here, runs the clock-driven step loop on the GPU and writes VTK
files; the last fluid file's positions are compared with the oracle's driver
loop over the same steps (the VTK text carries 9 significant digits)."""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_csv(path, arrays, keys):
    with open(path, "w") as fh:
        fh.write("# generated\n" + ",".join(keys) + "\n")
        for i in range(arrays[keys[0]].shape[0]):
            fh.write(",".join(repr(float(arrays[k][i])) for k in keys) + "\n")


def _read_vtk_points(path):
    lines = open(path).read().split("\n")
    i = next(k for k, t in enumerate(lines) if t.startswith("POINTS"))
    n = int(lines[i].split()[1])
    return np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)]).T


def test_run_simulation_program(oracle, rt, abi, pkg, tmp_path):
    from oracle import driver
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "run_simulation")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    case = pkg.cases.dam_break(block=(16, 10, 10), tank=(32, 10, 20), l0=0.01, seed=0)
    cfg = case.config
    cfg.time_step = 1.0e-4
    cfg.min_time_step = 1.0e-5
    cfg.max_time_step = 2.0e-3
    cfg.time_step_growth_factor = 1.5
    cfg.final_time = 1.0e-3
    cfg.output_interval = 2.5e-4
    _write_csv(tmp_path / "fluid.csv", case.fluid, ["x", "y", "z", "vx", "vy", "vz"])
    _write_csv(tmp_path / "wall.csv", case.walls,
               ["x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z"])
    files = rt.default_config_files()
    files.fluid_particle_file = str(tmp_path / "fluid.csv").encode()
    files.wall_particle_file = str(tmp_path / "wall.csv").encode()
    files.output_directory = str(tmp_path / "out").encode()
    files.vtk_file_prefix = b"dam"
    ini = tmp_path / "simulation.ini"
    rt.save_config(ini, cfg, files)
    print(f"\nrun_simulation n={case.n_fluid}", flush=True)
    run = subprocess.run([exe, str(ini)], capture_output=True, text=True, timeout=120)
    print(run.stdout[-2000:], run.stderr[-2000:], flush=True)
    assert run.returncode == 0
    assert "Simulation complete" in run.stdout
    fluid_files = sorted(glob.glob(str(tmp_path / "out" / "dam_fluid_*.vtk")))
    wall_files = glob.glob(str(tmp_path / "out" / "dam_wall_*.vtk"))
    assert len(wall_files) == 1 and os.path.basename(fluid_files[0]) == "dam_fluid_000000.vtk"
    ref_fluid, ref_status = driver.run(oracle, case, 100)
    steps = len(ref_status)
    assert ref_status[-1]["reached_final_time"]
    last = os.path.basename(fluid_files[-1])
    assert last == f"dam_fluid_{steps:06d}.vtk", (last, steps)
    pos = _read_vtk_points(fluid_files[-1])
    ref = np.stack([ref_fluid["x"], ref_fluid["y"], ref_fluid["z"]])
    err = rel_l2(pos, ref)
    print("steps", steps, "final positions rel L2 vs oracle loop", err, flush=True)
    assert err <= 1e-8
    text = open(fluid_files[-1]).read()
    for token in ("VECTORS velocity float", "SCALARS pressure float 1", "SCALARS surface_type int 1",
                  "VECTORS pressure_gradient float", "VECTORS ps_displacement float"):
        assert token in text
