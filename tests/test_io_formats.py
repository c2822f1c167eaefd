"""On-disk formats (host code behind the C-ABI): INI simulation config, particle
CSV input and legacy ASCII VTK output, against the inputs and known answers of
the reference's tests/unit/smoke_config.cpp, smoke_file_manager.cpp and
smoke_vtk_writer.cpp, plus round trips."""
import math
import os

import numpy as np
import pytest

from helpers import parse_vtk


def test_config_ini_known_answers(rt, abi, golden, tmp_path):
    fx = golden["config"]
    d = abi.default_config()
    assert math.isclose(d.support_radius / d.particle_spacing,
                        fx["default_ratios"]["support_radius_ratio"], abs_tol=fx["tolerance"])
    assert abs(d.particle_spacing ** 3 - fx["default_ratios"]["particle_volume"]) <= fx["tolerance"]
    ini = tmp_path / "smoke.ini"
    ini.write_text(fx["ini"])
    cfg, files = rt.load_config(ini)
    for k, v in fx["expected"].items():
        assert abs(getattr(cfg, k) - v) <= fx["tolerance"], k
    for k, v in fx["expected_files"].items():
        got = getattr(files, k)
        assert (got.decode() if isinstance(got, bytes) else got) == v, k
    assert cfg.splash_neighbor_threshold == 12 and cfg.gravity[2] == -9.8
    assert cfg.include_wall_neighbors == 1 and files.output_directory == b"/tmp/lsmps3d_output"
    # save -> load round trip: every numeric field and every file field
    saved = tmp_path / "saved.ini"
    rt.save_config(saved, cfg, files)
    cfg2, files2 = rt.load_config(saved)
    for name, _ in abi.Config._fields_:
        if name in ("cell_origin", "cell_dims", "weight_kernel", "reserved0"):
            continue  # not INI keys the reference saves
        a, b = getattr(cfg, name), getattr(cfg2, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    for name, _ in abi.ConfigFiles._fields_:
        assert getattr(files, name) == getattr(files2, name), name
    text = saved.read_text()
    assert text.startswith("# lsmps3d simulation config\n\n[geometry]\nparticle_spacing=0.02\n")
    assert "[files]\nfluid_particle_file=input/fluid_particles.csv\n" in text


def test_config_validation_and_unknown_keys(rt, abi, golden, tmp_path):
    fx = golden["config"]
    rt.validate_config(abi.default_config(), rt.default_config_files())
    for change in fx["invalid"]:
        c = abi.default_config()
        for k, v in change.items():
            setattr(c, k, v)
        with pytest.raises(rt.LsmpsError) as e:
            rt.validate_config(c)
        assert e.value.code == abi.LSMPS_ERR_INVALID_ARGUMENT
    bad = tmp_path / "bad.ini"
    bad.write_text(fx["bad_ini"])
    with pytest.raises(rt.LsmpsError) as e:
        rt.load_config(bad)
    assert "Unknown simulation config key 'geometry.unknown_key' on line 2" in str(e.value)
    with pytest.raises(rt.LsmpsError) as e:
        rt.load_config(tmp_path / "missing.ini")
    assert e.value.code == abi.LSMPS_ERR_IO
    for text, msg in (("[nope]\na=1\n", "Unknown simulation config section 'nope'"),
                      ("[geometry]\nparticle_spacing\n", "expected key=value"),
                      ("[geometry]\nparticle_spacing=0.01x\n", "Invalid real value"),
                      ("[surface]\ninclude_wall_neighbors=maybe\n", "Invalid boolean value"),
                      ("[]\n", "empty section")):
        p = tmp_path / "case.ini"
        p.write_text(text)
        with pytest.raises(rt.LsmpsError) as e:
            rt.load_config(p)
        assert msg in str(e.value), (text, str(e.value))


def test_particle_csv_known_answers(rt, abi, golden, tmp_path):
    fx = golden["file_manager"]
    (tmp_path / "fluid.csv").write_text(fx["fluid_csv"])
    (tmp_path / "wall.csv").write_text(fx["wall_csv"])
    f = rt.read_fluid_csv(tmp_path / "fluid.csv")
    w = rt.read_wall_csv(tmp_path / "wall.csv")
    ex = fx["expected"]
    assert f["x"].shape[0] == ex["fluid_count"] and w["x"].shape[0] == ex["wall_count"]
    assert np.allclose(f["vx"], ex["fluid_vx"], atol=fx["tolerance"])
    assert np.array_equal(f["pressure"], ex["fluid_pressure"])
    assert list(f["surface_type"]) == ex["fluid_surface_type"]
    assert abs(w["normal_y"][0] - ex["wall_normal_y0"]) <= fx["tolerance"]
    assert abs(w["vz"][1] - ex["wall_vz1"]) <= fx["tolerance"]
    for text in (fx["bad_csv"], fx["extra_column_csv"]):
        p = tmp_path / "bad.csv"
        p.write_text(text)
        with pytest.raises(rt.LsmpsError) as e:
            rt.read_fluid_csv(p)
        assert e.value.code == abi.LSMPS_ERR_INVALID_ARGUMENT
    # output files named as FileManager::write_fluid_result / write_wall_result
    out = tmp_path / "vtk"
    pf = rt.write_vtk(out, fx["vtk_prefix"] + "_fluid", fx["vtk_step"], f["x"], f["y"], f["z"],
                      vectors=[("velocity", (f["vx"], f["vy"], f["vz"]))],
                      real_scalars=[("pressure", f["pressure"])],
                      int_scalars=[("surface_type", f["surface_type"])])
    pw = rt.write_vtk(out, fx["vtk_prefix"] + "_wall", fx["vtk_step"], w["x"], w["y"], w["z"],
                      vectors=[("normal", (w["normal_x"], w["normal_y"], w["normal_z"])),
                               ("velocity", (w["vx"], w["vy"], w["vz"]))])
    assert [os.path.basename(pf), os.path.basename(pw)] == fx["expected_files"]
    tf, tw = open(pf).read(), open(pw).read()
    assert all(t in tf for t in fx["expected_fluid_tokens"])
    assert all(t in tw for t in fx["expected_wall_tokens"])


def test_csv_round_trip_of_a_case(rt, cases, tmp_path):
    """A generated case written as the reference's CSV (with comments, a header
    and blank lines) reads back exactly (repr of doubles round-trips)."""
    case = cases.dam_break(block=(6, 5, 4), tank=(10, 5, 8), l0=0.01, seed=2)
    fl, wl = case.fluid, case.walls
    lines = ["# generated", "x,y,z,vx,vy,vz", ""]
    lines += [",".join(repr(float(fl[k][i])) for k in ("x", "y", "z", "vx", "vy", "vz"))
              for i in range(case.n_fluid)]
    (tmp_path / "f.csv").write_text("\n".join(lines) + "\n")
    wk = ("x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z")
    wlines = [" , ".join(repr(float(wl[k][i])) for k in wk) + "  # wall" for i in range(case.n_wall)]
    (tmp_path / "w.csv").write_text("\n".join(wlines) + "\n")
    f = rt.read_fluid_csv(tmp_path / "f.csv")
    w = rt.read_wall_csv(tmp_path / "w.csv")
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(f[k], fl[k]), k
    for k in wk:
        assert np.array_equal(w[k], wl[k]), k


def test_vtk_writer_known_answers(rt, golden, tmp_path):
    fx = golden["vtk_writer"]
    p = fx["points"]
    path = rt.write_vtk(tmp_path, fx["prefix"], fx["step"], p["x"], p["y"], p["z"],
                        vectors=list((k, v) for k, v in fx["vectors"].items()),
                        real_scalars=list(fx["real_scalars"].items()),
                        int_scalars=list(fx["int_scalars"].items()))
    assert os.path.basename(path) == fx["expected_file"]
    text = open(path).read()
    for t in fx["expected_tokens"]:
        assert t in text, t
    parsed = parse_vtk(path)
    assert np.allclose(parsed["points"], np.array([p["x"], p["y"], p["z"]]).T)
    for k, v in fx["vectors"].items():
        assert np.allclose(parsed["vectors"][k], np.array(v).T)
    for k, v in fx["real_scalars"].items():
        assert np.allclose(parsed["real"][k], v)
    for k, v in fx["int_scalars"].items():
        assert list(parsed["int"][k]) == v
    vx, vy, vz = fx["vectors"][fx["mismatched_vector"]]
    with pytest.raises(rt.LsmpsError):
        rt.write_vtk(tmp_path, fx["prefix"], fx["step"] + 1, p["x"], p["y"], p["z"],
                     vectors=[(fx["mismatched_vector"], (vx, vy, vz[:-1]))])
    # 9 significant digits, as the reference's setprecision(9)
    path = rt.write_vtk(tmp_path, "prec", 0, [1.0 / 3.0], [2.0e-10], [123456.789012345],
                        point_fields=False)
    assert "0.333333333 2e-10 123456.789\n" in open(path).read()
    assert "VECTORS" not in open(path).read()
