"""The diagnostic drivers (tests/cpp/*_diagnostics, over the library's C++
module API) run on the GPU; their files and printed summaries are checked
against the oracle on the same particles (rebuilt bit for bit by
tests/diagnostic_geometries.py) and against closed forms.

The CSV and VTK text carries 9 significant digits, so real fields are compared
at 1e-8 relative to the field's scale; labels, neighbor counts and the printed
counts are exact."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import diagnostic_geometries as dg
from conftest import rel_l2
from helpers import make_case, parse_vtk

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")
CSV_TOL = 1e-8


def _run(name, *args):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), "run __graft_entry__.build()"
    print(f"\n{name} {' '.join(map(str, args))}", flush=True)
    res = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    print(res.stdout[-3000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    return res.stdout


def _read_csv(path):
    with open(path) as fh:
        header = fh.readline().strip().split(",")
        rows = [line.rstrip("\n").split(",") for line in fh]
    return {k: [r[i] for r in rows] for i, k in enumerate(header)}


def _near(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = max(float(np.max(np.abs(ref))), 1e-300)
    return float(np.max(np.abs(got - ref))) <= CSV_TOL * scale


def _stat(text, label):
    m = re.search(re.escape(label) + r":\s*([-+0-9.eE]+)", text)
    assert m, label
    return float(m.group(1))


def _section(stdout, name):
    start = stdout.index(f"Complex surface diagnostics: {name}")
    nxt = stdout.find("Complex surface diagnostics:", start + 1)
    return stdout[start:nxt if nxt >= 0 else len(stdout)]


def test_complex_surface_program(rt, oracle, cases, abi, tmp_path):
    out = _run("complex_surface_diagnostics", "all", tmp_path)
    for name in dg.CASE_NAMES:
        f, expected, w = dg.geometry(name)
        walls = {"x": w["x"], "y": w["y"], "z": w["z"], "normal_x": w["nx"],
                 "normal_y": w["ny"], "normal_z": w["nz"]}
        case = make_case(cases, abi, name, (f["x"], f["y"], f["z"]), walls=walls,
                         config=dg.case_config(cases))
        nb = oracle.neighbors(case)
        ref = oracle.classify(case, nb, diagnostics=True)
        csv = _read_csv(tmp_path / name / f"{name}_debug.csv")
        n = case.n_fluid
        assert len(csv["particle_id"]) == n, name
        for k in ("x", "y", "z"):
            assert [float(v) for v in csv[k]] == [float(f"{v:.9g}") for v in f[k]], (name, k)
        st = np.array([int(v) for v in csv["surface_type"]])
        assert np.array_equal(st, ref["surface_type"]), name
        for k in ("fluid_neighbor_count", "wall_neighbor_count"):
            assert np.array_equal(np.array([int(v) for v in csv[k]]), ref[k]), (name, k)
        for k in ("number_density", "number_density_ratio", "anisotropy", "air_open_ratio",
                  "air_anisotropy", "surface_normal_x", "surface_normal_y", "surface_normal_z"):
            assert _near([float(v) for v in csv[k]], ref[k]), (name, k)
        assert np.array_equal(np.array([int(v) for v in csv["expected_surface"]]), expected)
        vtk = parse_vtk(tmp_path / name / f"{name}_000000.vtk")
        assert np.array_equal(vtk["int"]["surface_type"], ref["surface_type"])
        wvtk = parse_vtk(tmp_path / name / f"{name}_walls_000000.vtk")
        assert wvtk["points"].shape[0] == case.n_wall
        # the printed statistics from the oracle's labels and normals
        text = _section(out, name)
        labels = ref["surface_type"]
        counts = np.bincount(labels, minlength=4)
        for k, label in enumerate(("Inner", "NearSurface", "Surface", "Splash")):
            assert _stat(text, "  " + label) == counts[k], (name, label)
        want = expected == 1
        got = labels == 2
        assert _stat(text, "Expected-surface particles") == int(want.sum())
        assert _stat(text, "Expected-surface detected") == int((want & got).sum())
        assert _stat(text, "False Surface particles") == int((~want & got).sum())
        both = want & got
        dot = (ref["surface_normal_x"] * f["nx"] + ref["surface_normal_y"] * f["ny"]
               + ref["surface_normal_z"] * f["nz"])[both]
        angle = float(np.mean(np.degrees(np.arccos(np.clip(dot, -1, 1))))) if both.any() else 0.0
        assert abs(_stat(text, "Mean normal angle error") - angle) <= 1e-5 * max(angle, 1.0)
        print(name, counts.tolist(), int(want.sum()), int(both.sum()), round(angle, 3), flush=True)


def _pressure_operator(case, nb, inverse, p, wall_scale):
    """grad p and lap p from the pressure Type-A inverses (numpy, FP64)."""
    f, w = case.fluid, case.walls
    re_ = case.config.support_radius
    rho = case.config.density
    g = np.array(case.config.gravity[:])
    n = case.n_fluid
    rhs = np.zeros((n, 9))
    for off, idx, src, is_wall in ((nb["fluid_offsets"], nb["fluid_indices"], f, False),
                                  (nb["wall_offsets"], nb["wall_indices"], w, True)):
        rows = np.repeat(np.arange(n), np.diff(off))
        cols = idx[:off[-1]]
        d = np.stack([src[k][cols] - f[k][rows] for k in ("x", "y", "z")], axis=1)
        r = np.linalg.norm(d, axis=1)
        wt = np.where(r >= re_, 0.0, 1.0 - r / re_)
        s = d / re_
        if not is_wall:
            vec = np.stack([s[:, 0], s[:, 1], s[:, 2], s[:, 0] ** 2, s[:, 1] ** 2, s[:, 2] ** 2,
                            s[:, 0] * s[:, 1], s[:, 1] * s[:, 2], s[:, 2] * s[:, 0]], axis=1)
            coef = wt * (p[cols] - p[rows])
        else:
            nrm = np.stack([w[k][cols] for k in ("normal_x", "normal_y", "normal_z")], axis=1)
            vec = np.stack([nrm[:, 0], nrm[:, 1], nrm[:, 2], 2 * s[:, 0] * nrm[:, 0],
                            2 * s[:, 1] * nrm[:, 1], 2 * s[:, 2] * nrm[:, 2],
                            s[:, 1] * nrm[:, 0] + s[:, 0] * nrm[:, 1],
                            s[:, 2] * nrm[:, 1] + s[:, 1] * nrm[:, 2],
                            s[:, 0] * nrm[:, 2] + s[:, 2] * nrm[:, 0]], axis=1)
            coef = wall_scale * wt * re_ * rho * (nrm @ g)
        keep = wt > 0
        np.add.at(rhs, rows[keep], vec[keep] * coef[keep][:, None])
    minv = inverse.reshape(n, 9, 9).transpose(0, 2, 1)  # column-major per particle
    c = np.einsum("nij,nj->ni", minv, rhs)
    return c[:, 0] / re_, c[:, 1] / re_, c[:, 2] / re_, 2 * (c[:, 3] + c[:, 4] + c[:, 5]) / re_ ** 2


def test_hydrostatic_program(rt, oracle, cases, abi, tmp_path):
    out = _run("hydrostatic_surface_diagnostics", tmp_path)
    f, w = dg.hydrostatic_box()
    walls = {"x": w["x"], "y": w["y"], "z": w["z"], "normal_x": w["nx"], "normal_y": w["ny"],
             "normal_z": w["nz"]}
    cfg = dg.case_config(cases, anisotropy_threshold=0.05)
    cfg.density = 1000.0
    cfg.gravity[0], cfg.gravity[1], cfg.gravity[2] = 0.0, 0.0, -9.81
    surface_z = 0.5 - 0.5 * dg.SPACING
    p_true = 1000.0 * 9.81 * np.maximum(surface_z - f["z"], 0.0)
    case = make_case(cases, abi, "hydrostatic_box", (f["x"], f["y"], f["z"]), walls=walls,
                     config=cfg, pressure=p_true)
    n = case.n_fluid
    assert n == 62500
    nb = oracle.neighbors(case)
    cls = oracle.classify(case, nb, diagnostics=True)
    st = cls["surface_type"]
    labelled = case.copy()
    labelled.fluid["surface_type"] = st
    vel = oracle.moments(labelled, nb, abi.MOMENT_VELOCITY_WALL_DIRICHLET_A)
    pre = oracle.moments(labelled, nb, abi.MOMENT_PRESSURE_WALL_NEUMANN_A)
    tmp = oracle.provision(labelled, nb, vel["inverse"])
    ppe = oracle.assemble(labelled, nb, tmp, vel["inverse"], pre["inverse"])
    pressure = oracle.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], ppe["rhs"])[0]

    vtk = parse_vtk(tmp_path / "hydrostatic_surface_000000.vtk")
    R, I = vtk["real"], vtk["int"]
    assert np.array_equal(I["surface_type"], st)
    assert np.array_equal(I["fluid_neighbor_count"], cls["fluid_neighbor_count"])
    assert np.array_equal(I["wall_neighbor_count"], cls["wall_neighbor_count"])
    rows = ppe["row_offsets"]
    assert np.array_equal(I["ppe_matrix_row_nnz"], np.diff(rows))
    assert _near(R["hydrostatic_pressure"], p_true)
    for k, ref in (("ppe_rhs", ppe["rhs"]), ("ppe_velocity_divergence", ppe["divergence"]),
                   ("ppe_wall_neumann_rhs", ppe["pressure_laplacian"])):
        assert _near(R[k], ref), k
    rid = np.repeat(np.arange(n), np.diff(rows))
    vals = ppe["values"][:rows[-1]]
    cols = ppe["col_indices"][:rows[-1]]
    assert _near(R["ppe_matrix_row_sum"], np.bincount(rid, weights=vals, minlength=n))
    diag = np.zeros(n)
    on = cols == rid
    diag[rid[on]] = vals[on]
    assert _near(R["ppe_matrix_diagonal"], diag)
    ap = np.bincount(rid, weights=vals * p_true[cols], minlength=n)
    resid = ap - ppe["rhs"]
    # A p_true - b cancels to ~1e-6 of its terms: compare on the terms' scale
    terms = np.bincount(rid, weights=np.abs(vals * p_true[cols]), minlength=n) + np.abs(ppe["rhs"])
    assert float(np.max(np.abs(R["ppe_hydrostatic_residual"] - resid))) <= CSV_TOL * float(terms.max())
    assert _near(R["ppe_hydrostatic_residual_plus_wall_rhs"], resid + ppe["pressure_laplacian"])
    assert _near(R["ppe_hydrostatic_residual_minus_two_wall_rhs"],
                 resid - 2.0 * ppe["pressure_laplacian"])
    # the solved pressure (both solves to a 1e-10 relative residual)
    assert rel_l2(R["ppe_pressure"], pressure) <= 1e-7
    gx, gy, gz, lap = _pressure_operator(case, nb, pre["inverse"], pressure,
                                         cfg.lsmps_wall_weight_scale)
    # natural scales: rho g for the gradient, rho g / re for the Laplacian (the
    # two solves differ by ~1e-9 relative, which the Laplacian magnifies by 1/re^2)
    scale_g = 1000.0 * 9.81
    for k, ref in (("ppe_pressure_gradient_x", gx), ("ppe_pressure_gradient_y", gy),
                   ("ppe_pressure_gradient_z", gz)):
        assert float(np.max(np.abs(R[k] - ref))) <= 1e-6 * scale_g, k
    scale_l = float(np.max(np.abs(lap))) + scale_g / cfg.support_radius
    assert float(np.max(np.abs(R["ppe_pressure_laplacian"] - lap))) <= 1e-6 * scale_l
    # printed summary
    counts = np.bincount(st, minlength=4)
    for k, label in enumerate(("Inner", "NearSurface", "Surface", "Splash")):
        assert _stat(out, "  " + label) == counts[k], label
    top = f["z"] >= 0.5 - 0.75 * dg.SPACING
    assert _stat(out, "Top-layer fluid particles") == int(top.sum())
    assert _stat(out, "Top-layer Surface particles") == int((top & (st == 2)).sum())
    assert _stat(out, "Below-top Surface particles") == int((~top & (st == 2)).sum())
    n0 = oracle.reference_number_density(dg.SPACING, 3.1 * dg.SPACING)
    assert abs(_stat(out, "reference number density n0") - n0) <= 1e-5 * n0
    print("labels", counts.tolist(), "max |p - p_hydro|", float(np.max(np.abs(pressure - p_true))),
          flush=True)
    # closed form: the hydrostatic field p = rho g (H - l0/2 - z) is linear in z,
    # zero on the top (Surface, Dirichlet) layer and has dp/dn = rho g.n on the
    # walls, so LSMPS reproduces it up to the lambda = 1e-8 regularisation
    rho_g = 1000.0 * 9.81
    p_scale = rho_g * surface_z
    assert float(np.max(np.abs(pressure - p_true))) <= 1e-7 * p_scale
    assert float(np.max(np.abs(R["ppe_pressure"] - p_true))) <= 1e-7 * p_scale
    inner = (st == 0) | (st == 1)
    assert float(np.max(np.abs(R["ppe_pressure_gradient_z"][inner] + rho_g))) <= 1e-5 * rho_g
    for k in ("ppe_pressure_gradient_x", "ppe_pressure_gradient_y"):
        assert float(np.max(np.abs(R[k][inner]))) <= 1e-5 * rho_g, k


def test_pipe_flow_program(tmp_path):
    out = _run("pipe_flow_diagnostics", tmp_path)
    vtk = parse_vtk(tmp_path / "pipe_flow_000000.vtk")
    pts = vtk["points"]
    h, radius, u_max = 0.025, 0.2, 1.0
    reach = int(math.ceil(radius / h))
    expect = [((i + 0.5) * h, j * h, k * h) for i in range(int(round(0.6 / h)))
              for j in range(-reach, reach + 1) for k in range(-reach, reach + 1)
              if (j * h) ** 2 + (k * h) ** 2 <= radius * radius]
    assert pts.shape[0] == len(expect)
    assert _stat(out, "fluid particles") == len(expect)
    segments = max(24, int(math.ceil(2 * math.pi * radius / h)))
    assert _stat(out, "wall particles") == int(round(0.6 / h)) * segments
    y, z = pts[:, 1], pts[:, 2]
    u = u_max * (1 - (y * y + z * z) / radius ** 2)
    assert np.max(np.abs(vtk["real"]["velocity_x"] - u)) <= 1e-7
    assert np.allclose(vtk["real"]["analytic_velocity_laplacian"], -4 * u_max / radius ** 2)
    grad = vtk["vectors"]["analytic_velocity_gradient"]
    assert np.max(np.abs(grad[:, 1] + 2 * u_max * y / radius ** 2)) <= 1e-7
    assert np.max(np.abs(grad[:, 2] + 2 * u_max * z / radius ** 2)) <= 1e-7
    assert abs(_stat(out, "analytic laplacian") - (-4 * u_max / radius ** 2)) <= 1e-4
