"""Closed-form checks of the oracle's particle shifting, anti-penetration and
velocity smoothing (pressure_correction.cu:278-367, 404-444, 446-515).

The reference's unit tests hold no known answer for these three kernels, so the
oracle is pinned by what the formulas fix on tiny hand-built stencils:
- shifting: a pair closer than ps_min_distance_ratio * l0 pushes each particle
  straight away from the other by scale * (min_distance - d); wall neighbours
  push along the wall normal by scale * (min_distance - d) (no 1/d); symmetric
  stencils cancel; the shift is capped at ps_max_displacement_ratio * l0 in its
  own direction; a Surface particle keeps only the part tangent to its fluid
  normal; Splash particles and a zero scale give no shift;
- anti-penetration: a particle 0.1 l0 off a wall ends exactly at the clearance
  0.25 l0 with its inward normal velocity removed and its outward velocity
  kept; two walls in a corner are resolved one after the other;
- smoothing: uniform and linear velocity fields on symmetric stencils are fixed
  points; a single neighbour gives (1 - alpha) v_i + alpha v_j; Splash and
  out-of-range neighbours are ignored.
Each expected value is written from the geometric statement, not from the
oracle's code, so a dropped term, a wrong sign or a wrong normalisation fails.
"""
import numpy as np
import pytest

from helpers import make_case

L0 = 0.01


def _cfg(abi, **kw):
    c = abi.default_config()
    c.particle_spacing = L0
    c.support_radius = 3.1 * L0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _csr(rows):
    offsets = np.zeros(len(rows) + 1, np.int32)
    for i, r in enumerate(rows):
        offsets[i + 1] = offsets[i] + len(r)
    idx = np.array([j for r in rows for j in r], np.int32)
    return offsets, idx


def _nb(fluid_rows, wall_rows=None):
    fo, fi = _csr(fluid_rows)
    wo, wi = _csr(wall_rows or [[] for _ in fluid_rows])
    return {"fluid_offsets": fo, "fluid_indices": fi, "wall_offsets": wo, "wall_indices": wi}


def _walls(points, normals):
    p = np.asarray(points, np.float64).reshape(-1, 3)
    nrm = np.asarray(normals, np.float64).reshape(-1, 3)
    k = p.shape[0]
    return {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2], "vx": np.zeros(k), "vy": np.zeros(k),
            "vz": np.zeros(k), "normal_x": nrm[:, 0], "normal_y": nrm[:, 1], "normal_z": nrm[:, 2]}


def _case(cases, abi, pts, cfg, walls=None, velocity=None, labels=None):
    p = np.asarray(pts, np.float64).reshape(-1, 3)
    return make_case(cases, abi, "closed_form", (p[:, 0], p[:, 1], p[:, 2]), walls=walls,
                     config=cfg, velocity=velocity, surface_type=labels)


# ---- particle shifting ----------------------------------------------------

def test_shifting_pair_push(oracle, abi, cases):
    cfg = _cfg(abi)
    d = 0.5 * L0
    case = _case(cases, abi, [[0, 0, 0], [d, 0, 0]], cfg)
    ps = oracle.particle_shifting(case, _nb([[1], [0]]))
    mag = cfg.ps_displacement_scale * (cfg.ps_min_distance_ratio * L0 - d)
    # each particle moves straight away from the other
    np.testing.assert_allclose([ps[0][0], ps[1][0], ps[2][0]], [-mag, 0, 0], atol=1e-18, rtol=1e-13)
    np.testing.assert_allclose([ps[0][1], ps[1][1], ps[2][1]], [mag, 0, 0], atol=1e-18, rtol=1e-13)


def test_shifting_oblique_pair_direction(oracle, abi, cases):
    cfg = _cfg(abi)
    u = np.array([1.0, -2.0, 2.0]) / 3.0
    d = 0.4 * L0
    case = _case(cases, abi, [[0, 0, 0], list(d * u)], cfg)
    ps = oracle.particle_shifting(case, _nb([[1], [0]]))
    mag = cfg.ps_displacement_scale * (cfg.ps_min_distance_ratio * L0 - d)
    np.testing.assert_allclose([ps[k][0] for k in range(3)], -mag * u, rtol=1e-12, atol=1e-18)


def test_shifting_beyond_min_distance_and_symmetric_cancel(oracle, abi, cases):
    cfg = _cfg(abi)
    md = cfg.ps_min_distance_ratio * L0
    # neighbour exactly at the minimum distance and one beyond: no push
    case = _case(cases, abi, [[0, 0, 0], [md, 0, 0], [0, 1.2 * md, 0]], cfg)
    ps = oracle.particle_shifting(case, _nb([[1, 2], [0], [0]]))
    assert ps[0][0] == 0.0 and ps[1][0] == 0.0 and ps[2][0] == 0.0
    # six neighbours at +-d on the axes: the pushes cancel exactly
    d = 0.5 * L0
    pts = [[0, 0, 0], [d, 0, 0], [-d, 0, 0], [0, d, 0], [0, -d, 0], [0, 0, d], [0, 0, -d]]
    case = _case(cases, abi, pts, cfg)
    rows = [[1, 2, 3, 4, 5, 6]] + [[0]] * 6
    ps = oracle.particle_shifting(case, _nb(rows))
    assert ps[0][0] == 0.0 and ps[1][0] == 0.0 and ps[2][0] == 0.0


def test_shifting_cap(oracle, abi, cases):
    cfg = _cfg(abi, ps_displacement_scale=10.0)
    u = np.array([2.0, 3.0, 6.0]) / 7.0
    d = 0.3 * L0
    case = _case(cases, abi, [[0, 0, 0], list(d * u)], cfg)
    ps = oracle.particle_shifting(case, _nb([[1], [0]]))
    cap = cfg.ps_max_displacement_ratio * L0
    got = np.array([ps[k][0] for k in range(3)])
    # the uncapped push (10 * 0.55 l0) exceeds the cap: length = cap, same direction
    np.testing.assert_allclose(np.linalg.norm(got), cap, rtol=1e-14)
    np.testing.assert_allclose(got, -cap * u, rtol=1e-13)


def test_shifting_wall_push_along_normal(oracle, abi, cases):
    cfg = _cfg(abi)
    dw = 0.6 * L0
    md = cfg.ps_min_distance_ratio * L0
    # wall particle below and slightly to the side; its normal points up (+z)
    wall = _walls([[0.3 * L0, 0, -np.sqrt(dw ** 2 - (0.3 * L0) ** 2)]], [[0, 0, 1]])
    case = _case(cases, abi, [[0, 0, 0]], cfg, walls=wall)
    ps = oracle.particle_shifting(case, _nb([[]], [[0]]))
    # the push is scale * (min_distance - d) along the wall normal, with no 1/d
    np.testing.assert_allclose([ps[0][0], ps[1][0], ps[2][0]],
                               [0.0, 0.0, cfg.ps_displacement_scale * (md - dw)],
                               rtol=1e-13, atol=1e-18)


def test_shifting_surface_keeps_tangential_part(oracle, abi, cases):
    cfg = _cfg(abi)
    md = cfg.ps_min_distance_ratio * L0
    d = 0.5 * L0
    dw = 0.7 * L0
    # fluid neighbour straight below: push +z, fluid normal +z; wall to the -x
    # side with normal +x: push +x
    wall = _walls([[-dw, 0, 0]], [[1, 0, 0]])
    labels = [abi.SURFACE, abi.INNER]
    case = _case(cases, abi, [[0, 0, 0], [0, 0, -d]], cfg, walls=wall, labels=labels)
    nb = _nb([[1], [0]], [[0], []])
    ps = oracle.particle_shifting(case, nb)
    # the Surface particle keeps only the tangential (x) part
    np.testing.assert_allclose([ps[0][0], ps[1][0], ps[2][0]],
                               [cfg.ps_displacement_scale * (md - dw), 0.0, 0.0],
                               rtol=1e-13, atol=1e-18)
    # the same particle labelled Inner keeps both parts
    case.fluid["surface_type"][0] = abi.INNER
    ps = oracle.particle_shifting(case, nb)
    np.testing.assert_allclose([ps[0][0], ps[1][0], ps[2][0]],
                               [cfg.ps_displacement_scale * (md - dw), 0.0,
                                cfg.ps_displacement_scale * (md - d)], rtol=1e-13, atol=1e-18)


def test_shifting_disabled_for_splash_and_zero_scale(oracle, abi, cases):
    d = 0.5 * L0
    cfg = _cfg(abi)
    case = _case(cases, abi, [[0, 0, 0], [d, 0, 0]], cfg, labels=[abi.SPLASH, abi.INNER])
    ps = oracle.particle_shifting(case, _nb([[1], [0]]))
    assert ps[0][0] == 0.0 and ps[0][1] != 0.0
    cfg0 = _cfg(abi, ps_displacement_scale=0.0)
    case = _case(cases, abi, [[0, 0, 0], [d, 0, 0]], cfg0)
    ps = oracle.particle_shifting(case, _nb([[1], [0]]))
    assert not np.any(np.concatenate(ps))


# ---- anti-penetration -----------------------------------------------------

def test_anti_penetration_floor(oracle, abi, cases):
    cfg = _cfg(abi)
    clearance = cfg.wall_clearance_ratio * L0
    wall = _walls([[0.2 * L0, -0.1 * L0, 0.0]], [[0, 0, 1]])
    # 0.1 l0 above the wall plane, moving down and sideways
    case = _case(cases, abi, [[0.0, 0.0, 0.1 * L0]], cfg, walls=wall,
                 velocity=([0.3], [-0.2], [-1.5]))
    out = oracle.anti_penetration(case, _nb([[]], [[0]]))
    np.testing.assert_allclose(out["z"][0], clearance, rtol=1e-14)
    assert out["x"][0] == 0.0 and out["y"][0] == 0.0
    assert out["vz"][0] == 0.0 and out["vx"][0] == 0.3 and out["vy"][0] == -0.2


def test_anti_penetration_outward_velocity_kept_and_clear_particle_untouched(oracle, abi, cases):
    cfg = _cfg(abi)
    clearance = cfg.wall_clearance_ratio * L0
    wall = _walls([[0.0, 0.0, 0.0]], [[0, 0, 1]])
    case = _case(cases, abi, [[0.0, 0.0, 0.1 * L0], [0.0, 0.0, 0.3 * L0]], cfg, walls=wall,
                 velocity=([0.0, 0.0], [0.0, 0.0], [0.7, -2.0]))
    out = oracle.anti_penetration(case, _nb([[], []], [[0], [0]]))
    np.testing.assert_allclose(out["z"][0], clearance, rtol=1e-14)
    assert out["vz"][0] == 0.7  # moving away from the wall: kept
    assert out["z"][1] == 0.3 * L0 and out["vz"][1] == -2.0  # beyond the clearance


def test_anti_penetration_corner_sequential(oracle, abi, cases):
    cfg = _cfg(abi)
    clearance = cfg.wall_clearance_ratio * L0
    # a floor (normal +z) and a side wall at x = 0 (normal +x)
    wall = _walls([[0.5 * L0, 0.0, 0.0], [0.0, 0.0, 0.5 * L0]], [[0, 0, 1], [1, 0, 0]])
    case = _case(cases, abi, [[0.05 * L0, 0.0, 0.05 * L0]], cfg, walls=wall,
                 velocity=([-1.0], [0.25], [-1.0]))
    out = oracle.anti_penetration(case, _nb([[]], [[0, 1]]))
    np.testing.assert_allclose([out["x"][0], out["z"][0]], [clearance, clearance], rtol=1e-14)
    assert out["vx"][0] == 0.0 and out["vz"][0] == 0.0 and out["vy"][0] == 0.25


def test_anti_penetration_disabled(oracle, abi, cases):
    cfg = _cfg(abi, wall_clearance_ratio=0.0)
    wall = _walls([[0.0, 0.0, 0.0]], [[0, 0, 1]])
    case = _case(cases, abi, [[0.0, 0.0, -0.1 * L0]], cfg, walls=wall,
                 velocity=([0.0], [0.0], [-1.0]))
    out = oracle.anti_penetration(case, _nb([[]], [[0]]))
    assert out["z"][0] == -0.1 * L0 and out["vz"][0] == -1.0


# ---- velocity smoothing ---------------------------------------------------

def _cube_stencil(h):
    pts = [[0.0, 0.0, 0.0]]
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx or dy or dz:
                    pts.append([dx * h, dy * h, dz * h])
    return np.array(pts)


@pytest.mark.parametrize("kernel", ["linear", "quadratic"])
def test_smoothing_fixed_points(oracle, abi, cases, kernel):
    cfg = _cfg(abi, weight_kernel=abi.WEIGHT_LINEAR if kernel == "linear" else abi.WEIGHT_QUADRATIC)
    pts = _cube_stencil(L0)
    n = pts.shape[0]
    rows = [list(range(1, n))] + [[0]] * (n - 1)
    # uniform velocity
    v = ([0.25] * n, [-0.5] * n, [1.0] * n)
    case = _case(cases, abi, pts, cfg, velocity=v)
    vx, vy, vz = oracle.smooth_velocity(case, _nb(rows))
    np.testing.assert_allclose([vx[0], vy[0], vz[0]], [0.25, -0.5, 1.0], rtol=1e-15)
    # linear field v = A x on the symmetric stencil: the weighted mean is v(0) = 0
    a = np.array([[1.0, 2.0, -1.0], [0.5, -3.0, 2.0], [4.0, 1.0, 0.25]])
    lin = pts @ a.T
    case = _case(cases, abi, pts, cfg, velocity=(lin[:, 0], lin[:, 1], lin[:, 2]))
    vx, vy, vz = oracle.smooth_velocity(case, _nb(rows))
    assert max(abs(vx[0]), abs(vy[0]), abs(vz[0])) <= 1e-15


def test_smoothing_single_neighbor_and_exclusions(oracle, abi, cases):
    cfg = _cfg(abi)
    alpha = cfg.velocity_smoothing_strength
    re = cfg.support_radius
    pts = [[0, 0, 0], [0.7 * L0, 0, 0], [0, 0.5 * L0, 0], [0, 0, re]]
    labels = [abi.INNER, abi.INNER, abi.SPLASH, abi.INNER]
    v = ([1.0, 3.0, 100.0, -50.0], [0.0, -1.0, 100.0, -50.0], [2.0, 2.0, 100.0, -50.0])
    case = _case(cases, abi, pts, cfg, velocity=v, labels=labels)
    # particle 0 sees an Inner neighbour, a Splash neighbour (ignored) and one
    # at exactly re (weight 0, ignored)
    rows = [[1, 2, 3], [0], [0], [0]]
    vx, vy, vz = oracle.smooth_velocity(case, _nb(rows))
    np.testing.assert_allclose([vx[0], vy[0], vz[0]],
                               [(1 - alpha) * 1.0 + alpha * 3.0, (1 - alpha) * 0.0 + alpha * -1.0,
                                2.0], rtol=1e-14)
    # the Splash particle keeps its velocity
    assert vx[2] == 100.0 and vy[2] == 100.0 and vz[2] == 100.0
    # no weighted neighbour: unchanged (particle 3's only neighbour is at re)
    assert vx[3] == -50.0


def test_smoothing_strength_clamped(oracle, abi, cases):
    cfg = _cfg(abi, velocity_smoothing_strength=7.0)  # clamped to 1: the neighbour mean
    pts = [[0, 0, 0], [0.5 * L0, 0, 0]]
    case = _case(cases, abi, pts, cfg, velocity=([1.0, 3.0], [0.0, 0.0], [0.0, 0.0]))
    vx, _, _ = oracle.smooth_velocity(case, _nb([[1], [0]]))
    assert vx[0] == 3.0 and vx[1] == 1.0
