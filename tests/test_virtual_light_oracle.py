"""The oracle's virtual-light diagnostic (compute_virtual_light_diagnostics,
src/surface/surface_detection.cu:350-475,555-578) against cases whose open
directions follow from the geometry, and against a brute-force formulation.

The reference's unit tests hold no known answer for this diagnostic, so it is
pinned here by closed forms: an isolated particle sees all 14 directions; a
particle inside a lattice sees none; the centre of a lattice's top face sees
exactly +z, (+x,+z)/sqrt2 and (+y,+z)/sqrt2 with the 30-degree default cone
(every neighbor has dz <= 0, and an in-plane neighbor projects at most 1/sqrt2
on the two diagonals); a wall particle straight above blocks +z alone, and only
when wall neighbors are included.
"""
import math

import numpy as np
import pytest

from helpers import make_case

# virtual_light_direction (surface_detection.cu:350-383)
_H = math.sqrt(0.5)
DIRECTIONS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                       [_H, _H, 0], [_H, -_H, 0], [-_H, _H, 0], [-_H, -_H, 0],
                       [_H, 0, _H], [_H, 0, -_H], [0, _H, _H], [0, _H, -_H]])


def _config(cases, l0=0.01):
    return cases.case_config(l0)


def _case(cases, abi, points, walls=None, config=None, name="light"):
    p = np.asarray(points, np.float64).reshape(-1, 3)
    cfg = config or _config(cases)
    w = walls or {}
    case = make_case(cases, abi, name, (p[:, 0], p[:, 1], p[:, 2]), walls=w, config=cfg)
    cases.derive_cell_grid(case.config, case.fluid, case.walls)
    return case


def _half_space_block(l0, n=9, depth=5):
    """A lattice of (2n+1)^2 x depth points with its top face at z = 0."""
    r = np.arange(-n, n + 1) * l0
    zs = -np.arange(depth) * l0
    gx, gy, gz = np.meshgrid(r, r, zs, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    centre = int(np.nonzero((gx.ravel() == 0) & (gy.ravel() == 0) & (gz.ravel() == 0))[0][0])
    return pts, centre


def test_isolated_particle_sees_every_direction(oracle, cases, abi):
    case = _case(cases, abi, [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    assert out["open_direction_count"].tolist() == [14, 14]
    assert np.all(out["open_fraction"] == 1.0)


def test_lattice_interior_sees_nothing(oracle, cases, abi):
    l0 = 0.01
    r = np.arange(-5, 6) * l0
    gx, gy, gz = np.meshgrid(r, r, r, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    case = _case(cases, abi, pts)
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    centre = int(np.nonzero(np.all(pts == 0.0, axis=1))[0][0])
    assert out["open_direction_count"][centre] == 0
    assert out["open_fraction"][centre] == 0.0


def test_top_face_centre_sees_the_upper_directions(oracle, cases, abi):
    l0 = 0.01
    pts, centre = _half_space_block(l0)
    case = _case(cases, abi, pts)
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    assert out["open_direction_count"][centre] == 3  # directions 4, 10, 12
    assert out["open_fraction"][centre] == 3.0 / 14.0


@pytest.mark.parametrize("include_walls, expected", [(1, 2), (0, 3)])
def test_wall_above_blocks_plus_z_only_with_walls(oracle, cases, abi, include_walls, expected):
    l0 = 0.01
    pts, centre = _half_space_block(l0)
    walls = {"x": [0.0], "y": [0.0], "z": [l0], "normal_x": [0.0], "normal_y": [0.0],
             "normal_z": [-1.0]}
    cfg = _config(cases, l0)
    cfg.include_wall_neighbors = include_walls
    case = _case(cases, abi, pts, walls=walls, config=cfg)
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    assert out["open_direction_count"][centre] == expected


def test_neighbor_beyond_support_radius_does_not_block(oracle, cases, abi):
    """A row entry farther than re (a stale list) is skipped (d2 > re^2)."""
    cfg = _config(cases)
    far = 1.01 * cfg.support_radius
    case = _case(cases, abi, [[0.0, 0.0, 0.0], [far, 0.0, 0.0]], config=cfg)
    nb = {"fluid_offsets": np.array([0, 1, 2], np.int32),
          "fluid_indices": np.array([1, 0], np.int32),
          "wall_offsets": np.zeros(3, np.int32), "wall_indices": np.zeros(1, np.int32)}
    out = oracle.virtual_light(case, nb)
    assert out["open_direction_count"].tolist() == [14, 14]
    near = _case(cases, abi, [[0.0, 0.0, 0.0], [0.5 * far, 0.0, 0.0]], config=cfg)
    out = oracle.virtual_light(near, nb)
    assert out["open_direction_count"].tolist() == [13, 13]


def test_cone_cosine_minus_one_blocks_everything(oracle, cases, abi):
    cfg = _config(cases)
    cfg.virtual_light_cone_cosine = -1.0
    case = _case(cases, abi, [[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [1.0, 1.0, 1.0]], config=cfg)
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    assert out["open_direction_count"].tolist() == [0, 0, 14]


def test_random_cloud_matches_brute_force(oracle, cases, abi):
    """Every particle of a random cloud against an all-pairs evaluation that
    forms the cosine as (d . u) / |d| and covers both fluid and wall neighbors."""
    rng = np.random.default_rng(11)
    l0 = 0.01
    pts = rng.uniform(0.0, 8 * l0, size=(300, 3))
    wp = rng.uniform(0.0, 8 * l0, size=(60, 3))
    walls = {"x": wp[:, 0], "y": wp[:, 1], "z": wp[:, 2], "normal_x": np.zeros(60),
             "normal_y": np.zeros(60), "normal_z": np.ones(60)}
    cfg = _config(cases, l0)
    cfg.virtual_light_cone_cosine = 0.8
    case = _case(cases, abi, pts, walls=walls, config=cfg)
    nb = oracle.neighbors(case)
    out = oracle.virtual_light(case, nb)
    re = cfg.support_radius
    for i in range(pts.shape[0]):
        blocked = np.zeros(14, bool)
        for src, skip_self in ((pts, True), (wp, False)):
            d = src - pts[i]
            r = np.linalg.norm(d, axis=1)
            ok = (r > 0) & (r <= re)
            if skip_self:
                ok[i] = False
            if not ok.any():
                continue
            cos = (d[ok] @ DIRECTIONS.T) / r[ok][:, None]
            blocked |= np.any(cos >= 0.8, axis=0)
        assert out["open_direction_count"][i] == 14 - int(blocked.sum()), i


def test_input_errors(oracle, cases, abi):
    cfg = _config(cases)
    case = _case(cases, abi, [[0.0, 0.0, 0.0]], config=cfg)
    nb = oracle.neighbors(case)
    cfg.virtual_light_cone_cosine = 1.5
    with pytest.raises(RuntimeError, match=r"cone cosine must be in \[-1, 1\]"):
        oracle.virtual_light(case, nb, config=cfg)
    cfg.virtual_light_cone_cosine = float("nan")  # the reference's test lets NaN through
    out = oracle.virtual_light(case, nb, config=cfg)
    assert out["open_direction_count"].tolist() == [14]
    cfg.virtual_light_cone_cosine = 0.5
    cfg.support_radius = 0.0
    with pytest.raises(RuntimeError, match="support radius must be positive"):
        oracle.virtual_light(case, nb, config=cfg)
