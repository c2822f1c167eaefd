"""The oracle's moment matrices reproduce quadratic fields exactly, for all
four kinds, on a jittered corner stencil with wall samples and Neumann rows
(moment_matrix.cu:152-309, kinds :494-538).

For phi(x) = a + g.x + x^T H x / 2, the LSMPS least-squares fit at particle i
has the exact Taylor coefficients as its solution whatever the weights, so
c = M^-1 rhs must give, in the scaled basis s = d / re,
  Type A  [sx sy sz sx^2 sy^2 sz^2 sx.sy sy.sz sz.sx]:
          re grad phi(x_i), re^2 H_kk / 2, re^2 H_xy, re^2 H_yz, re^2 H_zx
  Type B  [1, the same nine]: phi(x_i) first.
The right-hand side is written here from the sampling each kind describes:
  fluid samples  w p(d) (phi_j - phi_i)       (Type B: w p(d) phi_j)
  wall samples   w_s w p(d) (phi_w - phi_i)   velocity Type A (walls as samples)
  Neumann rows   w_s w q(d, n) re n.grad phi(x_w)  pressure Type A / B
  none           fluid-only Type A.
q = re dp/dn is the basis' derivative along the wall normal.  A wrong basis
order, a wrong sign or factor in q, a missing wall term or a transposed
inverse all break the reproduction; the only error left is the lambda = 1e-8
regularisation (relative 1e-8 times the conditioning of M).
"""
import numpy as np
import pytest

KINDS = ("velocity_a", "pressure_a", "pressure_b", "fluid_only_a")


def _basis(d, re, type_b):
    s = d / re
    sx, sy, sz = s
    b = [sx, sy, sz, sx * sx, sy * sy, sz * sz, sx * sy, sy * sz, sz * sx]
    return np.array(([1.0] if type_b else []) + b)


def _q(d, n, re, type_b):
    dx, dy, dz = d
    nx, ny, nz = n
    q = [nx, ny, nz, 2 * dx * nx / re, 2 * dy * ny / re, 2 * dz * nz / re,
         (dy * nx + dx * ny) / re, (dz * ny + dy * nz) / re, (dx * nz + dz * nx) / re]
    return np.array(([0.0] if type_b else []) + q)


def _weight(r, re, quadratic):
    if r >= re:
        return 0.0
    w = 1.0 - r / re
    return w * w if quadratic else w


@pytest.mark.parametrize("quadratic", [False, True])
def test_quadratic_reproduction_all_kinds(oracle, abi, cases, quadratic):
    kernel = abi.WEIGHT_QUADRATIC if quadratic else abi.WEIGHT_LINEAR
    case = cases.dam_break(block=(10, 9, 10), tank=(12, 9, 13), l0=0.01, seed=11, amplitude=0.2,
                           weight_kernel=kernel)
    cfg = case.config
    re = cfg.support_radius
    ws = cfg.lsmps_wall_weight_scale
    nb = oracle.neighbors(case)
    f, w = case.fluid, case.walls
    fx = np.stack([f["x"], f["y"], f["z"]], 1)
    wx = np.stack([w["x"], w["y"], w["z"]], 1)
    wn = np.stack([w["normal_x"], w["normal_y"], w["normal_z"]], 1)
    rng = np.random.default_rng(3)
    g = rng.normal(size=3)
    a0 = 0.7
    h = rng.normal(size=(3, 3))
    h = h + h.T

    def phi(x):
        return a0 + x @ g + 0.5 * np.einsum("...i,ij,...j->...", x, h, x)

    def grad(x):
        return g + x @ h

    kinds = {"velocity_a": abi.MOMENT_VELOCITY_WALL_DIRICHLET_A,
             "pressure_a": abi.MOMENT_PRESSURE_WALL_NEUMANN_A,
             "pressure_b": abi.MOMENT_PRESSURE_WALL_NEUMANN_B,
             "fluid_only_a": abi.MOMENT_FLUID_ONLY_A}
    # particles with walls of both kinds (floor and side) in range, plus interior ones
    fo, fi = nb["fluid_offsets"], nb["fluid_indices"]
    wo, wi = nb["wall_offsets"], nb["wall_indices"]
    with_walls = [i for i in range(case.n_fluid) if wo[i + 1] - wo[i] >= 20]
    interior = [i for i in range(case.n_fluid) if wo[i + 1] == wo[i] and fo[i + 1] - fo[i] >= 60]
    assert len(with_walls) >= 10 and len(interior) >= 1
    picks = with_walls[::max(1, len(with_walls) // 12)] + interior[:3]
    worst = {}
    for name, kind in kinds.items():
        mo = oracle.moments(case, nb, kind)
        size = mo["size"]
        type_b = size == 10
        worst[name] = 0.0
        for i in picks:
            xi = fx[i]
            rhs = np.zeros(size)
            for c in range(fo[i], fo[i + 1]):
                j = fi[c]
                d = fx[j] - xi
                wt = _weight(np.linalg.norm(d), re, quadratic)
                data = phi(fx[j]) if type_b else phi(fx[j]) - phi(xi)
                rhs += wt * _basis(d, re, type_b) * data
            if name != "fluid_only_a":
                for c in range(wo[i], wo[i + 1]):
                    j = wi[c]
                    d = wx[j] - xi
                    wt = ws * _weight(np.linalg.norm(d), re, quadratic)
                    if name == "velocity_a":
                        rhs += wt * _basis(d, re, False) * (phi(wx[j]) - phi(xi))
                    else:
                        rhs += wt * _q(d, wn[j], re, type_b) * (re * wn[j] @ grad(wx[j]))
            minv = mo["inverse"][i].T  # stored column-major: [col][row]
            coef = minv @ rhs
            gi = grad(xi)
            exact = [re * gi[0], re * gi[1], re * gi[2], re * re * h[0, 0] / 2,
                     re * re * h[1, 1] / 2, re * re * h[2, 2] / 2, re * re * h[0, 1],
                     re * re * h[1, 2], re * re * h[2, 0]]
            if type_b:
                exact = [phi(xi)] + exact
            exact = np.array(exact)
            err = np.linalg.norm(coef - exact) / np.linalg.norm(exact)
            worst[name] = max(worst[name], err)
    print("worst relative coefficient error per kind:", worst)
    for name in KINDS:
        assert worst[name] <= 1e-6, (name, worst[name])


def test_reproduction_detects_dropped_wall_term(oracle, abi, cases):
    """Guard for the test above: the pressure Type-A inverse applied to a
    right-hand side without its Neumann rows no longer reproduces the field at
    wall-adjacent particles (so those rows are really exercised)."""
    case = cases.dam_break(block=(10, 9, 10), tank=(12, 9, 13), l0=0.01, seed=11, amplitude=0.2)
    re = case.config.support_radius
    nb = oracle.neighbors(case)
    f = case.fluid
    fx = np.stack([f["x"], f["y"], f["z"]], 1)
    fo, fi, wo = nb["fluid_offsets"], nb["fluid_indices"], nb["wall_offsets"]
    i = max(range(case.n_fluid), key=lambda k: wo[k + 1] - wo[k])
    mo = oracle.moments(case, nb, abi.MOMENT_PRESSURE_WALL_NEUMANN_A)
    g = np.array([0.3, -0.2, 1.1])
    rhs = np.zeros(9)
    for c in range(fo[i], fo[i + 1]):
        d = fx[fi[c]] - fx[i]
        rhs += _weight(np.linalg.norm(d), re, False) * _basis(d, re, False) * (d @ g)
    coef = mo["inverse"][i].T @ rhs
    assert np.linalg.norm(coef[:3] - re * g) / np.linalg.norm(re * g) > 1e-3
