"""The FP64 CPU reference gradient of one step (oracle/adjoint.py) against
central finite differences of the same frozen step.

The step is differentiated with its neighbor lists and labels held fixed (they
are piecewise constant in the positions).  Directions are Rademacher (+-1 per
component) so that the directional derivative is large against rounding.  The
stress case reaches every branch: shifting (with its displacement cap),
anti-penetration, Dirichlet and splash rows; its jitter puts some pairs within
1e-6 l0 of a branch point, so it is differenced with eps = 1e-7 l0.
"""
import numpy as np
import pytest

from oracle.adjoint import FrozenStep

SPLASH, SURFACE = 3, 2


def _loss_factory(case, rng, n):
    vt = case.extra["v_target"].T
    cx = rng.standard_normal((3, n))

    def loss(f):
        return ((f["v_new"] - vt) ** 2).sum() + (cx * f["x_new"]).sum()

    def upstream(f):
        return 2.0 * (f["v_new"] - vt), cx

    return loss, upstream


def _check(fd, an, tol, floor=0.0):
    err = abs(fd - an) / max(abs(fd), floor, 1e-300)
    assert err <= tol, (fd, an, err)
    return err


def _fd_suite(orc, case, eps_x, tol, seed=0):
    fs = FrozenStep(orc, case)
    x0, v0, g0 = fs.state()
    n = x0.shape[1]
    rng = np.random.default_rng(seed)
    loss, upstream = _loss_factory(case, rng, n)
    fw = fs.forward(x0, v0, g0)
    vb, xb = upstream(fw)
    adj = fs.adjoint(fw, vb, xbar_out=xb)
    # the replayed update + anti-penetration reproduces the oracle's positions
    assert np.array_equal(adj["x_corrected"], fw["x_new"])
    gmax = np.abs(adj["gravity"]).max()
    for k in range(3):
        e = 1e-6 * np.abs(g0).max()
        gp, gm = g0.copy(), g0.copy()
        gp[k] += e
        gm[k] -= e
        fd = (loss(fs.forward(x0, v0, gp)) - loss(fs.forward(x0, v0, gm))) / (2 * e)
        _check(fd, adj["gravity"][k], tol, floor=1e-3 * gmax)
    for _ in range(2):
        d = rng.choice([-1.0, 1.0], size=x0.shape)
        fd = (loss(fs.forward(x0 + eps_x * d, v0, g0)) -
              loss(fs.forward(x0 - eps_x * d, v0, g0))) / (2 * eps_x)
        _check(fd, (adj["position"] * d).sum(), tol)
    for _ in range(2):
        d = rng.choice([-1.0, 1.0], size=x0.shape)
        e = 1e-7
        fd = (loss(fs.forward(x0, v0 + e * d, g0)) - loss(fs.forward(x0, v0 - e * d, g0))) / (2 * e)
        _check(fd, (adj["velocity"] * d).sum(), tol)
    p = fw["p"]
    d = np.where(p > 1e-3 * p.max(), rng.choice([-1.0, 1.0], size=p.shape), 0.0)
    e = 1e-6 * p.max()
    fd = (loss(fs.forward(x0, v0, g0, pressure=p + e * d)) -
          loss(fs.forward(x0, v0, g0, pressure=p - e * d))) / (2 * e)
    _check(fd, (adj["pressure"] * d).sum(), tol)
    return fs, fw, adj


@pytest.mark.parametrize("kernel", [0, 1], ids=["linear", "quadratic"])
def test_reference_gradient_matches_fd_on_stress_case(oracle, pkg, kernel):
    case = pkg.cases.adjoint_stress(n=12, weight_kernel=kernel)
    fs, fw, adj = _fd_suite(oracle, case, eps_x=1e-7 * case.extra["l0"], tol=1e-5)
    cfg = case.config
    st = fs.st
    # branch coverage of this case
    assert (st == SPLASH).sum() > 0 and (st == SURFACE).sum() > 0
    ps = np.stack(fw["cor"]["ps"])
    cap = cfg.ps_max_displacement_ratio * cfg.particle_spacing
    assert (np.sqrt((ps * ps).sum(0)) >= cap * (1 - 1e-12)).sum() > 0
    tv = np.stack([fw["tmp"][k] for k in ("vx", "vy", "vz")])
    grad = np.stack(fw["cor"]["grad"])
    nv = tv.copy()
    ns = st != SPLASH
    nv[:, ns] -= cfg.time_step / cfg.density * grad[:, ns]
    xp = fw["x"] + (0.5 * cfg.time_step * (fw["v"] + nv) + ps)
    assert (np.abs(adj["x_corrected"] - xp).sum(0) > 0).sum() > 0  # anti-penetration moved


def test_reference_gradient_matches_fd_on_sloshing(oracle, pkg):
    """Case C4's construction at 10^3 particles, eps = 1e-6 l0."""
    case = pkg.cases.sloshing(n=10, l0=0.005)
    _fd_suite(oracle, case, eps_x=1e-6 * case.extra["l0"], tol=1e-5, seed=1)
