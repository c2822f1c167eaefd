"""Adjoint of one step on the GPU (lsmps_adjoint_step through the C-ABI).

1. Against the FP64 CPU reference gradient (oracle/adjoint.py, itself pinned by
   central finite differences in test_adjoint_oracle.py) on the stress case that
   reaches every branch: 1e-7 relative L2 in FP64 mode, 1e-5 in mixed mode.
2. Case C4 (50^3 = 125,000-particle sloshing box), loss sum |v_new - v_target|^2,
   through the torch.autograd.Function of the test harness: gradients with
   respect to positions, velocities and body force against central finite
   differences of the GPU step with frozen geometry (eps = 1e-6 l0 for
   positions), within 1e-4 relative.  In mixed mode the operators are FP32
   factorisations plus one FP64 refinement step, so the mixed position and
   velocity gradients are compared with the FP64 ones (1e-4) and the mixed
   body-force gradient is differenced.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _stepper(rt, abi, case, precision, tolerance=None):
    ctx = rt.Context(precision)
    state = rt.DeviceState(case)
    opts = abi.default_solver_options()
    if tolerance is not None:
        opts.tolerance = tolerance
        opts.max_iterations = 40000
    st = rt.Stepper(ctx, state, case.config, 256, opts)
    return ctx, state, st


@pytest.mark.parametrize("precision,tol", [(0, 1e-7), (1, 1e-5)], ids=["fp64", "mixed"])
def test_adjoint_matches_reference_gradient(oracle, pkg, rt, abi, precision, tol):
    from oracle.adjoint import FrozenStep
    case = pkg.cases.adjoint_stress(n=12)
    print(f"\nadjoint stress case n={case.n_fluid} precision={precision}", flush=True)
    fs = FrozenStep(oracle, case)
    x0, v0, g0 = fs.state()
    fw = fs.forward(x0, v0, g0)
    vt = case.extra["v_target"].T
    cx = np.random.default_rng(0).standard_normal(x0.shape)
    ctx, state, st = _stepper(rt, abi, case, precision)
    st.set_mode(abi.STEP_RECORD)
    rep = st.step()
    assert rep["converged"]
    labels = state.fluid["surface_type"].cpu().numpy()
    assert np.array_equal(labels, fs.st)
    f = state.fluid
    v_new = torch.stack([f["vx"], f["vy"], f["vz"]]).cpu().numpy()
    x_new = torch.stack([f["x"], f["y"], f["z"]]).cpu().numpy()
    assert rel_l2(x_new, fw["x_new"]) < 1e-12
    vb = 2.0 * (v_new - vt)
    ref = fs.adjoint(fw, 2.0 * (fw["v_new"] - vt), xbar_out=cx)
    out = st.adjoint(torch.from_numpy(vb).cuda(), torch.from_numpy(cx).cuda())
    assert out["report"]["converged"]
    errs = {k: rel_l2(out[k].cpu().numpy(), ref[k]) for k in ("position", "velocity", "pressure")}
    errs["gravity"] = rel_l2(out["gravity"], ref["gravity"])
    print("relative L2 vs reference gradient:", errs, "A^T iterations",
          out["report"]["iterations"], flush=True)
    for k, e in errs.items():
        assert e <= tol, (k, e)
    st.close()
    ctx.close()


def _frozen_loss(st, state, cfg, labels, vt, x, v, g):
    f = state.fluid
    with torch.no_grad():
        for k, key in enumerate(("x", "y", "z")):
            f[key].copy_(x[k])
        for k, key in enumerate(("vx", "vy", "vz")):
            f[key].copy_(v[k])
        f["surface_type"].copy_(labels)
    c = cfg.copy()
    for k in range(3):
        c.gravity[k] = float(g[k])
    st.set_config(c)
    rep = st.step(raise_on_nonconvergence=False)
    v_new = torch.stack([f["vx"], f["vy"], f["vz"]])
    return float(((v_new - vt) ** 2).sum()), rep


def _fd_gravity(st, state, cfg, labels, vt, xd, vd, gd):
    fd = np.zeros(3)
    for k in range(3):
        e = 1e-6 * abs(gd).max()
        gp, gm = gd.copy(), gd.copy()
        gp[k] += e
        gm[k] -= e
        fd[k] = (_frozen_loss(st, state, cfg, labels, vt, xd, vd, gp)[0] -
                 _frozen_loss(st, state, cfg, labels, vt, xd, vd, gm)[0]) / (2 * e)
    return fd


def _c4_gradients(rt, abi, case, precision):
    from adjoint_harness import StepFunction
    ctx, state, st = _stepper(rt, abi, case, precision, tolerance=1e-12)
    dev = torch.device("cuda:0")
    fl = case.fluid
    x = torch.tensor(np.stack([fl["x"], fl["y"], fl["z"]]), device=dev, requires_grad=True)
    v = torch.tensor(np.stack([fl["vx"], fl["vy"], fl["vz"]]), device=dev, requires_grad=True)
    g = torch.tensor(list(case.config.gravity), dtype=torch.float64, requires_grad=True)
    vt = torch.tensor(case.extra["v_target"].T.copy(), device=dev)
    x_new, v_new = StepFunction.apply(x, v, g, st, state, case.config, abi)
    loss = ((v_new - vt) ** 2).sum()
    loss.backward()
    return ctx, state, st, x, v, g, vt, float(loss.detach())


@pytest.fixture(scope="module")
def c4_fp64(pkg, rt, abi):
    case = pkg.cases.sloshing(n=50, l0=0.005)
    print(f"\nC4 sloshing n={case.n_fluid} FP64", flush=True)
    ctx, state, st, x, v, g, vt, loss = _c4_gradients(rt, abi, case, abi.PRECISION_FP64)
    yield case, ctx, state, st, x, v, g, vt, loss
    st.close()
    ctx.close()


def test_adjoint_c4_fp64_matches_finite_differences(c4_fp64, abi):
    case, ctx, state, st, x, v, g, vt, loss = c4_fp64
    labels = state.fluid["surface_type"].clone()
    st.set_mode(abi.STEP_FROZEN_GEOMETRY)
    l0 = case.extra["l0"]
    cfg = case.config
    xd, vd, gd = x.detach(), v.detach(), g.detach().numpy()
    base, rep = _frozen_loss(st, state, cfg, labels, vt, xd, vd, gd)
    print("loss", loss, "frozen replay", base, "solve", rep, flush=True)
    assert abs(base - loss) <= 1e-12 * abs(loss)
    gen = torch.Generator(device="cpu").manual_seed(7)
    errs = {}
    fd_g = _fd_gravity(st, state, cfg, labels, vt, xd, vd, gd)
    an_g = g.grad.numpy()
    # the body-force gradient is a 3-vector: its error is measured against its norm
    errs["g"] = (fd_g, an_g, float(np.linalg.norm(fd_g - an_g) / np.linalg.norm(fd_g)))
    for name, eps, base_t, grad in (("x", 1e-6 * l0, xd, x.grad), ("v", 1e-7, vd, v.grad)):
        d = (torch.randint(0, 2, tuple(base_t.shape), generator=gen) * 2 - 1).to(torch.float64)
        d = d.to(base_t.device)
        if name == "x":
            lp = _frozen_loss(st, state, cfg, labels, vt, base_t + eps * d, vd, gd)[0]
            lm = _frozen_loss(st, state, cfg, labels, vt, base_t - eps * d, vd, gd)[0]
        else:
            lp = _frozen_loss(st, state, cfg, labels, vt, xd, base_t + eps * d, gd)[0]
            lm = _frozen_loss(st, state, cfg, labels, vt, xd, base_t - eps * d, gd)[0]
        fd = (lp - lm) / (2 * eps)
        an = float((grad * d).sum())
        errs[name] = (fd, an, abs(fd - an) / abs(fd))
    print("finite differences (fd, adjoint, rel err):", errs, flush=True)
    for k, (fd, an, e) in errs.items():
        assert e <= 1e-4, (k, fd, an, e)


def test_adjoint_c4_mixed(c4_fp64, pkg, rt, abi):
    case, _, _, _, x64, v64, g64, _, _ = c4_fp64
    print(f"\nC4 sloshing n={case.n_fluid} mixed", flush=True)
    ctx, state, st, x, v, g, vt, loss = _c4_gradients(rt, abi, case, abi.PRECISION_MIXED)
    ex = rel_l2(x.grad.cpu().numpy(), x64.grad.cpu().numpy())
    ev = rel_l2(v.grad.cpu().numpy(), v64.grad.cpu().numpy())
    eg = rel_l2(g.grad.numpy(), g64.grad.numpy())
    print("mixed vs FP64 adjoint: position", ex, "velocity", ev, "gravity", eg, flush=True)
    assert ex <= 1e-4 and ev <= 1e-4 and eg <= 1e-4
    labels = state.fluid["surface_type"].clone()
    st.set_mode(abi.STEP_FROZEN_GEOMETRY)
    gd = g.detach().numpy()
    xd, vd = x.detach(), v.detach()
    fd = _fd_gravity(st, state, case.config, labels, vt, xd, vd, gd)
    an = g.grad.numpy()
    err = float(np.linalg.norm(fd - an) / np.linalg.norm(fd))
    print(f"mixed body force: fd {fd} adjoint {an} rel err {err}", flush=True)
    assert err <= 1e-4
    st.close()
    ctx.close()
