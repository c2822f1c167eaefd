"""Captured IDR cycles (LSMPS_FLAG_GRAPH) against the direct launches, through
the C-ABI: both run under the same device-side loop control (two cycles in
flight), so pressures are bit-identical and the iteration and reliable-update
counts equal.  The graphs are replayed over several steps with
the program's own eager work and synchronisations between them, and every
replay is checked against the launch path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(cases):
    return [cases.hydrostatic(n=20, l0=0.01),
            cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0),
            cases.splash(n=24, l0=0.01, seed=3, top_layers=3)]


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_graph_cycles_match_launches(rt, abi, cases, precision, idx):
    case = _cases(cases)[idx]
    prec = abi.PRECISION_FP64 if precision == "FP64" else abi.PRECISION_MIXED
    print("graph", case.name, precision, flush=True)
    eager = rt.run_step_on_device(case, precision=prec)
    graph = rt.run_step_on_device(case, precision=prec, flags=abi.FLAG_GRAPH)
    re, rg = eager["report"], graph["report"]
    print("eager", re, "graph", rg, flush=True)
    assert re["graph_cycles"] == 0
    assert rg["graph_cycles"] >= 2
    for k in ("iterations", "refinements", "converged", "relative_residual"):
        assert re[k] == rg[k], k
    assert np.array_equal(eager["pressure"], graph["pressure"])
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(eager["fluid"][k], graph["fluid"][k]), k


def test_graph_replays_between_eager_work(rt, abi, cases):
    import torch
    case = cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0)
    ref = rt.run_step_on_device(case, precision=abi.PRECISION_MIXED)
    ctx = rt.Context(abi.PRECISION_MIXED, flags=abi.FLAG_GRAPH)
    state = rt.DeviceState(case)
    st = rt.Stepper(ctx, state, case.config, 256, abi.default_solver_options())
    snap = state.snapshot()
    junk = torch.ones(1 << 20, dtype=torch.float64, device="cuda:0")
    for replay in range(4):
        state.restore(snap)
        rep = st.step()
        torch.cuda.synchronize()
        print("replay", replay, rep, flush=True)
        assert rep["graph_cycles"] >= 2
        assert rep["iterations"] == ref["report"]["iterations"]
        p = state.fluid["pressure"].cpu().numpy()
        assert np.array_equal(p, ref["pressure"]), replay
        # eager work of the program between the replays
        junk.mul_(1.0001).add_(float(replay))
        torch.cuda.synchronize()
    st.close()
    ctx.close()


def test_graph_iteration_limit(rt, abi, cases):
    """A solve stopped by max_iterations: same iterate and counts in both paths."""
    case = cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0)
    opts = abi.default_solver_options()
    opts.max_iterations = 40
    eager = rt.run_step_on_device(case, precision=abi.PRECISION_FP64, solver=opts)
    graph = rt.run_step_on_device(case, precision=abi.PRECISION_FP64, solver=opts,
                                  flags=abi.FLAG_GRAPH)
    assert eager["report"]["converged"] == 0 and graph["report"]["converged"] == 0
    assert eager["report"]["iterations"] == graph["report"]["iterations"] >= 40
    assert np.array_equal(eager["pressure"], graph["pressure"])


def test_direct_loop_across_solves_with_timing(rt, abi, cases):
    """The direct launches' loop control and SpMV event sets carried across
    solves of one stepper: tolerances 1e-6, 1e-10, 1e-6 in turn (a halted
    cycle of each solve is still in flight when the next solve starts), with
    kernel timing on.  Each step equals a fresh context's step at its
    tolerance bit for bit, and the timed SpMV launches are counted."""
    import ctypes as C
    import torch
    case = cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=0)
    loose = abi.default_solver_options()
    loose.tolerance = 1e-6
    tight = abi.default_solver_options()
    refs = {1e-6: rt.run_step_on_device(case, precision=abi.PRECISION_MIXED, solver=loose),
            1e-10: rt.run_step_on_device(case, precision=abi.PRECISION_MIXED, solver=tight)}
    assert refs[1e-6]["report"]["iterations"] < refs[1e-10]["report"]["iterations"]
    ctx = rt.Context(abi.PRECISION_MIXED, flags=abi.FLAG_KERNEL_TIMING)
    state = rt.DeviceState(case)
    st = rt.Stepper(ctx, state, case.config, 256, loose)
    snap = state.snapshot()
    lib = rt.load()
    rt.check(lib.lsmps_context_kernel_stats(ctx.handle, None, None, None, 1))
    for tol, opts in ((1e-6, loose), (1e-10, tight), (1e-6, loose)):
        st.set_solver(opts)
        state.restore(snap)
        rep = st.step()
        torch.cuda.synchronize()
        print("tolerance", tol, rep, flush=True)
        ref = refs[tol]
        assert rep["graph_cycles"] == 0
        for k in ("iterations", "refinements", "converged", "relative_residual"):
            assert rep[k] == ref["report"][k], (tol, k)
        assert np.array_equal(state.fluid["pressure"].cpu().numpy(), ref["pressure"]), tol
    launches, ms, nbytes = C.c_int64(), C.c_double(), C.c_double()
    rt.check(lib.lsmps_context_kernel_stats(ctx.handle, C.byref(launches), C.byref(ms),
                                            C.byref(nbytes), 0))
    print("timed SpMV launches", launches.value, "ms", ms.value, flush=True)
    assert launches.value > 0 and ms.value > 0.0 and nbytes.value > 0.0
    st.close()
    ctx.close()
