"""Failure and edge paths of the HIP library, through the C-ABI.

- Neighbour capacity overflow (require_neighbor_capacity,
  src/neighbor/neighbor_search.cu:22-30,323-328): LSMPS_ERR_CAPACITY with the
  reference's message, raised on the host after the count pass and before the
  write pass touches the index buffer.
- A particle with no neighbours at all: its moment matrices are zero and the
  regularisation adds lambda * 1 * I (trace 0; moment_matrix.cu:297-308), so the
  factorisation succeeds with inverse I / lambda; its label is Splash and its
  PPE row is a Dirichlet row (ppe_matrix.cu:262-276).
- MomentMatrixViews that are not prepared Type-A matrices of the expected kind
  (validate_ppe_moment_matrix, ppe_matrix.cu:108-122;
  validate_pressure_moment_matrix, pressure_correction.cu:96-111):
  LSMPS_ERR_NOT_PREPARED with the reference's messages, nothing computed.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _tensors(torch, n, cap, fill):
    i32 = lambda k, v=0: torch.full((max(k, 1),), v, dtype=torch.int32, device="cuda:0")
    return i32(n + 1), i32(cap, fill)


def test_neighbor_capacity_overflow(rt, abi, cases):
    import torch
    case = cases.hydrostatic(n=8, l0=0.01)
    n, nw = case.n_fluid, case.n_wall
    cfg = case.config
    cells = int(cfg.cell_dims[0]) * int(cfg.cell_dims[1]) * int(cfg.cell_dims[2])
    ctx = rt.Context(abi.PRECISION_FP64)
    state = rt.DeviceState(case)
    i32 = lambda k: torch.zeros(max(k, 1), dtype=torch.int32, device="cuda:0")
    fc = [i32(n), i32(n), i32(cells), i32(cells)]
    wc = [i32(nw), i32(nw), i32(cells), i32(cells)]
    cap = 4 * n  # an interior lattice particle has ~122 fluid neighbours
    sentinel = -7
    fo, fi = _tensors(torch, n, cap, sentinel)
    wo, wi = _tensors(torch, n, cap, sentinel)
    ptr = rt.ptr
    fcl = abi.CellList(n, cells, *(abi.iptr(ptr(t)) for t in fc))
    wcl = abi.CellList(nw, cells, *(abi.iptr(ptr(t)) for t in wc))
    fnl = abi.NeighborList(n, cap, abi.iptr(ptr(fo)), abi.iptr(ptr(fi)))
    wnl = abi.NeighborList(n, cap, abi.iptr(ptr(wo)), abi.iptr(ptr(wi)))
    fs, ws = state.fluid_soa(), state.wall_soa()
    print("capacity overflow: n", n, "capacity", cap, flush=True)
    rc = ctx.lib.lsmps_build_neighbor_lists(ctx.handle, C.byref(fs), C.byref(ws), C.byref(cfg),
                                            C.byref(fcl), C.byref(wcl), C.byref(fnl),
                                            C.byref(wnl))
    msg = ctx.lib.lsmps_last_error().decode()
    assert rc == abi.LSMPS_ERR_CAPACITY, (rc, msg)
    assert "Neighbor list capacity exceeded: required" in msg and f"capacity {cap}" in msg, msg
    torch.cuda.synchronize()
    # the write pass never ran: every index slot still holds the sentinel
    assert bool((fi == sentinel).all().item())
    # the required total in the message is the true neighbour count
    offsets = fo[:n + 1].cpu().numpy()
    assert f"required {int(offsets[-1])}," in msg
    assert offsets[-1] > cap
    ctx.close()

    # the fused stepper reports the same condition
    with pytest.raises(rt.LsmpsError) as err:
        rt.run_step_on_device(case, precision=abi.PRECISION_FP64, max_neighbors=4)
    assert err.value.code == abi.LSMPS_ERR_CAPACITY


def _isolated_case(cases):
    """Hydrostatic 8^3 box plus one fluid particle 10 l0 above the free surface."""
    base = cases.hydrostatic(n=8, l0=0.01)
    f = {k: v.copy() for k, v in base.fluid.items()}
    for k in f:
        extra = np.array([0.0], dtype=f[k].dtype)
        if k in ("x", "y"):
            extra[0] = 0.04
        if k == "z":
            extra[0] = 0.18
        f[k] = np.concatenate([f[k], extra])
    c = cases.Case("hydrostatic_isolated", f, {k: v.copy() for k, v in base.walls.items()},
                   base.config.copy())
    cases.derive_cell_grid(c.config, c.fluid, c.walls)
    return c


def test_isolated_particle_regularised(rt, abi, cases, oracle):
    case = _isolated_case(cases)
    n = case.n_fluid
    iso = n - 1
    ref = oracle.pipeline(case)
    nb = ref["neighbors"]
    assert nb["fluid_offsets"][iso + 1] == nb["fluid_offsets"][iso]
    assert nb["wall_offsets"][iso + 1] == nb["wall_offsets"][iso]
    lam = case.config.lsmps_regularization
    # oracle: zero moments + lambda * I (trace 0), inverse I / lambda, info 0
    for kind in (abi.MOMENT_VELOCITY_WALL_DIRICHLET_A, abi.MOMENT_PRESSURE_WALL_NEUMANN_A,
                 abi.MOMENT_PRESSURE_WALL_NEUMANN_B, abi.MOMENT_FLUID_ONLY_A):
        c2 = case.copy()
        c2.fluid["surface_type"] = ref["surface"]["surface_type"]
        mo = oracle.moments(c2, nb, kind)
        size = mo["size"]
        assert mo["trace"][iso] == 0.0
        assert mo["regularization"][iso] == lam
        assert mo["info"][iso] == 0
        np.testing.assert_allclose(mo["inverse"][iso], np.eye(size) / lam, rtol=1e-14)
    assert ref["surface"]["surface_type"][iso] == abi.SPLASH

    print("isolated particle case n", n, flush=True)
    g = rt.run_step_on_device(case, precision=abi.PRECISION_FP64)
    assert np.array_equal(g["fluid_offsets"], nb["fluid_offsets"])
    assert np.array_equal(g["fluid_indices"], nb["fluid_indices"])
    assert np.array_equal(g["surface_type"], ref["surface"]["surface_type"])
    ppe = g["ppe"]
    r0 = ppe["row_offsets"][iso]
    assert ppe["col_indices"][r0] == iso and ppe["values"][r0] == 1.0
    assert ppe["rhs"][iso] == 0.0
    assert g["pressure"][iso] == 0.0
    assert g["report"]["converged"] == 1
    assert rel_l2(g["pressure"], ref["pressure"]) <= 1e-9
    # the Splash particle falls freely: u = u* = dt g (no pressure gradient)
    dt = case.config.time_step
    assert abs(g["fluid"]["vz"][iso] - dt * case.config.gravity[2]) <= 1e-15


def _view_call(rt, abi, ctx, state, case, vm, pm, which):
    import torch
    n = case.n_fluid
    cap = n * 8
    dev = "cuda:0"
    tmp = {k: torch.zeros(n, dtype=torch.float64, device=dev) for k in ("vx", "vy", "vz")}
    ptr = rt.ptr
    fs, wsoa = state.fluid_soa(), state.wall_soa()
    empty = abi.NeighborList(n, cap, abi.iptr(ptr(torch.zeros(n + 1, dtype=torch.int32, device=dev))),
                             abi.iptr(ptr(torch.zeros(cap, dtype=torch.int32, device=dev))))
    if which == "assemble":
        tf = abi.FluidSoA()
        tf.count = n
        tf.vx, tf.vy, tf.vz = (abi.dptr(ptr(tmp[k])) for k in ("vx", "vy", "vz"))
        tw = abi.WallSoA()
        tw.count = 0
        ws = abi.PpeWorkspace()
        return ctx.lib.lsmps_ppe_assemble_views(ctx.handle, C.byref(case.config), C.byref(fs),
                                                C.byref(wsoa), C.byref(empty), C.byref(empty),
                                                C.byref(tf), C.byref(tw), C.byref(vm),
                                                C.byref(pm), C.byref(ws))
    grad = abi.Vector3(n, *(abi.dptr(ptr(tmp[k])) for k in ("vx", "vy", "vz")))
    p = torch.zeros(n, dtype=torch.float64, device=dev)
    return ctx.lib.lsmps_compute_pressure_gradient_view(
        ctx.handle, C.byref(fs), C.byref(wsoa), C.byref(empty), C.byref(empty),
        abi.dptr(ptr(p)), C.byref(case.config), C.byref(pm), C.byref(grad))


def test_unprepared_views_rejected(rt, abi, cases):
    case = cases.hydrostatic(n=6, l0=0.01)
    n = case.n_fluid
    ctx = rt.Context(abi.PRECISION_FP64)
    state = rt.DeviceState(case)

    def view(kind, ready=1, size=9, basis=0, count=n, inverse=1):
        v = abi.MomentView()
        v.particle_count, v.matrix_size, v.basis_kind, v.kind, v.is_ready = count, size, basis, kind, ready
        v.inverse_matrices = abi.dptr(inverse)  # never dereferenced: validation fails first
        return v

    vel = abi.MOMENT_VELOCITY_WALL_DIRICHLET_A
    pre = abi.MOMENT_PRESSURE_WALL_NEUMANN_A
    bad = [
        ("assemble", view(vel, ready=0), view(pre), "PPE requires a prepared velocity Type-A moment matrix"),
        ("assemble", view(pre), view(pre), "PPE requires a prepared velocity Type-A moment matrix"),
        ("assemble", view(vel), view(pre, size=10, basis=1), "PPE requires a prepared pressure Type-A moment matrix"),
        ("assemble", view(vel), view(pre, count=n - 1), "PPE requires a prepared pressure Type-A moment matrix"),
        ("assemble", view(vel, inverse=0), view(pre), "PPE moment matrix inverse buffer is missing for velocity"),
        ("gradient", None, view(vel), "Pressure correction requires a prepared pressure Type-A moment matrix"),
        ("gradient", None, view(pre, ready=0), "Pressure correction requires a prepared pressure Type-A moment matrix"),
        ("gradient", None, view(pre, inverse=0), "Pressure correction pressure moment inverse buffer is missing"),
    ]
    for which, vm, pm, message in bad:
        rc = _view_call(rt, abi, ctx, state, case, vm if vm is not None else view(vel), pm, which)
        msg = ctx.lib.lsmps_last_error().decode()
        assert rc == abi.LSMPS_ERR_NOT_PREPARED, (which, rc, msg)
        assert message in msg, (message, msg)
    ctx.close()


def test_context_destroyed_before_its_objects(rt, abi, cases):
    """A context destroyed before the stepper and moment matrix created on it
    stays alive until they are destroyed; the next calls see no stale error."""
    case = cases.hydrostatic(n=6, l0=0.01)
    ctx = rt.Context(abi.PRECISION_FP64)
    state = rt.DeviceState(case)
    st = rt.Stepper(ctx, state, case.config, 256, abi.default_solver_options())
    m = C.c_void_p()
    rt.check(ctx.lib.lsmps_moment_create(ctx.handle, C.c_size_t(case.n_fluid),
                                         C.byref(case.config), C.byref(m)), "moment_create")
    lib = ctx.lib
    ctx.close()  # the caller's handle first
    rep = st.step()
    assert rep["converged"] == 1
    rt.check(lib.lsmps_moment_destroy(m), "moment_destroy")
    st.close()
    g = rt.run_step_on_device(case, precision=abi.PRECISION_FP64)
    assert g["report"]["converged"] == 1
