"""lsmps_build_neighbor_lists_noalloc (static per-GPU scratch, no device
allocation) against lsmps_build_neighbor_lists, through the C-ABI: cell lists
and both neighbour tables bit-identical, on two calls in a row where the
second is smaller (the shared scratch then holds the first call's contents),
and the host-side rejection of inputs beyond the static limits."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = -7


def _buffers(torch, abi, rt, case, max_neighbors=256):
    n, nw = case.n_fluid, case.n_wall
    cfg = case.config
    cells = int(cfg.cell_dims[0]) * int(cfg.cell_dims[1]) * int(cfg.cell_dims[2])
    i32 = lambda k: torch.full((max(k, 1),), SENTINEL, dtype=torch.int32, device="cuda:0")
    cap = n * max_neighbors
    t = {"fc": [i32(n), i32(n), i32(cells), i32(cells)],
         "wc": [i32(nw), i32(nw), i32(cells), i32(cells)],
         "fo": i32(n + 1), "fi": i32(cap), "wo": i32(n + 1), "wi": i32(cap)}
    ptr = rt.ptr
    t["fcl"] = abi.CellList(n, cells, *(abi.iptr(ptr(x)) for x in t["fc"]))
    t["wcl"] = abi.CellList(nw, cells, *(abi.iptr(ptr(x)) for x in t["wc"]))
    t["fnl"] = abi.NeighborList(n, cap, abi.iptr(ptr(t["fo"])), abi.iptr(ptr(t["fi"])))
    t["wnl"] = abi.NeighborList(n, cap, abi.iptr(ptr(t["wo"])), abi.iptr(ptr(t["wi"])))
    return t


def _call(lib, entry, ctx, state, case, t):
    fs, ws = state.fluid_soa(), state.wall_soa()
    fn = getattr(lib, entry)
    return fn(ctx.handle, C.byref(fs), C.byref(ws), C.byref(case.config), C.byref(t["fcl"]),
              C.byref(t["wcl"]), C.byref(t["fnl"]), C.byref(t["wnl"]))


def _host(t, n):
    f_off = t["fo"][:n + 1].cpu().numpy()
    w_off = t["wo"][:n + 1].cpu().numpy()
    out = {"fluid_offsets": f_off, "fluid_indices": t["fi"][:int(f_off[-1])].cpu().numpy(),
           "wall_offsets": w_off, "wall_indices": t["wi"][:int(w_off[-1])].cpu().numpy()}
    for side in ("fc", "wc"):
        for k, x in zip(("cell_ids", "sorted_indices", "cell_begin", "cell_end"), t[side]):
            out[f"{side}_{k}"] = x.cpu().numpy()
    return out


def test_noalloc_neighbor_lists_match(rt, abi, cases, oracle):
    import torch
    ctx = rt.Context(abi.PRECISION_FP64)
    lib = ctx.lib
    first = cases.dam_break(block=(30, 20, 20), tank=(60, 20, 40), l0=0.01, seed=4)
    second = cases.hydrostatic(n=10, l0=0.01)
    assert second.n_fluid < first.n_fluid and second.n_wall < first.n_wall
    for case in (first, second):
        print("noalloc", case.name, case.n_fluid, case.n_wall, flush=True)
        state = rt.DeviceState(case)
        a = _buffers(torch, abi, rt, case)
        b = _buffers(torch, abi, rt, case)
        rt.check(_call(lib, "lsmps_build_neighbor_lists", ctx, state, case, a),
                 "build_neighbor_lists")
        rt.check(_call(lib, "lsmps_build_neighbor_lists_noalloc", ctx, state, case, b),
                 "build_neighbor_lists_noalloc")
        torch.cuda.synchronize()
        ha, hb = _host(a, case.n_fluid), _host(b, case.n_fluid)
        for k in ha:
            assert np.array_equal(ha[k], hb[k]), (case.name, k)
        nb = oracle.neighbors(case)
        assert np.array_equal(hb["fluid_offsets"], nb["fluid_offsets"])
        assert np.array_equal(hb["fluid_indices"], nb["fluid_indices"])
        assert np.array_equal(hb["wall_indices"], nb["wall_indices"])
    ctx.close()


def test_noalloc_rejects_oversized_inputs(rt, abi, cases):
    import torch
    ctx = rt.Context(abi.PRECISION_FP64)
    case = cases.hydrostatic(n=6, l0=0.01)
    state = rt.DeviceState(case)
    big = case.copy()
    big.config.cell_dims[0], big.config.cell_dims[1], big.config.cell_dims[2] = 128, 128, 65
    t = _buffers(torch, abi, rt, case)
    cells = 128 * 128 * 65
    t["fcl"].cell_count = cells
    t["wcl"].cell_count = cells
    rc = _call(ctx.lib, "lsmps_build_neighbor_lists_noalloc", ctx, state, big, t)
    msg = ctx.lib.lsmps_last_error().decode()
    assert rc == abi.LSMPS_ERR_CAPACITY, (rc, msg)
    assert f"{cells} cells exceed the static scratch limit of 1048576" in msg, msg
    torch.cuda.synchronize()
    # rejected before any kernel ran: the outputs are untouched
    assert bool((t["fo"] == SENTINEL).all().item())
    assert bool((t["fc"][0] == SENTINEL).all().item())
    ctx.close()
