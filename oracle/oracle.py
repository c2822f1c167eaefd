"""ctypes wrapper of the FP64 CPU oracle (oracle/lsmps_oracle.cpp).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.  It operates on host numpy arrays of a
`cases.Case` and returns numpy arrays.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblsmps_oracle.so")


def build():
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _p(a):
    return a.ctypes.data if a is not None else 0


class Oracle:
    def __init__(self, abi, threads=None):
        if not os.path.exists(LIB_PATH):
            build()
        self.abi = abi
        self.lib = C.CDLL(LIB_PATH)
        L = self.lib
        L.orc_last_error.restype = C.c_char_p
        L.orc_reference_number_density.restype = C.c_double
        L.orc_reference_number_density.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_stage_seconds.restype = C.POINTER(C.c_double)
        L.orc_solver_setup_seconds.restype = C.c_double
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        if threads:
            L.orc_set_threads(int(threads))
        self.threads = L.orc_set_threads(0)

    # -- helpers -----------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed ({rc}): "
                               f"{self.lib.orc_last_error().decode()}")

    def fluid_soa(self, f):
        a = self.abi
        return a.FluidSoA(f["x"].shape[0], a.dptr(_p(f["x"])), a.dptr(_p(f["y"])),
                          a.dptr(_p(f["z"])), a.dptr(_p(f["vx"])), a.dptr(_p(f["vy"])),
                          a.dptr(_p(f["vz"])), a.dptr(_p(f.get("pressure"))),
                          a.iptr(_p(f.get("surface_type"))))

    def wall_soa(self, w):
        a = self.abi
        return a.WallSoA(w["x"].shape[0], a.dptr(_p(w["x"])), a.dptr(_p(w["y"])),
                         a.dptr(_p(w["z"])), a.dptr(_p(w.get("vx"))), a.dptr(_p(w.get("vy"))),
                         a.dptr(_p(w.get("vz"))), a.dptr(_p(w.get("normal_x"))),
                         a.dptr(_p(w.get("normal_y"))), a.dptr(_p(w.get("normal_z"))))

    def nlist(self, offsets, indices, capacity=None):
        a = self.abi
        cap = capacity if capacity is not None else indices.shape[0]
        return a.NeighborList(offsets.shape[0] - 1, cap, a.iptr(_p(offsets)), a.iptr(_p(indices)))

    # -- stages ------------------------------------------------------------
    def build_cell_list(self, x, y, z, grid):
        a = self.abi
        n = x.shape[0]
        cc = grid.dims[0] * grid.dims[1] * grid.dims[2]
        out = {k: np.zeros(n, np.int32) for k in ("cell_ids", "sorted_indices")}
        out["cell_begin"] = np.zeros(cc, np.int32)
        out["cell_end"] = np.zeros(cc, np.int32)
        cl = a.CellList(n, cc, *(a.iptr(_p(out[k])) for k in
                                 ("cell_ids", "sorted_indices", "cell_begin", "cell_end")))
        self._check(self.lib.orc_build_cell_list(C.c_size_t(n), a.dptr(_p(x)), a.dptr(_p(y)),
                                                 a.dptr(_p(z)), C.byref(grid), C.byref(cl)),
                    "build_cell_list")
        return out

    def neighbors(self, case, max_neighbors=256, config=None):
        a = self.abi
        cfg = config or case.config
        nf, nw = case.n_fluid, case.n_wall
        cc = cfg.cell_count()
        cap = max(nf * max_neighbors, 1)
        bufs = {
            "fc": [np.zeros(max(nf, 1), np.int32), np.zeros(max(nf, 1), np.int32),
                   np.zeros(cc, np.int32), np.zeros(cc, np.int32)],
            "wc": [np.zeros(max(nw, 1), np.int32), np.zeros(max(nw, 1), np.int32),
                   np.zeros(cc, np.int32), np.zeros(cc, np.int32)],
        }
        fc = a.CellList(nf, cc, *(a.iptr(_p(b)) for b in bufs["fc"]))
        wc = a.CellList(nw, cc, *(a.iptr(_p(b)) for b in bufs["wc"]))
        f_off = np.zeros(nf + 1, np.int32)
        f_idx = np.zeros(cap, np.int32)
        w_off = np.zeros(nf + 1, np.int32)
        w_idx = np.zeros(cap, np.int32)
        fn = self.nlist(f_off, f_idx)
        wn = self.nlist(w_off, w_idx)
        fs = self.fluid_soa(case.fluid)
        ws = self.wall_soa(case.walls)
        self._check(self.lib.orc_build_neighbor_lists(C.byref(fs), C.byref(ws), C.byref(cfg),
                                                      C.byref(fc), C.byref(wc), C.byref(fn),
                                                      C.byref(wn)), "build_neighbor_lists")
        return {
            "fluid_offsets": f_off, "fluid_indices": f_idx[:f_off[-1]].copy(),
            "wall_offsets": w_off, "wall_indices": w_idx[:w_off[-1]].copy(),
            "fluid_cells": {"cell_ids": bufs["fc"][0][:nf], "sorted_indices": bufs["fc"][1][:nf],
                            "cell_begin": bufs["fc"][2], "cell_end": bufs["fc"][3]},
            "wall_cells": {"cell_ids": bufs["wc"][0][:nw], "sorted_indices": bufs["wc"][1][:nw],
                           "cell_begin": bufs["wc"][2], "cell_end": bufs["wc"][3]},
        }

    def reference_number_density(self, l0, re, kernel=0):
        return self.lib.orc_reference_number_density(l0, re, kernel)

    def classify(self, case, nb, config=None, diagnostics=True):
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        fluid = dict(case.fluid)
        fluid["surface_type"] = np.zeros(n, np.int32)
        diag = None
        out = {}
        if diagnostics:
            out = {"fluid_neighbor_count": np.zeros(n, np.int32),
                   "wall_neighbor_count": np.zeros(n, np.int32)}
            for k in ("number_density", "number_density_ratio", "anisotropy", "air_open_ratio",
                      "air_anisotropy", "surface_normal_x", "surface_normal_y",
                      "surface_normal_z"):
                out[k] = np.zeros(n)
            diag = a.SurfaceDiagnostics(
                a.iptr(_p(out["fluid_neighbor_count"])), a.iptr(_p(out["wall_neighbor_count"])),
                *(a.dptr(_p(out[k])) for k in (
                    "number_density", "number_density_ratio", "anisotropy", "air_open_ratio",
                    "air_anisotropy", "surface_normal_x", "surface_normal_y",
                    "surface_normal_z")))
        fs = self.fluid_soa(fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        self._check(self.lib.orc_classify_surface(C.byref(fs), C.byref(ws), C.byref(fn),
                                                  C.byref(wn), C.byref(cfg),
                                                  C.byref(diag) if diag is not None else None),
                    "classify_surface")
        out["surface_type"] = fluid["surface_type"]
        return out

    def virtual_light(self, case, nb, config=None):
        """compute_virtual_light_diagnostics: open direction counts and fractions."""
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        count = np.zeros(n, np.int32)
        frac = np.zeros(n)
        fs = self.fluid_soa(case.fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        self._check(self.lib.orc_virtual_light(C.byref(fs), C.byref(ws), C.byref(fn), C.byref(wn),
                                               C.byref(cfg), a.iptr(_p(count)), a.dptr(_p(frac))),
                    "virtual_light")
        return {"open_direction_count": count, "open_fraction": frac}

    def moments(self, case, nb, kind, config=None, want_matrices=False):
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        size = 10 if kind == a.MOMENT_PRESSURE_WALL_NEUMANN_B else 9
        inv = np.zeros(n * size * size)
        mats = np.zeros(n * size * size) if want_matrices else None
        tr = np.zeros(n)
        rg = np.zeros(n)
        info = np.zeros(n, np.int32)
        fs = self.fluid_soa(case.fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        self._check(self.lib.orc_moment_prepare(
            C.c_int(kind), C.byref(fs), C.byref(ws), C.byref(fn), C.byref(wn), C.byref(cfg),
            a.dptr(_p(mats)), a.dptr(_p(inv)), a.dptr(_p(tr)), a.dptr(_p(rg)),
            a.iptr(_p(info))), "moment_prepare")
        out = {"inverse": inv.reshape(n, size, size), "trace": tr, "regularization": rg,
               "info": info, "size": size}
        if want_matrices:
            out["matrices"] = mats.reshape(n, size, size)
        return out

    def provision(self, case, nb, vel_inverse, config=None, surface_type=None):
        a = self.abi
        cfg = config or case.config
        n, nw = case.n_fluid, case.n_wall
        fluid = dict(case.fluid)
        if surface_type is not None:
            fluid["surface_type"] = surface_type
        tmp = {"x": None, "y": None, "z": None, "vx": np.zeros(n), "vy": np.zeros(n),
               "vz": np.zeros(n), "pressure": None, "surface_type": None}
        tmpw = {"x": None, "y": None, "z": None, "vx": np.zeros(nw), "vy": np.zeros(nw),
                "vz": np.zeros(nw)}
        lap = [np.zeros(n) for _ in range(3)]
        fs = self.fluid_soa(fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        ts = a.FluidSoA(n, a.dptr(0), a.dptr(0), a.dptr(0), a.dptr(_p(tmp["vx"])),
                        a.dptr(_p(tmp["vy"])), a.dptr(_p(tmp["vz"])), a.dptr(0), a.iptr(0))
        tw = a.WallSoA(nw, a.dptr(0), a.dptr(0), a.dptr(0), a.dptr(_p(tmpw["vx"])),
                       a.dptr(_p(tmpw["vy"])), a.dptr(_p(tmpw["vz"])), a.dptr(0), a.dptr(0),
                       a.dptr(0))
        vinv = np.ascontiguousarray(vel_inverse, dtype=np.float64)
        self._check(self.lib.orc_provision(C.byref(cfg), C.byref(fs), C.byref(ws), C.byref(fn),
                                           C.byref(wn), a.dptr(_p(vinv)), C.byref(ts),
                                           C.byref(tw), *(a.dptr(_p(l)) for l in lap)),
                    "provision")
        return {"vx": tmp["vx"], "vy": tmp["vy"], "vz": tmp["vz"], "wall_vx": tmpw["vx"],
                "wall_vy": tmpw["vy"], "wall_vz": tmpw["vz"], "laplacian": lap}

    def assemble(self, case, nb, tmp, vel_inverse, pre_inverse, config=None, surface_type=None,
                 neighbor_capacity=None):
        a = self.abi
        cfg = config or case.config
        n, nw = case.n_fluid, case.n_wall
        fluid = dict(case.fluid)
        if surface_type is not None:
            fluid["surface_type"] = surface_type
        cap = neighbor_capacity if neighbor_capacity is not None else nb["fluid_indices"].shape[0]
        nnz = n + cap
        rows = np.zeros(n + 1, np.int32)
        cols = np.zeros(max(nnz, 1), np.int32)
        vals = np.zeros(max(nnz, 1))
        rhs, div, lap = np.zeros(n), np.zeros(n), np.zeros(n)
        pres = np.zeros(n)
        ws_ = a.PpeWorkspace(a.Csr(n, n, nnz, a.iptr(_p(rows)), a.iptr(_p(cols)),
                                   a.dptr(_p(vals))), a.dptr(_p(rhs)), a.dptr(_p(pres)),
                             a.dptr(_p(div)), a.dptr(_p(lap)))
        fs = self.fluid_soa(fluid)
        wsoa = self.wall_soa(case.walls)
        f_idx = nb["fluid_indices"]
        if f_idx.shape[0] < cap:
            f_idx = np.concatenate([f_idx, np.zeros(cap - f_idx.shape[0], np.int32)])
        fn = self.nlist(nb["fluid_offsets"], f_idx, cap)
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        ts = a.FluidSoA(n, a.dptr(0), a.dptr(0), a.dptr(0), a.dptr(_p(tmp["vx"])),
                        a.dptr(_p(tmp["vy"])), a.dptr(_p(tmp["vz"])), a.dptr(0), a.iptr(0))
        tw = a.WallSoA(nw, a.dptr(0), a.dptr(0), a.dptr(0), a.dptr(_p(tmp["wall_vx"])),
                       a.dptr(_p(tmp["wall_vy"])), a.dptr(_p(tmp["wall_vz"])), a.dptr(0),
                       a.dptr(0), a.dptr(0))
        vinv = np.ascontiguousarray(vel_inverse, dtype=np.float64)
        pinv = np.ascontiguousarray(pre_inverse, dtype=np.float64)
        self._check(self.lib.orc_ppe_assemble(C.byref(cfg), C.byref(fs), C.byref(wsoa),
                                              C.byref(fn), C.byref(wn), C.byref(ts), C.byref(tw),
                                              a.dptr(_p(vinv)), a.dptr(_p(pinv)), C.byref(ws_)),
                    "ppe_assemble")
        used = rows[-1]
        return {"row_offsets": rows, "col_indices": cols[:used].copy(),
                "values": vals[:used].copy(), "nnz": nnz, "rhs": rhs, "divergence": div,
                "pressure_laplacian": lap}

    def coarse_aggregation(self, case, nb, labels, factor=4, config=None):
        """Aggregates of factor^3 fluid cells (the SCHWARZ_COARSE coarse space):
        the aggregate of each fluid row, -1 for Surface / Splash rows; aggregates
        are numbered in the linear order of their coarse cells."""
        cfg = config or case.config
        nx, ny, nz = (int(cfg.cell_dims[k]) for k in range(3))
        # the cell list holds the ids in sorted order: cell_ids[k] is the cell of
        # particle sorted_indices[k]
        fc = nb["fluid_cells"]
        cid = np.empty(case.n_fluid, np.int64)
        cid[fc["sorted_indices"]] = fc["cell_ids"]
        cx, cy, cz = cid % nx, (cid // nx) % ny, cid // (nx * ny)
        gx, gy = -(-nx // factor), -(-ny // factor)
        key = ((cz // factor) * gy + cy // factor) * gx + cx // factor
        active = (labels == 0) | (labels == 1)
        agg = np.full(case.n_fluid, -1, np.int32)
        uniq, inv = np.unique(key[active], return_inverse=True)
        agg[active] = inv.astype(np.int32)
        return agg, int(uniq.shape[0])

    def coarse_apply(self, row_offsets, col_indices, values, v, agg, count):
        """z = P Ac^-1 P^T v with Ac = P^T A P over the aggregation."""
        a = self.abi
        n = v.shape[0]
        rows = np.ascontiguousarray(row_offsets, np.int32)
        cols = np.ascontiguousarray(col_indices, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        vv = np.ascontiguousarray(v, np.float64)
        z = np.zeros(n)
        csr = a.Csr(n, n, vals.shape[0], a.iptr(_p(rows)), a.iptr(_p(cols)), a.dptr(_p(vals)))
        g = np.ascontiguousarray(agg, np.int32)
        self.lib.orc_set_coarse(a.iptr(_p(g)), C.c_int64(n), C.c_int32(count))
        try:
            self._check(self.lib.orc_coarse_apply(C.byref(csr), a.dptr(_p(vv)), a.dptr(_p(z))),
                        "coarse_apply")
        finally:
            self.lib.orc_set_coarse(None, C.c_int64(0), C.c_int32(0))
        return z

    def solve(self, row_offsets, col_indices, values, rhs, options=None, coarse=None):
        """PPE solve; `coarse` = (aggregate per row, count) enables the coarse
        space of a SCHWARZ_COARSE solve."""
        a = self.abi
        n = rhs.shape[0]
        if coarse is not None:
            g = np.ascontiguousarray(coarse[0], np.int32)
            self.lib.orc_set_coarse(a.iptr(_p(g)), C.c_int64(n), C.c_int32(int(coarse[1])))
        try:
            return self._solve(row_offsets, col_indices, values, rhs, options)
        finally:
            if coarse is not None:
                self.lib.orc_set_coarse(None, C.c_int64(0), C.c_int32(0))

    def _solve(self, row_offsets, col_indices, values, rhs, options=None):
        a = self.abi
        n = rhs.shape[0]
        opt = options or a.default_solver_options()
        rows = np.ascontiguousarray(row_offsets, np.int32)
        cols = np.ascontiguousarray(col_indices, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros(n)
        csr = a.Csr(n, n, vals.shape[0], a.iptr(_p(rows)), a.iptr(_p(cols)), a.dptr(_p(vals)))
        rep = a.SolveReport()
        rc = self.lib.orc_ppe_solve(C.byref(csr), a.dptr(_p(b)), a.dptr(_p(x)), C.byref(opt),
                                    C.byref(rep))
        if rc not in (0, a.LSMPS_ERR_NOT_CONVERGED):
            self._check(rc, "ppe_solve")
        out = rep.as_dict()
        out["setup_seconds"] = self.lib.orc_solver_setup_seconds()
        return x, out

    def schwarz_apply(self, row_offsets, col_indices, values, v):
        """z = M^-1 v for the Schwarz preconditioner of the CSR matrix; returns
        (z, blocks, rows summed over blocks)."""
        a = self.abi
        n = v.shape[0]
        rows = np.ascontiguousarray(row_offsets, np.int32)
        cols = np.ascontiguousarray(col_indices, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        vv = np.ascontiguousarray(v, np.float64)
        z = np.zeros(n)
        csr = a.Csr(n, n, vals.shape[0], a.iptr(_p(rows)), a.iptr(_p(cols)), a.dptr(_p(vals)))
        nb, nr = C.c_int(), C.c_int()
        self._check(self.lib.orc_schwarz_apply(C.byref(csr), a.dptr(_p(vv)), a.dptr(_p(z)),
                                               C.byref(nb), C.byref(nr)), "schwarz_apply")
        return z, nb.value, nr.value

    def particle_shifting(self, case, nb, config=None):
        """compute_particle_shifting_kernel alone; returns [ps_x, ps_y, ps_z]."""
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        ps = [np.zeros(n) for _ in range(3)]
        v3 = a.Vector3(n, *(a.dptr(_p(x)) for x in ps))
        fs = self.fluid_soa(case.fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        self._check(self.lib.orc_particle_shifting(C.byref(cfg), C.byref(fs), C.byref(ws),
                                                   C.byref(fn), C.byref(wn), C.byref(v3)),
                    "particle_shifting")
        return ps

    def anti_penetration(self, case, nb, config=None):
        """anti_penetration_kernel alone on a copy of the fluid; returns the new fluid dict."""
        cfg = config or case.config
        fluid = {k: v.copy() for k, v in case.fluid.items()}
        fs = self.fluid_soa(fluid)
        ws = self.wall_soa(case.walls)
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        self._check(self.lib.orc_anti_penetration(C.byref(cfg), C.byref(fs), C.byref(ws),
                                                  C.byref(wn)), "anti_penetration")
        return fluid

    def smooth_velocity(self, case, nb, config=None):
        """smooth_velocity_kernel + copy on a copy of the fluid; returns (vx, vy, vz)."""
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        fluid = {k: v.copy() for k, v in case.fluid.items()}
        sm = [np.zeros(n) for _ in range(3)]
        v3 = a.Vector3(n, *(a.dptr(_p(x)) for x in sm))
        fs = self.fluid_soa(fluid)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        self._check(self.lib.orc_smooth_velocity(C.byref(cfg), C.byref(fs), C.byref(fn),
                                                 C.byref(v3)), "smooth_velocity")
        return fluid["vx"], fluid["vy"], fluid["vz"]

    def correction(self, case, nb, tmp, pressure, pre_inverse, config=None, surface_type=None):
        """DevicePressureCorrection::apply on a copy of the case's fluid; returns the new state."""
        a = self.abi
        cfg = config or case.config
        n = case.n_fluid
        fluid = {k: v.copy() for k, v in case.fluid.items()}
        if surface_type is not None:
            fluid["surface_type"] = surface_type.copy()
        p = np.ascontiguousarray(pressure, np.float64).copy()
        vecs = {k: [np.zeros(n) for _ in range(3)] for k in ("grad", "ps", "smooth")}
        v3 = {k: a.Vector3(n, *(a.dptr(_p(x)) for x in v)) for k, v in vecs.items()}
        fs = self.fluid_soa(fluid)
        ws = self.wall_soa(case.walls)
        fn = self.nlist(nb["fluid_offsets"], nb["fluid_indices"])
        wn = self.nlist(nb["wall_offsets"], nb["wall_indices"])
        ts = a.FluidSoA(n, a.dptr(0), a.dptr(0), a.dptr(0), a.dptr(_p(tmp["vx"])),
                        a.dptr(_p(tmp["vy"])), a.dptr(_p(tmp["vz"])), a.dptr(0), a.iptr(0))
        pinv = np.ascontiguousarray(pre_inverse, dtype=np.float64)
        self._check(self.lib.orc_correction_apply(C.byref(cfg), C.byref(fs), C.byref(ws),
                                                  C.byref(fn), C.byref(wn), C.byref(ts),
                                                  a.dptr(_p(p)), a.dptr(_p(pinv)),
                                                  C.byref(v3["grad"]), C.byref(v3["ps"]),
                                                  C.byref(v3["smooth"])), "correction")
        return {"fluid": fluid, "pressure": p, "grad": vecs["grad"], "ps": vecs["ps"]}

    def pipeline(self, case, options=None, max_neighbors=256):
        """All stages of one step, returning every intermediate (for parity tests)."""
        a = self.abi
        nb = self.neighbors(case, max_neighbors)
        cls = self.classify(case, nb)
        st = cls["surface_type"]
        c2 = case.copy()
        c2.fluid["surface_type"] = st
        vel = self.moments(c2, nb, a.MOMENT_VELOCITY_WALL_DIRICHLET_A)
        pre = self.moments(c2, nb, a.MOMENT_PRESSURE_WALL_NEUMANN_A)
        tmp = self.provision(c2, nb, vel["inverse"])
        ppe = self.assemble(c2, nb, tmp, vel["inverse"], pre["inverse"],
                            neighbor_capacity=case.n_fluid * max_neighbors)
        opt = options or a.default_solver_options()
        coarse = None
        if opt.preconditioner == a.PRECOND_SCHWARZ_COARSE:
            coarse = self.coarse_aggregation(case, nb, st, opt.coarse_cells or 4)
        p, rep = self.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], ppe["rhs"],
                            opt, coarse=coarse)
        cor = self.correction(c2, nb, tmp, p, pre["inverse"])
        return {"neighbors": nb, "surface": cls, "velocity_moment": vel,
                "pressure_moment": pre, "provision": tmp, "ppe": ppe, "pressure": p,
                "report": rep, "correction": cor}

    def step(self, case, options=None, max_neighbors=256):
        """orc_step on a copy of the case; returns (new fluid dict, report, seconds per stage)."""
        a = self.abi
        fluid = {k: v.copy() for k, v in case.fluid.items()}
        fs = self.fluid_soa(fluid)
        ws = self.wall_soa(case.walls)
        opt = options or a.default_solver_options()
        rep = a.SolveReport()
        rc = self.lib.orc_step(C.byref(case.config), C.byref(fs), C.byref(ws),
                               C.c_size_t(max_neighbors), C.byref(opt), C.byref(rep),
                               *([None] * 8))
        if rc not in (0, a.LSMPS_ERR_NOT_CONVERGED):
            self._check(rc, "step")
        secs = self.lib.orc_stage_seconds()
        return fluid, rep.as_dict(), [secs[k] for k in range(6)]
