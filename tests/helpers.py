"""Builders shared by the CPU (oracle) and GPU (parity) tests."""
import numpy as np


def config_from(abi, overrides):
    cfg = abi.default_config()
    for k, v in overrides.items():
        if isinstance(v, list):
            arr = getattr(cfg, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def make_case(cases_mod, abi, name, fluid_xyz, walls=None, config=None, velocity=None,
              surface_type=None, pressure=None):
    x, y, z = (np.asarray(a, np.float64) for a in fluid_xyz)
    n = x.shape[0]
    vx, vy, vz = velocity if velocity is not None else (np.zeros(n),) * 3
    fluid = {"x": x.copy(), "y": y.copy(), "z": z.copy(),
             "vx": np.asarray(vx, np.float64).copy(), "vy": np.asarray(vy, np.float64).copy(),
             "vz": np.asarray(vz, np.float64).copy(),
             "pressure": np.zeros(n) if pressure is None else np.asarray(pressure, np.float64).copy(),
             "surface_type": np.zeros(n, np.int32) if surface_type is None
             else np.asarray(surface_type, np.int32).copy()}
    walls = walls or {}
    nw = len(walls.get("x", []))
    w = {}
    for k in ("x", "y", "z", "vx", "vy", "vz", "normal_x", "normal_y", "normal_z"):
        w[k] = np.asarray(walls.get(k, np.zeros(nw)), np.float64).copy()
    return cases_mod.Case(name, fluid, w, config or abi.default_config())


def all_pairs_neighbors(n, n_wall_rows=None):
    """All-pairs fluid CSR without self and an empty wall CSR (the unit tests' fixture)."""
    offsets = np.arange(n + 1, dtype=np.int32) * (n - 1)
    indices = np.array([j for i in range(n) for j in range(n) if j != i], dtype=np.int32)
    return {"fluid_offsets": offsets, "fluid_indices": indices,
            "wall_offsets": np.zeros(n + 1, np.int32), "wall_indices": np.zeros(0, np.int32)}


def lattice(points):
    p = np.asarray(points, np.float64)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def brute_force_neighbors(px, py, pz, sx, sy, sz, radius, exclude_self):
    """Neighbor sets by O(N*M) distance tests (same FP64 expression as the oracle)."""
    r2 = radius * radius
    rows = []
    for i in range(px.shape[0]):
        dx = px[i] - sx
        dy = py[i] - sy
        dz = pz[i] - sz
        d2 = dx * dx + dy * dy + dz * dz
        mask = d2 <= r2
        if exclude_self:
            mask[i] = False
        rows.append(np.nonzero(mask)[0])
    return rows


def parse_vtk(path):
    """Legacy ASCII VTK file -> points, vectors, real and int scalars (numpy)."""
    toks = open(path).read().split("\n")
    out = {"vectors": {}, "real": {}, "int": {}}
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.startswith("POINTS"):
            n = int(t.split()[1])
            out["points"] = np.array([[float(v) for v in toks[i + 1 + k].split()] for k in range(n)])
            i += n
        elif t.startswith("VECTORS"):
            name = t.split()[1]
            out["vectors"][name] = np.array([[float(v) for v in toks[i + 1 + k].split()]
                                             for k in range(n)])
            i += n
        elif t.startswith("SCALARS"):
            _, name, typ, _ = t.split()
            vals = [toks[i + 2 + k] for k in range(n)]
            out["int" if typ == "int" else "real"][name] = np.array(
                [int(v) for v in vals] if typ == "int" else [float(v) for v in vals])
            i += n + 1
        i += 1
    return out
