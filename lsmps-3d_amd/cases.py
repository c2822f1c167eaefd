"""Deterministic synthetic particle cases for the per-step LSMPS pipeline.

Geometry conventions follow the reference's generator
(tool/generate_hydrostatic_box.cpp:50-99): fluid particles sit at cell centres
(i + 1/2) l0, wall particles on the integer lattice, wall normals point into the
fluid, and corner/edge wall particles take the x-face normal first, then the
y-face normal, then the bottom normal (generate_hydrostatic_box.cpp:72-87).
The reference generator writes a single wall layer; these cases use `layers`
layers stacked outward.  The cell grid follows derive_cell_grid_from_particles
(src/app/simulation_driver.cu:83-108).

Randomness comes from SplitMix64 (Steele, Lea & Flood 2014) so every case is
reproducible bit for bit on any host: u = (x >> 11) * 2^-53.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """Outputs start+1 .. start+count of the SplitMix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed, count, start=0):
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, count):
    """Box-Muller standard normals from consecutive uniform pairs."""
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = np.maximum(u[0::2], 2.0 ** -53)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:count]


@dataclass
class Case:
    name: str
    fluid: dict            # x y z vx vy vz pressure (float64) surface_type (int32)
    walls: dict            # x y z vx vy vz normal_x normal_y normal_z (float64)
    config: abi.Config
    extra: dict = field(default_factory=dict)

    @property
    def n_fluid(self):
        return int(self.fluid["x"].shape[0])

    @property
    def n_wall(self):
        return int(self.walls["x"].shape[0])

    def copy(self):
        return Case(self.name, {k: v.copy() for k, v in self.fluid.items()},
                    {k: v.copy() for k, v in self.walls.items()}, self.config.copy(),
                    dict(self.extra))


def lattice_block(nx, ny, nz, l0, origin=(0.0, 0.0, 0.0)):
    """Fluid block at cell centres, z-major then y then x (generate_hydrostatic_box.cpp:58-67)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = (ix.ravel().astype(np.float64) + 0.5) * l0 + origin[0]
    y = (iy.ravel().astype(np.float64) + 0.5) * l0 + origin[1]
    z = (iz.ravel().astype(np.float64) + 0.5) * l0 + origin[2]
    return x, y, z


def tank_walls(nx, ny, nz, l0, layers=3, origin=(0.0, 0.0, 0.0)):
    """Open-top tank [0,nx]x[0,ny]x[0,nz] (units of l0) with `layers` wall layers.

    Wall particles are the integer lattice points of the thickened box that are
    not strictly inside the fluid region.  Normal priority: x faces, then y
    faces, then the bottom (generate_hydrostatic_box.cpp:72-87).
    """
    k = layers - 1
    iz, iy, ix = np.meshgrid(np.arange(-k, nz + 1), np.arange(-k, ny + k + 1),
                             np.arange(-k, nx + k + 1), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    interior = (ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0)
    ix, iy, iz = ix[~interior], iy[~interior], iz[~interior]
    n = ix.shape[0]
    nxv = np.zeros(n)
    nyv = np.zeros(n)
    nzv = np.zeros(n)
    xlo, xhi = ix <= 0, ix >= nx
    ylo = ~xlo & ~xhi & (iy <= 0)
    yhi = ~xlo & ~xhi & (iy >= ny)
    bot = ~xlo & ~xhi & ~ylo & ~yhi
    nxv[xlo] = 1.0
    nxv[xhi] = -1.0
    nyv[ylo] = 1.0
    nyv[yhi] = -1.0
    nzv[bot] = 1.0
    return {
        "x": ix.astype(np.float64) * l0 + origin[0],
        "y": iy.astype(np.float64) * l0 + origin[1],
        "z": iz.astype(np.float64) * l0 + origin[2],
        "vx": np.zeros(n), "vy": np.zeros(n), "vz": np.zeros(n),
        "normal_x": nxv, "normal_y": nyv, "normal_z": nzv,
    }


def derive_cell_grid(config, fluid, walls):
    """derive_cell_grid_from_particles (src/app/simulation_driver.cu:83-108)."""
    xs = [a for a in (fluid["x"], walls["x"]) if a.size]
    ys = [a for a in (fluid["y"], walls["y"]) if a.size]
    zs = [a for a in (fluid["z"], walls["z"]) if a.size]
    if not xs:
        raise ValueError("Cannot derive cell grid from empty particle input")
    lo = [min(float(a.min()) for a in v) for v in (xs, ys, zs)]
    hi = [max(float(a.max()) for a in v) for v in (xs, ys, zs)]
    pad = config.support_radius
    for k in range(3):
        config.cell_origin[k] = lo[k] - pad
    for k in range(3):
        span = (hi[k] + pad) - config.cell_origin[k]
        config.cell_dims[k] = int(math.ceil(span / config.cell_size)) + 1
    return config


def case_config(l0, re_factor=3.1, weight_kernel=abi.WEIGHT_LINEAR):
    c = abi.default_config()
    c.particle_spacing = l0
    c.support_radius = re_factor * l0
    c.cell_size = c.support_radius
    c.near_surface_radius = 1.5 * l0
    c.weight_kernel = weight_kernel
    return c


def _fluid_dict(x, y, z, vx=None, vy=None, vz=None):
    n = x.shape[0]
    return {
        "x": np.ascontiguousarray(x, dtype=np.float64),
        "y": np.ascontiguousarray(y, dtype=np.float64),
        "z": np.ascontiguousarray(z, dtype=np.float64),
        "vx": np.zeros(n) if vx is None else np.ascontiguousarray(vx, dtype=np.float64),
        "vy": np.zeros(n) if vy is None else np.ascontiguousarray(vy, dtype=np.float64),
        "vz": np.zeros(n) if vz is None else np.ascontiguousarray(vz, dtype=np.float64),
        "pressure": np.zeros(n),
        "surface_type": np.zeros(n, dtype=np.int32),
    }


def jitter(x, y, z, l0, amplitude, seed):
    """Uniform +-amplitude*l0 per coordinate, particle-major draw order."""
    n = x.shape[0]
    u = uniform01(seed, 3 * n).reshape(n, 3)
    d = (2.0 * u - 1.0) * (amplitude * l0)
    return x + d[:, 0], y + d[:, 1], z + d[:, 2]


def block_in_tank(name, block, tank, l0, jitter_amplitude=0.0, seed=0, layers=3,
                  weight_kernel=abi.WEIGHT_LINEAR):
    """A lattice fluid block in the (0,0,0) corner of an open tank."""
    x, y, z = lattice_block(*block, l0)
    if jitter_amplitude > 0.0:
        x, y, z = jitter(x, y, z, l0, jitter_amplitude, seed)
    walls = tank_walls(*tank, l0, layers=layers)
    fluid = _fluid_dict(x, y, z)
    cfg = case_config(l0, weight_kernel=weight_kernel)
    derive_cell_grid(cfg, fluid, walls)
    return Case(name, fluid, walls, cfg, {"block": block, "tank": tank, "l0": l0})


def hydrostatic(n=20, l0=0.01, layers=3, weight_kernel=abi.WEIGHT_LINEAR):
    """20^3 hydrostatic box, open top, 3 wall layers (case C1)."""
    return block_in_tank(f"hydrostatic_{n}", (n, n, n), (n, n, n), l0, 0.0, 0, layers,
                         weight_kernel)


def dam_break(block=(100, 50, 50), tank=(200, 50, 100), l0=0.005, seed=0, amplitude=0.05,
              weight_kernel=abi.WEIGHT_LINEAR):
    """Jittered dam-break block in the corner of a tank (case C2; M1 at 100^3)."""
    nb = block[0] * block[1] * block[2]
    return block_in_tank(f"dam_break_{nb}", block, tank, l0, amplitude, seed, 3, weight_kernel)


def metric_case():
    """M1: 100x100x100 = 1,000,000 fluid in a 200x100x200-cell tank, l0 = 5 mm."""
    return dam_break(block=(100, 100, 100), tank=(200, 100, 200), l0=0.005, seed=0)


def sloshing(n=50, l0=0.005, seed=2, tank_height=None):
    """Case C4: z-stretched box with surface H + 0.1 H sin(2 pi x / Lx), u_x = 0.1 sin(pi z / H)."""
    x, y, z = lattice_block(n, n, n, l0)
    lx = n * l0
    h = n * l0
    z = z * (1.0 + 0.1 * np.sin(2.0 * np.pi * x / lx))
    vx = 0.1 * np.sin(np.pi * z / h)
    walls = tank_walls(n, n, tank_height or int(1.5 * n), l0)
    fluid = _fluid_dict(x, y, z, vx=vx)
    cfg = case_config(l0)
    derive_cell_grid(cfg, fluid, walls)
    target = normal(seed, 3 * x.shape[0]).reshape(-1, 3) * 0.01
    return Case(f"sloshing_{n}", fluid, walls, cfg, {"v_target": target, "l0": l0})


def adjoint_stress(n=12, l0=0.005, seed=5, weight_kernel=abi.WEIGHT_LINEAR):
    """Small jittered block that reaches every branch of the step and of its
    adjoint: pairs closer than the shifting distance (jitter 0.3 l0), particles
    inside the wall clearance, random velocities (some into the walls), a
    shifting scale large enough to reach the displacement cap, and a few top
    particles lifted by 3.5 l0 (splash labels next to surface rows: four top
    particles at least 3.5 l0 from the side walls and 5 l0 from each other)."""
    assert n >= 10
    x, y, z = lattice_block(n, n, n, l0)
    x, y, z = jitter(x, y, z, l0, 0.3, seed)
    z = z.copy()
    lo, hi = 3, n - 4
    pick = [((n - 1) * n + iy) * n + ix for iy in (lo, hi) for ix in (lo, hi)]
    z[pick] += 3.5 * l0
    v = normal(seed + 1, 3 * x.shape[0]).reshape(-1, 3) * 0.05
    walls = tank_walls(n, n, n + 4, l0)
    fluid = _fluid_dict(x, y, z, vx=v[:, 0].copy(), vy=v[:, 1].copy(), vz=v[:, 2].copy())
    cfg = case_config(l0, weight_kernel=weight_kernel)
    cfg.ps_displacement_scale = 0.5
    derive_cell_grid(cfg, fluid, walls)
    target = normal(seed + 2, 3 * x.shape[0]).reshape(-1, 3) * 0.01
    return Case(f"adjoint_stress_{n}", fluid, walls, cfg, {"v_target": target, "l0": l0})


def splash(n=160, l0=0.0025, seed=3, top_layers=8):
    """Case C5: cube block whose top layers are lifted by U(2,10) l0 into droplets."""
    x, y, z = lattice_block(n, n, n, l0)
    top = z > (n - top_layers) * l0
    lift = 2.0 + 8.0 * uniform01(seed, int(top.sum()))
    z = z.copy()
    z[top] += lift * l0
    walls = tank_walls(n, n, n + 16, l0)
    fluid = _fluid_dict(x, y, z)
    cfg = case_config(l0)
    derive_cell_grid(cfg, fluid, walls)
    return Case(f"splash_{n}", fluid, walls, cfg, {"l0": l0})


def batched_tanks(count=64, seed=1, l0=0.005, min_fluid=2000, max_fluid=60000,
                  amplitude=0.05):
    """C3: `count` independent tanks, fluid counts log-uniform in [min_fluid,
    max_fluid], block aspect ratios x:y:z = U(1,4):1:U(1,2), alternating
    hydrostatic (block fills the open tank, no jitter) and dam-break (block in
    the corner of a 2x-wider, 2x-taller tank, jittered) with shared l0 / re.

    Draws (SplitMix64, `seed`): three uniforms per tank in tank order (size,
    x ratio, z ratio); dam-break tank k is jittered with seed `seed` + 1 + k.
    Every tank is generated at the origin with its own cell grid; the batched
    case concatenates them, tank by tank.
    """
    u = uniform01(seed, 3 * count).reshape(count, 3)
    tanks = []
    for k in range(count):
        n_target = math.exp(math.log(min_fluid) + u[k, 0] * (math.log(max_fluid) - math.log(min_fluid)))
        ax = 1.0 + 3.0 * u[k, 1]
        az = 1.0 + 1.0 * u[k, 2]
        ny = max(2, int(round((n_target / (ax * az)) ** (1.0 / 3.0))))
        nx = max(2, int(round(ax * ny)))
        nz = max(2, int(round(az * ny)))
        if k % 2 == 0:
            t = block_in_tank(f"tank{k}_hydrostatic", (nx, ny, nz), (nx, ny, nz), l0)
        else:
            t = block_in_tank(f"tank{k}_dam_break", (nx, ny, nz), (2 * nx, ny, 2 * nz), l0,
                              amplitude, seed + 1 + k)
        tanks.append(t)
    fluid = {key: np.concatenate([t.fluid[key] for t in tanks]) for key in tanks[0].fluid}
    walls = {key: np.concatenate([t.walls[key] for t in tanks]) for key in tanks[0].walls}
    fo = np.concatenate([[0], np.cumsum([t.n_fluid for t in tanks])]).astype(np.int64)
    wo = np.concatenate([[0], np.cumsum([t.n_wall for t in tanks])]).astype(np.int64)
    cfg = tanks[0].config.copy()  # shared physics; per-tank grids in `extra`
    return Case(f"batched_{count}", fluid, walls, cfg,
                {"fluid_offsets": fo, "wall_offsets": wo,
                 "grids": [t.config.grid() for t in tanks], "tanks": tanks, "l0": l0})


def by_name(name):
    table = {
        "C1": hydrostatic,
        "C2": dam_break,
        "C3": batched_tanks,
        "M1": metric_case,
        "C4": sloshing,
        "C5": splash,
    }
    return table[name]()
