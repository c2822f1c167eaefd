"""Test infrastructure: the inputs of the diagnostic drivers, rebuilt in Python
so the oracle can be run on exactly the particles the programs used.

The programs are tests/cpp/{complex,hydrostatic}_surface_diagnostics; their
geometries follow the reference's tests/reference/complex_surface_diagnostics.cu
:150-394 and hydrostatic_surface_diagnostics.cu:62-125.  Every coordinate is
formed with the same IEEE double operations in the same order (math.sin /
math.cos / math.sqrt are the C library's, as std::sin etc. are in the
programs), so the positions are bit-identical and the oracle's neighbor lists
and labels can be compared bit for bit.
"""
import math

import numpy as np

PI = 3.14159265358979323846
SPACING = 0.02
BOX = 1.0
CASE_NAMES = ("inclined_plane", "sine_wave", "droplet", "cylinder", "stepped_box")


class _Points:
    def __init__(self):
        self.x, self.y, self.z, self.nx, self.ny, self.nz = ([] for _ in range(6))

    def add(self, px, py, pz, ux, uy, uz):
        self.x.append(px)
        self.y.append(py)
        self.z.append(pz)
        self.nx.append(ux)
        self.ny.append(uy)
        self.nz.append(uz)

    def arrays(self):
        return {k: np.array(getattr(self, k), np.float64) for k in ("x", "y", "z", "nx", "ny", "nz")}


def _unit(a, b, c):
    n = math.sqrt(a * a + b * b + c * c)
    if n <= 0:
        return a, b, c
    return a / n, b / n, c / n


def _centre(i, h):
    return (float(i) + 0.5) * h


def box_walls(w, n, h):
    for k in range(n + 1):
        for j in range(n + 1):
            w.add(0.0, float(j) * h, float(k) * h, 1.0, 0.0, 0.0)
            w.add(float(n) * h, float(j) * h, float(k) * h, -1.0, 0.0, 0.0)
        for i in range(1, n):
            w.add(float(i) * h, 0.0, float(k) * h, 0.0, 1.0, 0.0)
            w.add(float(i) * h, float(n) * h, float(k) * h, 0.0, -1.0, 0.0)
    for j in range(1, n):
        for i in range(1, n):
            w.add(float(i) * h, float(j) * h, 0.0, 0.0, 0.0, 1.0)


def _centres(n, h):
    for k in range(n):
        z = _centre(k, h)
        for j in range(n):
            y = _centre(j, h)
            for i in range(n):
                yield _centre(i, h), y, z


def geometry(name, h=SPACING):
    """(fluid arrays incl. expected normals, expected flags, wall arrays)."""
    n = int(round(BOX / h))
    f, w, expected = _Points(), _Points(), []
    box_walls(w, n, h)

    def fluid(x, y, z, on, ux=0.0, uy=0.0, uz=0.0):
        f.add(x, y, z, ux, uy, uz)
        expected.append(1 if on else 0)

    if name == "inclined_plane":
        slope = 0.25
        u = _unit(-slope, 0.0, 1.0)
        for x, y, z in _centres(n, h):
            top = 0.35 + slope * x
            if z <= top:
                fluid(x, y, z, top - z <= h, *u)
    elif name == "sine_wave":
        a = 0.08
        for x, y, z in _centres(n, h):
            sx, sy = math.sin(2 * PI * x), math.sin(2 * PI * y)
            cx, cy = math.cos(2 * PI * x), math.cos(2 * PI * y)
            top = 0.5 + a * sx * sy
            if z <= top:
                u = _unit(-a * 2 * PI * cx * sy, -a * 2 * PI * sx * cy, 1.0)
                fluid(x, y, z, top - z <= h, *u)
    elif name == "droplet":
        c0, c1, c2, radius = 0.5, 0.5, 0.25, 0.28
        for x, y, z in _centres(n, h):
            dx, dy, dz = x - c0, y - c1, z - c2
            r = math.sqrt(dx * dx + dy * dy + dz * dz)
            if r <= radius and z >= 0:
                fluid(x, y, z, radius - r <= h, *_unit(dx, dy, dz))
    elif name == "cylinder":
        c0, c1, radius, level = 0.5, 0.5, 0.12, 0.5
        for k in range(n + 1):
            z = float(k) * h
            for s in range(48):
                theta = float(s * 2) * PI / 48
                ux, uy = math.cos(theta), math.sin(theta)
                w.add(c0 + radius * ux, c1 + radius * uy, z, ux, uy, 0.0)
        for x, y, z in _centres(n, h):
            if z > level:
                continue
            r = math.sqrt((x - c0) * (x - c0) + (y - c1) * (y - c1))
            if r >= radius + 0.5 * h:
                fluid(x, y, z, level - z <= h, 0.0, 0.0, 1.0)
    elif name == "stepped_box":
        level, step_x, step_top = 0.5, 0.5, 0.18
        for j in range(n + 1):
            for k in range(int(round(step_top / h)) + 1):
                w.add(step_x, float(j) * h, float(k) * h, 1.0, 0.0, 0.0)
        for i in range(int(round(step_x / h)) + 1):
            for j in range(n + 1):
                w.add(float(i) * h, float(j) * h, step_top, 0.0, 0.0, 1.0)
        for x, y, z in _centres(n, h):
            bottom = step_top if x < step_x else 0.0
            if bottom <= z <= level:
                fluid(x, y, z, level - z <= h, 0.0, 0.0, 1.0)
    else:
        raise ValueError(name)
    return f.arrays(), np.array(expected, np.int32), w.arrays()


def hydrostatic_box(h=SPACING, size=BOX, level=0.5):
    """The half-full box: fluid at cell centres below `level`, five wall faces."""
    n = int(round(size / h))
    layers = int(round(level / h))
    f = _Points()
    for k in range(layers):
        for j in range(n):
            for i in range(n):
                f.add(_centre(i, h), _centre(j, h), _centre(k, h), 0.0, 0.0, 0.0)
    w = _Points()
    box_walls(w, n, h)
    return f.arrays(), w.arrays()


def case_config(cases, h=SPACING, anisotropy_threshold=0.033):
    """The programs' surface settings (cell grid from -re over 1 m + 2 re)."""
    re = 3.1 * h
    dim = int(math.ceil((BOX + 2.0 * re) / re))
    c = cases.abi.default_config()
    c.particle_spacing = h
    c.support_radius = re
    c.near_surface_radius = 2.0 * h
    for k in range(3):
        c.cell_origin[k] = -re
        c.cell_dims[k] = dim
    c.cell_size = re
    c.splash_neighbor_threshold = 12
    c.number_density_ratio_threshold = 0.85
    c.air_open_ratio_threshold = 0.33
    c.air_anisotropy_threshold = anisotropy_threshold
    c.include_wall_neighbors = 1
    c.wall_normal_independence_threshold = 0.25
    return c
