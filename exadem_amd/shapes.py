"""Polyhedron shape files (.shp) and the fixed-size shape record of the device path.

`read_shps` follows read_shps / read_shp (src/shape/include/exaDEM/shape_reader.hpp:37-149): a file
holds blocks between lines "<" and ">", with keys name, radius, preCompDone (must be "y"), nv + vertex
lines, ne + edge lines, nf + "count i0 i1 ..." face lines, volume, I/m; other keys (obb.*, position,
orientation, na) are read past. At ">" the bounding box is rebuilt from the vertices as build_OBB does
(src/shape/include/exaDEM/shape_prepro.hpp:26-97): principal axes of the vertex covariance, extents of
the vertex projections, enlarged by the Minkowski radius.

The device record (`xdem_shape`) holds at most 8 vertices, 12 edges and 8 faces of at most 4
vertices, and a box whose axes are the body axes. Shapes beyond those limits, or whose vertex
covariance is not diagonal in the body frame, raise ValueError.
"""
import math

import numpy as np

from .native import Shape

MAX_V, MAX_E, MAX_F, MAX_FV = 8, 12, 8, 4


class ShapeDef:
    """A shape as read from a .shp block, before packing into the device record."""

    def __init__(self, name="undefined"):
        self.name = name
        self.radius = 0.0
        self.vertices = []
        self.edges = []
        self.faces = []
        self.volume = 0.0
        self.inertia_on_mass = (0.0, 0.0, 0.0)

    def obb(self):
        """(center, axes[3x3 rows], extent) of build_OBB over the vertices plus the Minkowski radius."""
        v = np.asarray(self.vertices, np.float64)
        mu = v.mean(axis=0)
        c = v.T @ v - len(v) * np.outer(mu, mu)
        scale = max(1e-300, float(np.abs(c).max()))
        off = c - np.diag(np.diag(c))
        if np.abs(off).max() <= 1e-12 * scale:
            axes = np.eye(3)
        else:
            _, vec = np.linalg.eigh(c)
            axes = vec.T
        proj = v @ axes.T
        lo, hi = proj.min(axis=0), proj.max(axis=0)
        center = axes.T @ (0.5 * (hi + lo))
        extent = 0.5 * (hi - lo) + self.radius
        return center, axes, extent

    def to_record(self):
        if not (1 <= len(self.vertices) <= MAX_V) or len(self.edges) > MAX_E or len(self.faces) > MAX_F:
            raise ValueError(f"shape {self.name}: at most {MAX_V} vertices, {MAX_E} edges and {MAX_F} faces")
        if any(not (3 <= len(f) <= MAX_FV) for f in self.faces):
            raise ValueError(f"shape {self.name}: faces need 3 to {MAX_FV} vertices")
        center, axes, extent = self.obb()
        if np.abs(axes - np.eye(3)).max() > 0 or np.abs(center).max() > 1e-12:
            raise ValueError(f"shape {self.name}: the bounding box must be centred on the body axes")
        s = Shape()
        s.nv, s.ne, s.nf = len(self.vertices), len(self.edges), len(self.faces)
        s.minkowski = self.radius
        s.volume = self.volume
        for a in range(3):
            s.inertia_on_mass[a] = self.inertia_on_mass[a]
            s.obb_extent[a] = float(extent[a])
        for k, v in enumerate(self.vertices):
            for a in range(3):
                s.vertices[k][a] = float(v[a])
        for k, (e0, e1) in enumerate(self.edges):
            s.edges[k][0], s.edges[k][1] = int(e0), int(e1)
        for k, f in enumerate(self.faces):
            s.face_size[k] = len(f)
            for a, vi in enumerate(f):
                s.faces[k][a] = int(vi)
        return s


def _read_block(tokens, pos):
    shp = ShapeDef()
    n = len(tokens)

    def take(count):
        nonlocal pos
        if pos + count > n:
            raise ValueError("read_shp: unexpected end of file or malformed shape definition")
        out = tokens[pos:pos + count]
        pos += count
        return out

    while pos < n:
        key = take(1)[0]
        if key == "name":
            shp.name = take(1)[0]
        elif key == "preCompDone":
            if take(1)[0] != "y":
                raise ValueError("read_shp: preCompDone must be 'y' (volume and I/m are taken from the file)")
        elif key == "radius":
            shp.radius = float(take(1)[0])
        elif key == "volume":
            shp.volume = float(take(1)[0])
        elif key == "I/m":
            shp.inertia_on_mass = tuple(float(x) for x in take(3))
        elif key == "position":
            take(3)
        elif key == "nv":
            nv = int(take(1)[0])
            if nv == 0:
                raise ValueError("read_shp: nv is 0")
            shp.vertices = [tuple(float(x) for x in take(3)) for _ in range(nv)]
        elif key == "ne":
            ne = int(take(1)[0])
            shp.edges = [tuple(int(x) for x in take(2)) for _ in range(ne)]
        elif key == "nf":
            nf = int(take(1)[0])
            faces = []
            for _ in range(nf):
                k = int(take(1)[0])
                if k == 0:
                    raise ValueError("read_shp: empty face")
                faces.append(tuple(int(x) for x in take(k)))
            shp.faces = faces
        elif key == ">":
            return shp, pos
    raise ValueError("read_shp: unexpected end of file or malformed shape definition")


def read_shps(path_or_text):
    """All shapes of a .shp file (or of its text), in file order."""
    text = path_or_text
    if "\n" not in path_or_text and "<" not in path_or_text:
        with open(path_or_text) as fh:
            text = fh.read()
    lines = text.splitlines()
    shapes = []
    k = 0
    while k < len(lines):
        if lines[k].strip() == "<":
            tokens = " ".join(lines[k + 1:]).split()
            shp, used = _read_block(tokens, 0)
            shapes.append(shp)
            # advance past the closing ">" line
            count = 0
            k += 1
            while k < len(lines):
                count += len(lines[k].split())
                k += 1
                if count >= used:
                    break
            continue
        k += 1
    return shapes


def cube(c=1.0, r=0.02):
    """create_cube (src/shape/include/exaDEM/shape/basic_shape/cube.hpp:14-85): vertex, edge and face
    tables and the closed-form volume and inertia of the cube of edge c dilated by r."""
    pi = math.pi
    shp = ShapeDef("cube")
    shp.radius = r
    h = 0.5 * c
    shp.vertices = [(-h, -h, -h), (h, -h, -h), (h, h, -h), (-h, h, -h), (-h, -h, h), (h, -h, h), (h, h, h), (-h, h, h)]
    shp.edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    shp.faces = [(3, 2, 1, 0), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (4, 7, 3, 0), (1, 2, 6, 5)]
    volume = c ** 3 + 6.0 * c * c * r + 3.0 * pi * r * r * c + (4.0 / 3.0) * pi * r ** 3
    inertia = (1.0 / 6.0) * c ** 5
    d = 0.5 * (c + r)
    inertia += 6.0 * ((1.0 / 12.0) * c * c * r * (c * c + r * r) + c * c * r * d * d)
    inertia += 3.0 * pi * r * r * c * (c * c / 12.0 + r * r / 2.0)
    inertia += (8.0 / 15.0) * pi * r ** 5
    shp.volume = volume
    shp.inertia_on_mass = (inertia / volume,) * 3
    return shp


def octahedron(R=0.8, r=0.02):
    """Regular octahedron with its six vertices on the body axes at distance R, dilated by r. Volume by
    the Steiner formula V + S r + (sum over edges of L (pi - dihedral) / 2) r^2 + 4/3 pi r^3; inertia
    over mass of the undilated solid, R^2 / 5."""
    shp = ShapeDef("octahedron")
    shp.radius = r
    shp.vertices = [(R, 0.0, 0.0), (-R, 0.0, 0.0), (0.0, R, 0.0), (0.0, -R, 0.0), (0.0, 0.0, R), (0.0, 0.0, -R)]
    shp.edges = [(0, 2), (0, 3), (0, 4), (0, 5), (1, 2), (1, 3), (1, 4), (1, 5), (2, 4), (2, 5), (3, 4), (3, 5)]
    shp.faces = [(xv, yv, zv) for xv in (0, 1) for yv in (2, 3) for zv in (4, 5)]
    a = R * math.sqrt(2.0)
    dihedral = math.acos(-1.0 / 3.0)
    shp.volume = ((4.0 / 3.0) * R ** 3 + 4.0 * math.sqrt(3.0) * R * R * r + 6.0 * a * (math.pi - dihedral) * r * r
                  + (4.0 / 3.0) * math.pi * r ** 3)
    shp.inertia_on_mass = (R * R / 5.0,) * 3
    return shp


def bounding_radius(shp, h=1.0):
    """compute_max_rcut (src/shape/include/exaDEM/shape.hpp:479-491): largest vertex norm plus the
    Minkowski radius, times the homothety."""
    return h * (max(math.sqrt(sum(x * x for x in v)) for v in shp.vertices) + shp.radius)
