"""Polyhedra: shape tables, the .shp reader, and the oracle's detection, bounding-box and neighbour
rules, each checked against geometry fixed by hand, brute force or the reference's own shape file."""
import math

import numpy as np
import pytest

import pyoracle
from exadem_amd import inputs, shapes

GOLDEN_SHP = __file__.rsplit("/", 1)[0] + "/golden/octahedron.shp"
IDQ = (1.0, 0.0, 0.0, 0.0)


def quat_axis(axis, angle):
    a = np.asarray(axis, float)
    a /= np.linalg.norm(a)
    s = math.sin(angle / 2)
    return (math.cos(angle / 2), a[0] * s, a[1] * s, a[2] * s)


def rotmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


# ------------------------------------------------------------------------------------ shape tables
def test_shp_reader_on_reference_octahedron():
    """example/polyhedra/rotating_drum/octahedron.shp: the stored volume, I/m and bounding box."""
    (o,) = shapes.read_shps(GOLDEN_SHP)
    assert o.name == "Octahedron" and o.radius == 0.1
    assert (len(o.vertices), len(o.edges), len(o.faces)) == (6, 12, 8)
    assert o.volume == pytest.approx(1.0 / 6.0, abs=1e-16)
    assert o.inertia_on_mass == pytest.approx((0.05, 0.05, 0.05), abs=1e-16)
    # the file also stores the box that build_OBB produced; the reader rebuilds it from the vertices
    text = open(GOLDEN_SHP).read().split()
    k = text.index("obb.extent")
    stored = [float(x) for x in text[k + 1:k + 4]]
    _, axes, ext = o.obb()
    assert np.array_equal(axes, np.eye(3))
    assert ext == pytest.approx(stored, abs=1e-15)
    rec = o.to_record()
    assert list(rec.obb_extent) == pytest.approx(stored, abs=1e-15)
    # every face is a triangle of equal edges (regular octahedron of edge 0.46215...)
    v = np.array(o.vertices)
    L = [np.linalg.norm(v[a] - v[b]) for a, b in o.edges]
    assert max(L) - min(L) < 1e-12


def test_shp_reader_errors():
    with pytest.raises(ValueError):
        shapes.read_shps("<\nname x\npreCompDone n\n>\n")
    with pytest.raises(ValueError):
        shapes.read_shps("<\nname x\nnv 3\n0 0 0\n1 0 0\n")


def _steiner(shp):
    """Volume of the polytope dilated by r from its tables: V + S r + sum_e L_e (pi - theta_e)/2 r^2
    + 4/3 pi r^3, with V and S from the faces and the dihedral angles from the face normals."""
    v = np.array(shp.vertices)
    c = v.mean(axis=0)
    V = S = 0.0
    normals = []
    for f in shp.faces:
        p = v[list(f)]
        for k in range(1, len(f) - 1):
            tri = (p[0], p[k], p[k + 1])
            area_vec = 0.5 * np.cross(tri[1] - tri[0], tri[2] - tri[0])
            S += np.linalg.norm(area_vec)
            V += abs(np.dot(tri[0] - c, np.cross(tri[1] - c, tri[2] - c))) / 6.0
        n = np.cross(p[1] - p[0], p[2] - p[0])
        n /= np.linalg.norm(n)
        if np.dot(n, p[0] - c) < 0:
            n = -n
        normals.append(n)
    E = 0.0
    for a, b in shp.edges:
        adj = [k for k, f in enumerate(shp.faces) if a in f and b in f]
        assert len(adj) == 2
        exterior = math.acos(np.clip(np.dot(normals[adj[0]], normals[adj[1]]), -1, 1))
        E += np.linalg.norm(v[a] - v[b]) * exterior / 2.0
    r = shp.radius
    return V + S * r + E * r * r + 4.0 / 3.0 * math.pi * r ** 3


@pytest.mark.parametrize("make", [lambda: shapes.cube(1.0, 0.3), lambda: shapes.cube(0.7, 0.02),
                                  lambda: shapes.octahedron(0.8, 0.3), lambda: shapes.octahedron(0.8, 0.02)])
def test_dilated_volume_from_tables(make):
    """The closed forms (cube.hpp for the cube, Steiner for the octahedron) against the volume
    computed from the vertex/edge/face tables."""
    shp = make()
    assert shp.volume == pytest.approx(_steiner(shp), rel=1e-12)


def test_inertia_closed_forms():
    c = shapes.cube(1.3, 0.0)
    assert c.volume == pytest.approx(1.3 ** 3, rel=1e-15)
    assert c.inertia_on_mass[0] == pytest.approx(1.3 ** 2 / 6.0, rel=1e-14)
    # octahedron |x|+|y|+|z| <= R: I/m = R^2/5 (Monte Carlo, 2e6 samples)
    R = 0.8
    rng = np.random.default_rng(3)
    p = rng.uniform(-R, R, (2_000_000, 3))
    p = p[np.abs(p).sum(axis=1) <= R]
    assert np.mean(p[:, 1] ** 2 + p[:, 2] ** 2) == pytest.approx(shapes.octahedron(R, 0.0).inertia_on_mass[0],
                                                                  rel=4e-3)
    # the rounded cube's volume (cube.hpp) against Monte Carlo. Its inertia formula (cube.hpp:27-41)
    # gives every face slab the off-axis term V d^2, also the two faces normal to the axis, so it
    # is an approximation: +2.9% at r = 0.02 c, which the shape keeps, as the reference does.
    for cc, r, tol in ((1.0, 0.3, None), (1.0, 0.02, (0.025, 0.035))):
        shp = shapes.cube(cc, r)
        h = 0.5 * cc + r
        q = rng.uniform(-h, h, (3_000_000, 3))
        d = np.linalg.norm(np.maximum(np.abs(q) - 0.5 * cc, 0.0), axis=1)
        q = q[d <= r]
        assert len(q) / 3_000_000 * (2 * h) ** 3 == pytest.approx(shp.volume, rel=3e-3)
        if tol:
            excess = shp.inertia_on_mass[0] / np.mean(q[:, 1] ** 2 + q[:, 2] ** 2) - 1.0
            assert tol[0] < excess < tol[1]


# ------------------------------------------------------------------------------------ detection
@pytest.fixture(scope="module")
def cube_oct():
    return shapes.cube(1.0, 0.02).to_record(), shapes.octahedron(0.8, 0.02).to_record()


def test_vertex_face_closed_form(cube_oct):
    cube, octa = cube_oct
    h = 0.03
    # octahedron vertex 5 (0,0,-0.8) sits at height h above the cube's top face (face 1, z = +0.5)
    po = (0.1, -0.05, 0.5 + 0.8 + h)
    c = pyoracle.poly_detect(octa, po, IDQ, cube, (0, 0, 0), IDQ, 2, 5, 1)
    assert c["contact"]
    assert c["dn"] == pytest.approx(h - 0.04, abs=1e-15)
    assert np.allclose(c["n"], (0, 0, 1), atol=1e-15)
    assert np.allclose(c["cp"], (0.1, -0.05, 0.5 + h - (0.02 + 0.5 * (h - 0.04))), atol=1e-15)
    # projection outside the face: no contact; farther than the radii: no contact
    assert not pyoracle.poly_detect(octa, (0.6, 0.0, 1.33), IDQ, cube, (0, 0, 0), IDQ, 2, 5, 1)["contact"]
    assert not pyoracle.poly_detect(octa, (0.0, 0.0, 1.35), IDQ, cube, (0, 0, 0), IDQ, 2, 5, 1)["contact"]
    # rotated pair: the same configuration seen in a rotated frame keeps dn and rotates n
    q = quat_axis((1, 2, 3), 0.7)
    M = rotmat(q)
    pr = M @ np.array(po)
    cr = pyoracle.poly_detect(octa, pr, q, cube, (0, 0, 0), q, 2, 5, 1)
    assert cr["dn"] == pytest.approx(c["dn"], abs=1e-14)
    assert np.allclose(cr["n"], M @ c["n"], atol=1e-14)


def test_edge_edge_closed_form(cube_oct):
    cube, _ = cube_oct
    g = 0.035
    # cube A edge 0 runs along x at (y,z) = (-0.5,-0.5); cube B (identity) edge 1 runs along y at
    # x = B.x + 0.5, z = B.z - 0.5. Put it across A's edge midpoint, a gap g below it.
    pb = (-0.5, -0.45, -g)
    c = pyoracle.poly_detect(cube, (0, 0, 0), IDQ, cube, pb, IDQ, 3, 0, 1)
    assert c["contact"]
    assert c["dn"] == pytest.approx(g - 0.04, abs=1e-15)
    assert np.allclose(c["n"], (0, 0, 1), atol=1e-15)
    assert np.allclose(c["cp"], (0.0, -0.5, -0.5 - 0.02 - 0.5 * (g - 0.04)), atol=1e-15)
    # parallel edges (A edge 0 and B edge 4, both along x): never an edge-edge contact
    assert not pyoracle.poly_detect(cube, (0, 0, 0), IDQ, cube, (0.2, 0.0, -1.0 - 0.01), IDQ, 3, 0, 4)["contact"]
    # crossing point at an edge end (t = 0) is excluded
    assert not pyoracle.poly_detect(cube, (0, 0, 0), IDQ, cube, (-0.5, 0.0, -g), IDQ, 3, 0, 1)["contact"]


def test_vertex_edge_and_vertex_vertex_closed_form(cube_oct):
    cube, octa = cube_oct
    d = 0.03
    # octahedron vertex 0 (+0.8,0,0) at distance d from the midpoint of cube edge 0, along -y
    po = (-0.8, -0.5 - d, -0.5)
    c = pyoracle.poly_detect(octa, po, IDQ, cube, (0, 0, 0), IDQ, 1, 0, 0)
    assert c["contact"] and c["dn"] == pytest.approx(d - 0.04, abs=1e-15)
    assert np.allclose(c["n"], (0, -1, 0), atol=1e-15)
    # beyond the edge ends: none
    assert not pyoracle.poly_detect(octa, (0.51 - 0.8 + 0.0, -0.5 - d, -0.5), IDQ, cube, (0, 0, 0), IDQ, 1, 0, 0)[
        "contact"]
    # vertex-vertex along a diagonal
    e = 0.01
    u = np.array([1.0, 1.0, 1.0]) / math.sqrt(3)
    pb = np.array([0.5, 0.5, 0.5]) + u * e + np.array([0.8, 0.0, 0.0])  # octahedron vertex 1 (-0.8,0,0) there
    c = pyoracle.poly_detect(cube, (0, 0, 0), IDQ, octa, pb, IDQ, 0, 6, 1)
    assert c["contact"] and c["dn"] == pytest.approx(e - 0.04, abs=1e-14)
    assert np.allclose(c["n"], -u, atol=1e-14)


def _lp_overlap_margin(boxes):
    """Largest t such that some point lies in every box shrunk by t (scipy linprog); > 0 means the
    boxes overlap."""
    from scipy.optimize import linprog
    A, b = [], []
    for c, axes, ext in boxes:
        for k in range(3):
            a = axes[k]
            # a.x - a.c + t <= e ; -a.x + a.c + t <= e
            A.append(list(a) + [1.0])
            b.append(ext[k] + a @ c)
            A.append(list(-a) + [1.0])
            b.append(ext[k] - a @ c)
    res = linprog([0, 0, 0, -1.0], A_ub=A, b_ub=b, bounds=[(None, None)] * 3 + [(None, 10.0)])
    return -res.fun


def test_obb_intersection_against_linear_program(cube_oct):
    """The separating-axis box test against a linear-programming feasibility check on random poses."""
    cube, octa = cube_oct
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(400):
        qa = inputs.random_quaternions(int(rng.integers(1 << 30)), 1)[:, 0]
        qb = inputs.random_quaternions(int(rng.integers(1 << 30)), 1)[:, 0]
        pb = rng.uniform(-1.6, 1.6, 3)
        enl = 0.05
        boxes = []
        for shp, p, q in ((cube, np.zeros(3), qa), (octa, pb, qb)):
            M = rotmat(q)
            ext = np.array(list(shp.obb_extent)) + enl
            boxes.append((p, M.T, ext))
        margin = _lp_overlap_margin(boxes)
        if abs(margin) < 1e-6:
            continue
        got = pyoracle.obb_intersect(cube, np.zeros(3), qa, octa, pb, qb, enl)
        assert got == (margin > 0), (margin, pb)
        checked += 1
    assert checked > 350


def _keys(ii, jj, tt, sw, si, sj):
    return set(zip(ii.tolist(), jj.tolist(), tt.tolist(), sw.tolist(), si.tolist(), sj.tolist()))


def test_feature_list_contains_every_detected_contact():
    """nbh_polyhedron's filters are supersets of detection: every feature pair in contact (found by
    running the detection over all feature pairs of all particle pairs) is in the interaction list."""
    w = inputs.polyhedra(n=600, spacing=1.5, jitter=0.1, seed=9, periodic=(False, False, False), z0=1.2)
    o = pyoracle.Oracle(w.particles, w.domain, w.walls)
    o.set_shapes(w.shapes)
    o.rebuild_poly(w.rcut_inc)
    ii, jj, tt, sw, si, sj, _ = o.poly_interactions()
    pp = tt < 4
    listed = _keys(ii[pp], jj[pp], tt[pp], sw[pp], si[pp], sj[pp])
    P = w.particles
    pos = np.stack([P["rx"], P["ry"], P["rz"]], 1)
    q = np.stack([P["qw"], P["qx"], P["qy"], P["qz"]], 1)
    typ = P["type"]
    rb = P["radius"]
    found = 0
    for a in range(w.n):
        for b in range(a + 1, w.n):
            if np.linalg.norm(pos[a] - pos[b]) > rb[a] + rb[b] + 0.1:
                continue
            sa, sb = w.shapes[typ[a]], w.shapes[typ[b]]
            for ftype, na, nb in ((0, sa.nv, sb.nv), (1, sa.nv, sb.ne), (2, sa.nv, sb.nf), (3, sa.ne, sb.ne)):
                for i in range(na):
                    for j in range(nb):
                        if pyoracle.poly_detect(sa, pos[a], q[a], sb, pos[b], q[b], ftype, i, j)["contact"]:
                            assert (a, b, ftype, 0, i, j) in listed
                            found += 1
            for ftype, nb, na in ((1, sb.nv, sa.ne), (2, sb.nv, sa.nf)):
                for j in range(nb):
                    for i in range(na):
                        if pyoracle.poly_detect(sb, pos[b], q[b], sa, pos[a], q[a], ftype, j, i)["contact"]:
                            assert (a, b, ftype, 1, i, j) in listed
                            found += 1
    assert found > 50
    # pairs are held once, by the smaller id
    assert np.all(ii[pp] < jj[pp])
