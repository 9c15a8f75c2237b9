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
    w = inputs.polyhedra(density=2500.0, n=600, spacing=1.5, jitter=0.1, seed=9, periodic=(False, False, False), z0=1.2)
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


# ------------------------------------------------------------------------ contact assembly
# The force and moment that polyhedron contacts put onto both bodies (polyhedron/contact.hpp:124-211 for
# particle pairs: +f on the item's particle i, the vertex side, -f on j, moments (cp - r) x (+-f) from each
# centre, meff = mi mj / (mi + mj); contact.hpp:219-297 for a vertex against a driver: f on the particle,
# moment (cp - r) x f). Checked against rigid-body statics and kinematics worked out here, against the
# conservation of momentum and angular momentum, and with planted mistakes that must be caught.
KN = 1e5


def _poly_set(specs, shapes_defs):
    """Particle dictionary for polyhedra: specs = [(type, position, quaternion, velocity, spin)], ids in
    list order, density 2500, homothety 1."""
    n = len(specs)
    pos = np.array([s[1] for s in specs], float)
    rb = np.array([shapes.bounding_radius(shapes_defs[s[0]]) for s in specs])
    p = inputs.sphere_set(pos[:, 0], pos[:, 1], pos[:, 2], rb, 2500.0)
    typ = np.array([s[0] for s in specs], np.uint32)
    p["type"] = typ
    p["mass"] = 2500.0 * np.array([shapes_defs[t].volume for t in typ])
    for a in ("inx", "iny", "inz"):
        p[a] = p["mass"] * np.array([shapes_defs[t].inertia_on_mass[0] for t in typ])
    for k, key in enumerate(("qw", "qx", "qy", "qz")):
        p[key] = np.array([s[2][k] for s in specs], float)
    for k, key in enumerate(("vx", "vy", "vz")):
        p[key] = np.array([s[3][k] for s in specs], float)
    for k, key in enumerate(("vrotx", "vroty", "vrotz")):
        p[key] = np.array([s[4][k] for s in specs], float)
    return p


def _first_pass(p, shapes_defs, walls=(), box=((0, 0, 0), (12, 12, 12)), dt=1e-5, rcut_inc=0.1):
    """nbh_polyhedron + reset_force_moment + one contact_polyhedron pass; returns the particle fields by
    id and the interaction list."""
    from exadem_amd.native import contact_params, domain
    cell = 2 * float(np.max(p["radius"])) + rcut_inc
    o = pyoracle.Oracle(p, domain(box[0], box[1], (False, False, False), cell), list(walls))
    o.set_shapes([d.to_record() for d in shapes_defs])
    o.rebuild_poly(rcut_inc)
    o.reset_force_moment()
    cp = contact_params(kn=KN, mu=0.3, damp_rate=0.9)
    o.contact_poly(cp, cp, dt)
    q = o.particles()
    order = np.argsort(q["id"])
    out = {k: v[order] for k, v in q.items()}
    return out, o.poly_interactions()


def _vec(q, key, i):
    return np.array([q[key + a][i] for a in "xyz"])


def test_cube_face_down_on_a_floor():
    """A cube (edge 1, Minkowski radius 0.02) at rest on a floor wall at overlap d: its four bottom vertices
    each feel kn d, so F = (0, 0, 4 kn d), and the symmetric contact points leave no moment."""
    cube = shapes.cube(1.0, 0.02)
    d = 2e-3
    c = (5.0, 5.0, 0.5 + 0.02 - d)
    p = _poly_set([(0, c, IDQ, (0, 0, 0), (0, 0, 0))], [cube])
    q, inter = _first_pass(p, [cube], walls=[{"offset": 0.0, "normal": (0.0, 0.0, 1.0)}])
    F, M = _vec(q, "f", 0), _vec(q, "mom", 0)
    assert F[2] == pytest.approx(4 * KN * d, rel=1e-12)
    assert abs(F[0]) + abs(F[1]) < 1e-9 * F[2]
    assert np.max(np.abs(M)) < 1e-9 * F[2]
    # the vertex-wall candidates are the four bottom vertices (type 5, partner = wall 0)
    assert sorted(inter[4][inter[2] == 5].tolist()) == [0, 1, 2, 3]


@pytest.mark.parametrize("moving", [False, True])
def test_cube_on_one_vertex(moving):
    """A tilted cube touching the floor with one vertex only. At rest: F = kn d along the floor normal
    and M = (cp - r) x F with cp = vertex - n (r_m + dn / 2). Moving and spinning: the contact point's
    velocity v + w x (cp - r) sets the damping and friction, and still M = (cp - r) x F."""
    cube = shapes.cube(1.0, 0.02)
    rm, d, dt = 0.02, 1e-3, 1e-5
    qc = quat_axis((1.0, -2.0, 0.5), 0.9)
    R = rotmat(qc)
    body = np.array(cube.vertices)
    z = (body @ R.T)[:, 2]
    low = int(np.argmin(z))
    assert np.sort(z)[1] - z[low] > 0.15  # one vertex well below the others
    c = np.array([5.0, 5.0, 0.0])
    c[2] = rm - d - z[low]
    v = np.array([0.3, -0.2, -0.1]) if moving else np.zeros(3)
    w = np.array([1.0, 2.0, -0.5]) if moving else np.zeros(3)
    p = _poly_set([(0, c, qc, v, w)], [cube])
    q, _ = _first_pass(p, [cube], walls=[{"offset": 0.0, "normal": (0.0, 0.0, 1.0)}], dt=dt)
    F, M = _vec(q, "f", 0), _vec(q, "mom", 0)
    vert = c + R @ body[low]
    n = np.array([0.0, 0.0, 1.0])
    dn = vert[2] - rm
    assert dn == pytest.approx(-d, abs=1e-12)
    cpt = vert - n * (rm + 0.5 * dn)
    m = float(p["mass"][0])
    if not moving:
        assert F == pytest.approx([0.0, 0.0, KN * d], rel=1e-12, abs=1e-9)
    else:
        vc = v + np.cross(w, cpt - c)  # velocity of the body at the contact point; the floor is at rest
        vrel = -vc
        vn = vrel @ n
        damp = 0.9 * 2.0 * math.sqrt(KN * m)
        fn = max(0.0, KN * d + damp * vn) * n
        ft = 0.8 * KN * dt * (vrel - vn * n)
        assert np.linalg.norm(ft) < 0.3 * KN * d  # below the Coulomb cap
        assert F == pytest.approx(fn + ft, rel=1e-9, abs=1e-9 * np.linalg.norm(fn))
    assert M == pytest.approx(np.cross(cpt - c, F), rel=1e-9, abs=1e-9 * np.linalg.norm(F))


@pytest.mark.parametrize("small_owns", [True, False])
def test_cube_on_cube_face_contact(small_owns):
    """A small cube (edge 0.5) turned about z and set off-centre on a big cube's top face (edge 1), pressing
    by d and closing at speed u. Its four bottom vertices touch the big top face: four vertex-face items,
    held by the lower id, so swap = 0 when the small cube owns the pair and 1 when the big one does. Each
    contact gives f = (kn d + damp u) z with damp = 0.9 * 2 sqrt(kn meff), meff = ms mb / (ms + mb), +f on
    the small cube and -f on the big one; the small cube's moments cancel, the big cube takes
    -4 (cp_mean - c_big) x f. Both orders give the same forces and moments."""
    small, big = shapes.cube(0.5, 0.02), shapes.cube(1.0, 0.02)
    rm, d, u, dt = 0.02, 1e-3, 0.01, 1e-5
    qs = quat_axis((0.0, 0.0, 1.0), 0.3)
    cb = np.array([5.0, 5.0, 5.0])
    cs = np.array([5.1, 4.95, 5.0 + 0.5 + 0.25 + 2 * rm - d])
    spec_s = (0, cs, qs, (0.0, 0.0, -u), (0.0, 0.0, 0.0))
    spec_b = (1, cb, IDQ, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    specs = [spec_s, spec_b] if small_owns else [spec_b, spec_s]
    p = _poly_set(specs, [small, big])
    q, inter = _first_pass(p, [small, big], dt=dt)
    i_s, i_b = (0, 1) if small_owns else (1, 0)
    ii, jj, tt, sw, si, sj, _ = inter
    pp = tt < 4
    assert sorted(tt[pp].tolist()) == [2, 2, 2, 2] or np.count_nonzero((tt == 2) & pp) == 4
    vf = pp & (tt == 2)
    assert np.all(sw[vf] == (0 if small_owns else 1))
    ms, mb = float(p["mass"][i_s]), float(p["mass"][i_b])
    meff = ms * mb / (ms + mb)
    fz = KN * d + 0.9 * 2.0 * math.sqrt(KN * meff) * u
    Fs, Fb = _vec(q, "f", i_s), _vec(q, "f", i_b)
    Ms, Mb = _vec(q, "mom", i_s), _vec(q, "mom", i_b)
    assert Fs == pytest.approx([0.0, 0.0, 4 * fz], rel=1e-12, abs=1e-9 * fz)
    assert Fb == pytest.approx(-Fs, rel=1e-14, abs=1e-12 * fz)
    assert np.max(np.abs(Ms)) < 1e-9 * fz
    verts = cs + np.array(small.vertices)[:4] @ rotmat(qs).T
    cp_mean = verts.mean(axis=0) - np.array([0.0, 0.0, rm - 0.5 * d])
    f = np.array([0.0, 0.0, fz])
    assert Mb == pytest.approx(-4 * np.cross(cp_mean - cb, f), rel=1e-9, abs=1e-9 * fz)


def _loose_packing(n=343, seed=21):
    """Gravity-free, wall-free cubes and octahedra on a tight jittered lattice (many VV/VE/VF/EE contacts of
    both swap sides) with random velocities and spins, non-periodic."""
    w = inputs.polyhedra(density=2500.0, n=n, spacing=1.42, jitter=0.08, seed=seed, periodic=(False, False, False), z0=1.5,
                         sigma_v=0.3)
    p = {k: v.copy() for k, v in w.particles.items()}
    spin = inputs.normal(seed + 7, 3, 3 * w.n, 0.0, 2.0).reshape(3, w.n)
    p["vrotx"], p["vroty"], p["vrotz"] = (np.ascontiguousarray(spin[a]) for a in range(3))
    return w, p


def _balance(w, p):
    """Total force and total angular momentum source sum(r x F + M) of one pass, each relative to its
    sum of magnitudes, and the (type, swap) pairs of the contacts met."""
    from exadem_amd.native import contact_params
    o = pyoracle.Oracle(p, w.domain, [])
    o.set_shapes(w.shapes)
    o.rebuild_poly(w.rcut_inc)
    o.reset_force_moment()
    cp = contact_params(kn=KN, mu=0.3, damp_rate=0.9)
    o.contact_poly(cp, cp, 5e-5)  # a long step: friction forces of a few percent of the normal ones
    q = o.particles()
    r = np.stack([q["rx"], q["ry"], q["rz"]], 1)
    F = np.stack([q["fx"], q["fy"], q["fz"]], 1)
    M = np.stack([q["momx"], q["momy"], q["momz"]], 1)
    L = np.cross(r, F) + M
    fs = np.linalg.norm(F.sum(0)) / np.sum(np.linalg.norm(F, axis=1))
    ls = np.linalg.norm(L.sum(0)) / np.sum(np.linalg.norm(np.cross(r, F), axis=1) + np.linalg.norm(M, axis=1))
    _, _, tt, sw, _, _, fr = o.poly_interactions()
    act = np.any(fr != 0, axis=1) & (tt < 4)
    return fs, ls, set(zip(tt[act].tolist(), sw[act].tolist()))


def test_packing_conserves_momentum_and_angular_momentum():
    """Every contact puts +f and -f at one contact point into two bodies, so over a whole packing without
    gravity, walls or periodic images sum F = 0 and sum (r x F + M) = 0 to round-off, whatever the shapes,
    feature types, swap sides, velocities and spins."""
    w, p = _loose_packing()
    fs, ls, kinds = _balance(w, p)
    assert {(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (3, 0)} <= kinds or len(kinds) >= 5, kinds
    assert fs < 1e-13
    assert ls < 1e-13


@pytest.mark.parametrize("mutation", [5, 6, 7])
def test_planted_assembly_mistakes_are_caught(mutation):
    """Planted in the oracle: 5 the swap side ignored (the vertex side of a swap item pushed the wrong way,
    moments from the wrong centres), 6 the vertex side's moment arm taken from the vertex instead of the
    contact point, 7 meff = m_i. Each must fail the conservation test or the cube-on-cube statics."""
    caught = []
    pyoracle.set_mutation(mutation)
    try:
        w, p = _loose_packing()
        fs, ls, _ = _balance(w, p)
        caught.append(fs > 1e-12 or ls > 1e-12)  # round-off leaves ~1e-17
        for small_owns in (True, False):
            try:
                test_cube_on_cube_face_contact(small_owns)
                caught.append(False)
            except AssertionError:
                caught.append(True)
    finally:
        pyoracle.set_mutation(0)
    assert any(caught), mutation
