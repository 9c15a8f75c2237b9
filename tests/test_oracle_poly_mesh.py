"""Polyhedra against the triangle-mesh driver in the oracle (nbh_polyhedron_driver.hpp:60-200,
ContactLawRShapeDriverFunc in polyhedron/contact.hpp:300-432), checked against geometry computed here by
other means:

- a tilted cube whose lowest vertex touches the tray's apex vertex, a ridge edge or a flat face by d:
  kn d along the vertex, edge-bisector or face normal (the ridge edge is listed once per adjacent face);
- the apex pressing into a cube's bottom face (type 12): kn d upwards at rest; moving, the stored friction
  is the negated driver-side friction, so the next pass starts from the opposite sign (kept as the
  reference does; the planted mistake that drops the negation is caught);
- a cube of homothety 2 over the apex: the mesh-vertex gate uses the shape box without the homothety, so
  the apex is not a candidate (the planted mistake that drops the gate is caught);
- random poses of cubes and octahedra pressed into the tray's bump: the contact set, forces and moments
  against a brute-force evaluation over every particle-feature / mesh-feature pair, with closest points
  from bounded least squares (scipy) instead of the detection formulas.
"""
import math

import numpy as np
import pytest
from scipy.optimize import lsq_linear

import pyoracle
import mesh_cases as mc
from exadem_amd import inputs, shapes
from exadem_amd.native import contact_params, domain

KN, RV = 1e5, 0.1


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


def one(shape, c, q, h=1.0, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)):
    """A single polyhedron of type 0 at c with orientation q and homothety h, density 2500."""
    p = inputs.sphere_set(np.array([c[0]]), np.array([c[1]]), np.array([c[2]]),
                          np.array([shapes.bounding_radius(shape, h)]), 2500.0,
                          np.array([float(v[0])]), np.array([float(v[1])]), np.array([float(v[2])]))
    p["type"] = np.zeros(1, np.uint32)
    p["homothety"] = np.array([h])
    p["mass"] = np.array([2500.0 * shape.volume * h ** 3])
    for a in ("inx", "iny", "inz"):
        p[a] = p["mass"] * shape.inertia_on_mass[0] * h * h
    for k, key in enumerate(("qw", "qx", "qy", "qz")):
        p[key] = np.array([float(q[k])])
    for k, key in enumerate(("vrotx", "vroty", "vrotz")):
        p[key] = np.array([float(w[k])])
    return p


def passes(p, shape, rm, dt=1e-5, n=1, mutation=0):
    """nbh_polyhedron against the tray mesh, then n times (reset_force_moment, contact_polyhedron) at fixed
    positions; returns [(F, M)] per pass and the interaction rows after the last pass."""
    cell = 2 * float(np.max(p["radius"])) + RV
    o = pyoracle.Oracle(p, domain((0.0, 0.0, -1.0), (20.0, 20.0, 7.0), (False, False, False), cell), [])
    o.set_shapes([shape.to_record()])
    m = mc.tray_mesh(rm)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], rm)
    cp = contact_params(kn=KN, mu=0.3, damp_rate=0.9)
    out = []
    pyoracle.set_mutation(mutation)
    try:
        o.rebuild_poly(RV)
        for _ in range(n):
            o.reset_force_moment()
            o.contact_poly(cp, cp, dt)
            q = o.particles()
            out.append((np.array([q["fx"][0], q["fy"][0], q["fz"][0]]),
                        np.array([q["momx"][0], q["momy"][0], q["momz"][0]])))
        rows = o.poly_interactions()
    finally:
        pyoracle.set_mutation(0)
    return out, rows


CUBE = shapes.cube(1.0, 0.02)
QTILT = quat_axis((1.0, 1.0, 0.0), 0.9553)  # a body diagonal almost vertical: one vertex 0.58 below the next


@pytest.mark.parametrize("where,kind", [("vertex", 7), ("edge", 8), ("face", 9)])
def test_lowest_vertex_probes(where, kind):
    """The cube's lowest vertex at the sphere probes of the tray with radius r_m of the cube: overlap d with
    the apex vertex, a ridge edge or a flat face, and no other feature in contact."""
    rm, ra, d = 0.01, 0.02, 1e-3
    R = rotmat(QTILT)
    body = np.array(CUBE.vertices)
    low = int(np.argmin((body @ R.T)[:, 2]))
    P = mc.probes(ra, d, rm)[where]
    c = P - R @ body[low]
    [(F, M)], rows = passes(one(CUBE, c, QTILT), CUBE, rm)
    apex = np.array([10.0, 10.0, 1.0])
    direction = {"vertex": P - apex, "face": np.array([0.0, 0.0, 1.0]),
                 "edge": P - (apex + np.array([12.0, 10.0, 0.0])) / 2}[where]
    n = direction / np.linalg.norm(direction)
    mult = 2 if where == "edge" else 1  # the ridge edge is listed once per adjacent face
    assert F == pytest.approx(mult * KN * d * n, rel=1e-9, abs=1e-9 * KN * d)
    cpt = P - n * (ra - 0.5 * d)
    assert M == pytest.approx(np.cross(cpt - c, F), rel=1e-9, abs=1e-9 * KN * d)
    _, jj, tt, _, si, _, _ = rows
    assert np.all(jj[tt >= 7] == 6) and np.count_nonzero((tt == kind) & (si == low)) == mult


def _apex_under_face(h, u=0.0, n=1, mutation=0):
    rm, ra, d = 0.01, 0.02 * h, 1e-3
    c = np.array([10.2, 9.9, 1.0 + 0.5 * h + ra + rm - d])
    return c, passes(one(CUBE, c, (1.0, 0.0, 0.0, 0.0), h=h, v=(u, 0.0, 0.0)), CUBE, rm, n=n, mutation=mutation)


def test_apex_into_cube_face():
    """Type 12 at rest: F = kn d z and M = (cp - c) x F with cp = apex + z (r_m - d/2)."""
    c, ([(F, M)], rows) = _apex_under_face(1.0)
    f = np.array([0.0, 0.0, KN * 1e-3])
    assert F == pytest.approx(f, rel=1e-9, abs=1e-9 * KN * 1e-3)
    cpt = np.array([10.0, 10.0, 1.0 + 0.01 - 0.5e-3])
    assert M == pytest.approx(np.cross(cpt - c, f), rel=1e-9, abs=1e-9 * KN * 1e-3)
    tt, si = rows[2], rows[4]
    bottom = [k for k, f in enumerate(CUBE.faces) if all(CUBE.vertices[v][2] < 0 for v in f)]
    assert np.count_nonzero(tt == 12) == 1 and si[tt == 12][0] == bottom[0]
    assert np.count_nonzero(tt >= 7) == np.count_nonzero(tt == 12)


@pytest.mark.parametrize("mutation", [0, 9])
def test_apex_friction_is_stored_negated(mutation):
    """The cube slides along x at u over the apex. The law runs with the mesh vertex as body i, so its
    tangential force is kt dt u x on the driver and -kt dt u x on the cube; the reference then stores the
    negated friction, which the next pass takes as the driver's previous friction: the sliding force on
    the cube alternates between -kt dt u and 0 at fixed positions. Planted mistake 9 (no negation) makes
    the second pass -2 kt dt u and must be caught."""
    u, dt = 0.1, 1e-5
    a = 0.8 * KN * dt * u
    stored = []
    for n in (1, 2):
        _, (out, rows) = _apex_under_face(1.0, u=u, n=n, mutation=mutation)
        stored.append(rows[6][rows[2] == 12][0])
    checks = [out[0][0][0] == pytest.approx(-a, rel=1e-12), stored[0] == pytest.approx([-a, 0.0, 0.0], abs=1e-15),
              abs(out[1][0][0]) < 1e-15, np.all(stored[1] == 0.0)]
    if mutation == 0:
        assert all(checks)
    else:
        assert not all(checks)


@pytest.mark.parametrize("mutation", [0, 10])
def test_mesh_vertex_gate_ignores_homothety(mutation):
    """At homothety 2 the apex is 0.38 below the unscaled, rVerlet-enlarged shape box, so the reference
    lists no apex candidate although the apex presses into the scaled cube's face; with the gate dropped
    (planted mistake 10) the candidate and its force appear."""
    _, ([(F, _)], rows) = _apex_under_face(2.0, mutation=mutation)
    n12 = np.count_nonzero(rows[2] == 12)
    if mutation == 0:
        assert n12 == 0 and np.all(F == 0.0)
    else:
        assert n12 == 1 and F[2] == pytest.approx(KN * 1e-3, rel=1e-9)


# ------------------------------------------------------------------------------ brute force
def _seg_point(a, b, p):
    """Parameter in [0, 1] of the point of segment ab closest to p (bounded least squares)."""
    return float(lsq_linear((b - a)[:, None], p - a, bounds=(0.0, 1.0), method="bvls").x[0])


def _seg_seg(p0, p1, a, b):
    r = lsq_linear(np.stack([p1 - p0, -(b - a)], 1), a - p0, bounds=([0.0, 0.0], [1.0, 1.0]), method="bvls")
    return float(r.x[0]), float(r.x[1])


def _in_polygon(p, poly):
    """Projection of p into the plane of the convex polygon, and whether it falls strictly inside
    (least-squares coordinates on two edge vectors, then edge half-planes)."""
    o = poly[0]
    B = np.stack([poly[1] - o, poly[-1] - o], 1)
    coef = np.linalg.lstsq(B, p - o, rcond=None)[0]
    proj = o + B @ coef
    nrm = np.cross(poly[1] - o, poly[-1] - o)
    inside = all(np.dot(np.cross(poly[(k + 1) % len(poly)] - poly[k], proj - poly[k]), nrm) > 1e-12
                 for k in range(len(poly)))
    return proj, inside


def brute_force(shape, c, q, rm, mesh):
    """Every (type, particle feature, mesh feature) in contact with the particle point, mesh point and dn,
    from closest points of the features themselves."""
    R = rotmat(q)
    ra = shape.radius
    V = c + np.array(shape.vertices) @ R.T
    mv, me, mf = mesh["vertices"], mesh["edges"], mesh["faces"]
    out = []
    near_v = [k for k in range(len(mv)) if np.linalg.norm(mv[k] - c) < 2.0]
    near_e = [k for k in range(len(me)) if min(np.linalg.norm(mv[me[k][0]] - c), np.linalg.norm(mv[me[k][1]] - c)) < 3.0]
    near_f = [k for k in range(len(mf)) if min(np.linalg.norm(mv[i] - c) for i in mf[k]) < 3.0]

    def add(t, i, k, xp, xm):
        dn = np.linalg.norm(xp - xm) - (ra + rm)
        if dn < 0:
            out.append((t, i, k, xp, xm, dn))

    inner = lambda s: 1e-7 < s < 1 - 1e-7
    for i, p in enumerate(V):
        for k in near_v:
            add(7, i, k, p, mv[k])
        for k in near_e:
            a, b = mv[me[k][0]], mv[me[k][1]]
            t = _seg_point(a, b, p)
            if inner(t):
                add(8, i, k, p, a + t * (b - a))
        for k in near_f:
            proj, ins = _in_polygon(p, mv[mf[k]])
            if ins:
                add(9, i, k, p, proj)
    for i, (e0, e1) in enumerate(shape.edges):
        for k in near_e:
            a, b = mv[me[k][0]], mv[me[k][1]]
            s, t = _seg_seg(V[e0], V[e1], a, b)
            if inner(s) and inner(t):
                add(10, i, k, V[e0] + s * (V[e1] - V[e0]), a + t * (b - a))
    for k in near_v:
        m = mv[k]
        for i, (e0, e1) in enumerate(shape.edges):
            t = _seg_point(V[e0], V[e1], m)
            if inner(t):
                add(11, i, k, V[e0] + t * (V[e1] - V[e0]), m)
        for i, face in enumerate(shape.faces):
            proj, ins = _in_polygon(m, V[list(face)])
            if ins:
                add(12, i, k, proj, m)
    return out


@pytest.mark.parametrize("seed", range(10))
def test_random_poses_against_brute_force(seed):
    """A cube or an octahedron in a random orientation over the bump, lowered 3e-3 below its first touch. Moving slowly, the items with friction are exactly the brute-force
    contacts; at rest, F = sum kn |dn| n and M = sum (cp - c) x f over them, with n from the mesh point to
    the particle point and cp = x_particle - n (r + dn / 2)."""
    rng = np.random.default_rng(seed)
    shape = CUBE if seed % 2 == 0 else shapes.octahedron(0.8, 0.02)
    rm = 0.01
    mesh = mc.tray_mesh(rm)
    if seed < 8:
        qv = rng.normal(size=4)
        q = tuple(qv / np.linalg.norm(qv))
        c = np.array([10.0, 10.0, 0.0]) + np.append(rng.uniform(-1.2, 1.2, 2), 0.0)
    else:  # nearly face down over the apex: the apex against a particle face
        q = quat_axis(rng.normal(size=3), 0.15)
        c = np.array([10.0, 10.0, 0.0]) + np.append(rng.uniform(-0.2, 0.2, 2), 0.0)
    R = rotmat(q)
    low = float(np.min((np.array(shape.vertices) @ R.T)[:, 2]))
    # placement only: the height of first touch, found by bisection on the oracle's force (the checks
    # below are made at the final pose against the brute force)
    touches = lambda z: np.any(passes(one(shape, (c[0], c[1], z), q), shape, rm)[0][0][0] != 0)
    hi, lo = 1.2 - low, 1.2 - low
    while not touches(lo):
        lo -= 0.02
    for _ in range(30):
        mid = 0.5 * (hi + lo)
        hi, lo = (hi, mid) if touches(mid) else (mid, lo)
    c[2] = hi - 3e-3
    contacts = brute_force(shape, c, q, rm, mesh)
    kinds = {t for t, *_ in contacts}
    # the contact set: the items whose friction is non-zero after one pass while sliding slowly
    vel = rng.normal(size=3) * 1e-3
    _, rows = passes(one(shape, c, q, v=vel), shape, rm)
    _, jj, tt, _, si, sj, fr = rows
    got = sorted((int(t), int(i), int(k)) for t, i, k, f in zip(tt, si, sj, fr) if t >= 7 and np.any(f != 0))
    want = sorted((t, i, k) for t, i, k, *_ in contacts)
    assert got == want, (kinds, got, want)
    [(F, M)], _ = passes(one(shape, c, q), shape, rm)
    ra = shape.radius
    Fb, Mb = np.zeros(3), np.zeros(3)
    for t, i, k, xp, xm, dn in contacts:
        n = (xp - xm) / np.linalg.norm(xp - xm)
        f = -KN * dn * n
        Fb += f
        Mb += np.cross(xp - n * (ra + 0.5 * dn) - c, f)
    scale = np.sum([-KN * dn for *_, dn in contacts])
    assert F == pytest.approx(Fb, abs=1e-8 * scale)
    assert M == pytest.approx(Mb, abs=1e-8 * scale)
