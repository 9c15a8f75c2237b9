"""Triangle-mesh driver inputs shared by the CPU and GPU mesh tests: a heightfield tray (grid spacing 2 over
[0, 20]^2 at z = 0, each cell split along its (i, j)-(i+1, j+1) diagonal) whose centre vertex is raised to
z = 1, so the mesh has flat faces, a convex apex vertex and convex ridge edges."""
import numpy as np

from exadem_amd import inputs, native, shapes, stl


def tray_triangles(n=10, h=2.0, apex=1.0):
    z = np.zeros((n + 1, n + 1))
    z[n // 2, n // 2] = apex
    P = lambda i, j: (i * h, j * h, z[i, j])
    tri = []
    for i in range(n):
        for j in range(n):
            tri.append([P(i, j), P(i + 1, j), P(i + 1, j + 1)])
            tri.append([P(i, j), P(i + 1, j + 1), P(i, j + 1)])
    return np.array(tri, np.float64)


def tray_mesh(rm=0.0, mesh_id=6):
    v, e, f = stl.build_shape(tray_triangles())
    return {"mesh_id": mesh_id, "vertices": v, "edges": e, "faces": f, "minkowski": rm}


def face_normal(t):
    n = np.cross(t[1] - t[0], t[2] - t[0])
    n = n / np.linalg.norm(n)
    return n if n[2] > 0 else -n


def probes(R=0.5, overlap=0.01, rm=0.0):
    """Sphere centres touching the apex vertex, a ridge edge and a flat face with the given overlap."""
    tri = tray_triangles()
    apex = np.array([10.0, 10.0, 1.0])
    d = R + rm - overlap
    # ridge edge apex -> (12, 10, 0): bisector of the normals of its two faces
    fa = [t for t in tri if any(np.allclose(p, apex) for p in t) and any(np.allclose(p, (12, 10, 0)) for p in t)]
    nb = sum(face_normal(t) for t in fa)
    nb = nb / np.linalg.norm(nb)
    mid = (apex + np.array([12.0, 10.0, 0.0])) / 2
    return {"vertex": apex + np.array([0, 0, d]), "edge": mid + nb * d, "face": np.array([3.3, 4.6, d])}


def workload(name, p, box=(20.0, 20.0, 7.0), gravity=(0.0, 0.0, -9.81), steps=60, dt=1e-4):
    cp = native.contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    dom = native.domain((0.0, 0.0, -1.0), box, (False, False, False), 1.1)
    return inputs.Workload(name, p, dom, [], cp, cp, dt, steps, 0.1, gravity)


def single(center, R=0.5, v=(0.0, 0.0, 0.0)):
    return inputs.sphere_set(np.array([center[0]]), np.array([center[1]]), np.array([center[2]]), np.full(1, R),
                             2500.0, np.array([v[0]]), np.array([v[1]]), np.array([v[2]]))


def tray_bed(seed=21, R=0.5):
    """Two jittered layers (spacing 0.98, so neighbours overlap) over the flat part of the tray, the
    probes on the bump, velocities N(0, 0.3)."""
    g = np.arange(1.0, 19.0, 0.98)
    x, y = np.meshgrid(g, g, indexing="ij")
    x, y = x.ravel(), y.ravel()
    far = np.hypot(x - 10.0, y - 10.0) > 3.2
    x, y = x[far], y[far]
    xs = np.concatenate([x, x + 0.3])
    ys = np.concatenate([y, y + 0.2])
    zs = np.concatenate([np.full(x.size, R - 0.01), np.full(x.size, R - 0.01 + 0.98)])
    pr = probes(R)
    for k in ("vertex", "edge"):
        xs, ys, zs = np.append(xs, pr[k][0]), np.append(ys, pr[k][1]), np.append(zs, pr[k][2])
    n = xs.size
    jit = inputs.uniform(seed, 1, 2 * n, -0.02, 0.02).reshape(2, n)
    jit[:, -2:] = 0.0
    v = inputs.normal(seed + 1, 2, 3 * n, 0.0, 0.3).reshape(3, n)
    return inputs.sphere_set(xs + jit[0], ys + jit[1], zs, np.full(n, R), 2500.0, v[0], v[1], v[2])


def tray_polyhedra(seed=3, vz=-3.0):
    """Cubes (edge 1) and octahedra (circumradius 0.8), Minkowski radius 0.02, density 1, alternating, in
    random orientations on a 7 x 7 grid of spacing 2.2 over the tray (x, y in 3.4 .. 16.6), two layers:
    the lower one 0.03 to 0.1 above the tray under it (the bump included), the upper one 1.8 higher;
    falling at vz plus N(0, 0.3) per component, spins N(0, 1). Gravity, no walls, dt 1e-4 (the contact
    time is about 0.01 s at kn 1e5 for these masses)."""
    defs = [shapes.cube(1.0, 0.02), shapes.octahedron(0.8, 0.02)]
    recs = [d.to_record() for d in defs]
    g = 3.4 + 2.2 * np.arange(7)
    x, y = np.meshgrid(g, g, indexing="ij")
    x, y = x.ravel(), y.ravel()
    # highest tray point under a footprint of radius 0.9: the bump falls by 1/2 per unit of max(|dx|, |dy|)
    top = np.clip(1.0 - (np.maximum(np.abs(x - 10.0), np.abs(y - 10.0)) - 0.9) / 2.0, 0.0, 1.0)
    x, y, z = np.concatenate([x, x]), np.concatenate([y, y]), np.concatenate([top + 0.9, top + 2.7])
    n = x.size
    t = (np.arange(n) % 2).astype(np.uint32)
    v = inputs.normal(seed, 2, 3 * n, 0.0, 0.3).reshape(3, n)
    rb = np.array([shapes.bounding_radius(defs[k]) for k in t])
    p = inputs.sphere_set(x, y, z, rb, 1.0, v[0], v[1], v[2] + vz)
    p["type"] = t
    p["mass"] = np.array([recs[k].volume for k in t])
    for a in ("inx", "iny", "inz"):
        p[a] = p["mass"] * np.array([recs[k].inertia_on_mass[0] for k in t])
    q = inputs.normal(seed + 1, 2, 4 * n, 0.0, 1.0).reshape(4, n)
    q = q / np.linalg.norm(q, axis=0)
    for k, key in enumerate(("qw", "qx", "qy", "qz")):
        p[key] = np.ascontiguousarray(q[k])
    w = inputs.normal(seed + 2, 2, 3 * n, 0.0, 1.0).reshape(3, n)
    for k, key in enumerate(("vrotx", "vroty", "vrotz")):
        p[key] = np.ascontiguousarray(w[k])
    wl = workload("tray polyhedra", p, dt=1e-4, steps=300)
    wl.domain = native.domain((0.0, 0.0, -1.0), (20.0, 20.0, 7.0), (False, False, False), 2 * float(rb.max()) + 0.1)
    wl.shapes = recs
    return wl
