"""Seeded synthetic inputs for the sphere workloads.

Particle fields follow set_fields (src/field/set_fields.cu:323-327): m = rho * 4/3 pi r^3 and an
isotropic inertia 0.4 m r^2. Random draws come from the counter-based SplitMix64 streams of
libxdem_host.so, so a workload is a pure function of its seeds.
"""
import ctypes as C
import math
import os

import numpy as np

from .native import LIBDIR, domain

_host = None


def host():
    global _host
    if _host is None:
        path = os.path.join(LIBDIR, "libxdem_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with exadem_amd/build.py")
        lib = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        lib.xdem_gen_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, dp]
        lib.xdem_gen_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, dp]
        lib.xdem_gen_pack_periodic.argtypes = [C.c_uint64, C.c_int64, dp, dp, dp, C.c_double, C.c_int64,
                                               C.c_double, dp, dp, dp, dp]
        lib.xdem_gen_pack_periodic.restype = C.c_int64
        _host = lib
    return _host


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def uniform(seed, tag, n, a, b):
    out = np.empty(n, np.float64)
    host().xdem_gen_uniform(seed, tag, n, a, b, _p(out))
    return out


def normal(seed, tag, n, mean, sigma):
    out = np.empty(n, np.float64)
    host().xdem_gen_normal(seed, tag, n, mean, sigma, _p(out))
    return out


def sphere_set(rx, ry, rz, radius, density, vx=None, vy=None, vz=None):
    """Full particle field dictionary (input order, ids 0..n-1) for spheres."""
    n = rx.shape[0]
    pi = 4 * math.atan(1)
    r = radius
    V = ((4.0) / (3.0)) * pi * r * r * r
    m = V * density
    inertia = 0.4 * m * r * r
    z = np.zeros(n)
    return {
        "id": np.arange(n, dtype=np.uint64), "type": np.zeros(n, np.uint32), "group": np.zeros(n, np.uint32),
        "rx": np.ascontiguousarray(rx, np.float64), "ry": np.ascontiguousarray(ry, np.float64),
        "rz": np.ascontiguousarray(rz, np.float64),
        "vx": z.copy() if vx is None else np.ascontiguousarray(vx), "vy": z.copy() if vy is None else np.ascontiguousarray(vy),
        "vz": z.copy() if vz is None else np.ascontiguousarray(vz),
        "fx": z.copy(), "fy": z.copy(), "fz": z.copy(), "mass": m, "radius": np.ascontiguousarray(r, np.float64),
        "homothety": np.ones(n), "qw": np.ones(n), "qx": z.copy(), "qy": z.copy(), "qz": z.copy(),
        "momx": z.copy(), "momy": z.copy(), "momz": z.copy(), "vrotx": z.copy(), "vroty": z.copy(),
        "vrotz": z.copy(), "arotx": z.copy(), "aroty": z.copy(), "arotz": z.copy(),
        "inx": inertia.copy(), "iny": inertia.copy(), "inz": inertia.copy(),
    }


class Workload:
    """Particles + domain + walls + contact law + time step of one configuration."""

    def __init__(self, name, particles, dom, walls, cp, cpd, dt, steps, rcut_inc, gravity, meta=None):
        self.name = name
        self.particles = particles
        self.domain = dom
        self.walls = walls
        self.cp = cp
        self.cpd = cpd
        self.dt = dt
        self.steps = steps
        self.rcut_inc = rcut_inc
        self.gravity = gravity
        self.meta = meta or {}

    @property
    def n(self):
        return int(self.particles["rx"].shape[0])


def box_walls(hi):
    """Six stationary Surface walls bounding [0,hi] (register_surface with normal/offset)."""
    walls = []
    for a in range(3):
        nrm = [0.0, 0.0, 0.0]
        nrm[a] = 1.0
        walls.append({"offset": 0.0, "normal": tuple(nrm)})
        walls.append({"offset": float(hi[a]), "normal": tuple(nrm)})
    return walls


def deposition(n=10000, spacing=1.2, jitter=0.05, seed=1, radius=0.5, density=2500.0, box=(40.0, 40.0, 80.0),
               start=1.5, dt=1e-5, steps=1000, rcut_inc=0.1):
    """Plain FP64 sphere deposition: jittered simple-cubic lattice in a walled box under gravity.

    Sites at start + k*spacing on x and y while <= box - start, layers along z; the first n sites in
    x-fastest order are kept. spacing < 2*radius gives a contact-rich variant (initial overlaps)."""
    from .native import contact_params
    nx = int(math.floor((box[0] - 2 * start) / spacing + 1e-9)) + 1
    ny = int(math.floor((box[1] - 2 * start) / spacing + 1e-9)) + 1
    nz = int(math.ceil(n / (nx * ny)))
    k = np.arange(nx * ny * nz)[:n]
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    jit = uniform(seed, 1, 3 * n, -jitter, jitter).reshape(3, n)
    rx = start + ix * spacing + jit[0]
    ry = start + iy * spacing + jit[1]
    rz = start + iz * spacing + jit[2]
    p = sphere_set(rx, ry, rz, np.full(n, radius), density)
    dom = domain((0.0, 0.0, 0.0), box, (False, False, False), 2 * radius + rcut_inc)
    cp = contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    return Workload("deposition", p, dom, box_walls(box), cp, cp, dt, steps, rcut_inc, (0.0, 0.0, -9.81),
                    {"spacing": spacing, "jitter": jitter, "seed": seed})


def polydisperse(n=1_000_000, phi=0.55, rmin=0.5, rmax=1.0, seed_radius=2, seed_vel=3, seed_pos=2, sigma_v=0.1,
                 box_xy=200.0, density=2500.0, dt=5e-6, steps=2000, rcut_inc=0.1, gravity=(0.0, 0.0, -9.81),
                 relax_sweeps=150, tol=None):
    """Polydisperse spheres in a periodic box at solid fraction phi.

    Radii U[rmin, rmax] (seed_radius). At n = 1e6 the box is 200 x 200 x Lz with Lz fixed by phi;
    smaller n scale all three edges by (n/1e6)^(1/3). Positions: RSA at radius scale 0.8, then
    full radii and Jacobi overlap relaxation (max_overlap reported in meta). Velocities N(0, sigma_v)."""
    from .native import contact_params
    r = uniform(seed_radius, 1, n, rmin, rmax)
    vol = float(np.sum((4.0 / 3.0) * math.pi * r ** 3))
    s = (n / 1e6) ** (1.0 / 3.0)
    bx = by = box_xy * s
    bz = vol / (phi * bx * by)
    lo = np.zeros(3)
    hi = np.array([bx, by, bz])
    x = np.empty(n)
    y = np.empty(n)
    z = np.empty(n)
    mo = C.c_double()
    tol = 1e-3 * rmin if tol is None else tol
    sweeps = host().xdem_gen_pack_periodic(seed_pos, n, _p(r), _p(lo), _p(hi), 0.8, relax_sweeps, tol, _p(x), _p(y),
                                           _p(z), C.byref(mo))
    v = normal(seed_vel, 2, 3 * n, 0.0, sigma_v).reshape(3, n)
    p = sphere_set(x, y, z, r, density, v[0], v[1], v[2])
    dom = domain((0.0, 0.0, 0.0), tuple(hi), (True, True, True), 2 * rmax + rcut_inc)
    cp = contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    return Workload("polydisperse", p, dom, [], cp, cp, dt, steps, rcut_inc, gravity,
                    {"phi": phi, "box": tuple(hi), "relax_sweeps": int(sweeps), "max_overlap": mo.value})


def precision_spheres(n=4_000_000, phi=0.58, median=0.6, sigma=0.2, rmin=0.3, rmax=1.2, seed_radius=5, seed_vel=6,
                      seed_pos=5, sigma_v=0.1, density=2500.0, dt=5e-6, steps=500, rcut_inc=0.1,
                      gravity=(0.0, 0.0, -9.81), relax_sweeps=40):
    """Input of the precision comparison (FP32-force against FP64): n spheres with lognormal radii
    R = median exp(sigma Z), Z ~ N(0, 1) (seed_radius), clipped to [rmin, rmax], in a periodic cube
    whose edge gives solid fraction phi. Positions: RSA at radius scale 0.8, then full radii and
    relax_sweeps Jacobi overlap-relaxation sweeps (at phi 0.58 the relaxation leaves overlaps of a few
    percent of a radius; max_overlap in meta). Velocities N(0, sigma_v), the contact law of the 1M
    benchmark."""
    from .native import contact_params
    z = normal(seed_radius, 1, n, 0.0, 1.0)
    r = np.clip(median * np.exp(sigma * z), rmin, rmax)
    vol = float(np.sum((4.0 / 3.0) * math.pi * r ** 3))
    edge = (vol / phi) ** (1.0 / 3.0)
    lo = np.zeros(3)
    hi = np.array([edge, edge, edge])
    x = np.empty(n)
    y = np.empty(n)
    zz = np.empty(n)
    mo = C.c_double()
    sweeps = host().xdem_gen_pack_periodic(seed_pos, n, _p(r), _p(lo), _p(hi), 0.8, relax_sweeps, 1e-3 * rmin, _p(x),
                                           _p(y), _p(zz), C.byref(mo))
    v = normal(seed_vel, 2, 3 * n, 0.0, sigma_v).reshape(3, n)
    p = sphere_set(x, y, zz, r, density, v[0], v[1], v[2])
    dom = domain((0.0, 0.0, 0.0), tuple(hi), (True, True, True), 2 * rmax + rcut_inc)
    cp = contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    return Workload("precision", p, dom, [], cp, cp, dt, steps, rcut_inc, gravity,
                    {"phi": phi, "box": tuple(hi), "relax_sweeps": int(sweeps), "max_overlap": mo.value})


# ---------------------------------------------------------------------------------------------- polyhedra
def random_quaternions(seed, n):
    """Uniform random unit quaternions (Shoemake), (w, x, y, z)."""
    u = uniform(seed, 4, 3 * n, 0.0, 1.0).reshape(3, n)
    u1, u2, u3 = u
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    return np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3)])


def polyhedra(n=200_000, spacing=1.7, jitter=0.05, seed=4, density=1.0, minkowski=0.02, dt=2e-6, steps=1000,
              rcut_inc=0.1, nx=None, gravity=(0.0, 0.0, -9.81), sigma_v=0.0, z0=None, periodic=(True, True, False)):
    """Cubes (edge 1.0) and octahedra (circumradius 0.8), 50/50 by id parity, random orientations, on a
    jittered simple-cubic lattice; periodic in x and y, floor wall at z = 0 and lid above.

    Density 1 is the reference's default for polyhedra (set_fields.cu:84-85, "If not defined, density is
    1"): with kn = 1e5 the weight of a particle then presses a contact by about 1e-4, well inside the
    Minkowski shells (0.02 each), where vertex-face detection holds."""
    from .native import contact_params
    from . import shapes as shp
    defs = [shp.cube(1.0, minkowski), shp.octahedron(0.8, minkowski)]
    shapes = [d.to_record() for d in defs]
    rb = [shp.bounding_radius(d) for d in defs]
    if nx is None:
        nx = max(3, int(round((n / 2.0) ** (1.0 / 3.0))))
    ny = nx
    nz = int(math.ceil(n / (nx * ny)))
    k = np.arange(n)
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    jit = uniform(seed, 1, 3 * n, -jitter, jitter).reshape(3, n)
    L = nx * spacing
    rx = (ix + 0.5) * spacing + jit[0]
    ry = (iy + 0.5) * spacing + jit[1]
    rz = (max(rb) + 0.05 if z0 is None else z0) + iz * spacing + jit[2]
    types = (k % 2).astype(np.uint32)
    p = sphere_set(rx, ry, rz, np.array([rb[t] for t in types]), density)
    vol = np.array([shapes[t].volume for t in types])
    im = np.array([shapes[t].inertia_on_mass[0] for t in types])
    p["mass"] = density * vol
    for a in ("inx", "iny", "inz"):
        p[a] = p["mass"] * im
    q = random_quaternions(seed, n)
    p["qw"], p["qx"], p["qy"], p["qz"] = (np.ascontiguousarray(q[i]) for i in range(4))
    p["type"] = types
    if sigma_v > 0:
        v = normal(seed + 1, 2, 3 * n, 0.0, sigma_v).reshape(3, n)
        p["vx"], p["vy"], p["vz"] = (np.ascontiguousarray(v[i]) for i in range(3))
    top = float(rz.max() + max(rb) + 2.0)
    cell = 2 * max(rb) + rcut_inc
    dom = domain((0.0, 0.0, 0.0), (L, L, top), tuple(periodic), cell)
    walls = [{"offset": 0.0, "normal": (0.0, 0.0, 1.0)}, {"offset": top, "normal": (0.0, 0.0, 1.0)}]
    cp = contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    w = Workload("polyhedra", p, dom, walls, cp, cp, dt, steps, rcut_inc, gravity,
                 {"spacing": spacing, "box": (L, L, top), "nx": nx})
    w.shapes = shapes
    return w


def _start_with_accelerations(sim):
    """Rebuild and one force pass ending with force_to_accel, so that the next step kicks with accelerations
    (the first_iteration of the reference decks leaves forces in place of accelerations, which at the
    overlaps of a growing packing would kick the particles by F dt / 2 instead of F dt / 2m)."""
    sim.rebuild()
    sim.ctx.reset_force_moment()
    sim.compute_force()
    sim.ctx.force_to_accel()


def polyhedra_dense(n=200_000, phi=0.45, seed=4, dt=2e-6, growth=1.015, relax_steps=200, final_relax=3,
                    settle_steps=0, dt_grow=5e-4, h0=0.8, log=None, cache=None):
    """Polyhedra at solid fraction phi of the box, grown with the GPU DEM.

    Starts from the jittered lattice of `polyhedra` (spacing 1.55, periodic in x and y, floor and lid)
    with every particle scaled by h0 (0.8: bounding spheres 1.42 across, so that no two particles
    overlap whatever their orientations) and grows every particle's homothety by `growth` per stage (mass x h^3,
    inertia x h^5, bounding radius x h, cells re-sized to 2 max radius + rcut_inc) until the particles'
    volume fills phi of the box; each stage starts at rest and relaxes for relax_steps DEM steps of
    dt_grow without gravity, and final_relax more such stages follow at the final size. The result stands
    for the configuration's random packing at phi, with gravity acting from the run's first step;
    settle_steps > 0 lets it settle onto the floor under gravity first. Velocities are zeroed; dt is set
    to `dt`. meta gives the achieved solid fraction of the box, that of the bed (particle volume over the
    box area times the bed's top), the homothety and the smallest overlap (dn) of the last pass.

    cache: an .npz path; when it holds the result for the same arguments it is loaded instead of grown,
    otherwise the grown particles are written there (so that a profiled run can skip the growth)."""
    import copy as _copy
    import json as _json
    import os as _os
    from .sim import PolySim
    w = polyhedra(n=n, spacing=1.55, jitter=0.05, seed=seed, dt=dt)
    key = _json.dumps([n, phi, seed, dt, growth, relax_steps, final_relax, settle_steps, dt_grow, h0,
                       float(np.sum(w.particles["mass"]))])
    if cache and _os.path.exists(cache):
        with np.load(cache) as z:
            meta = _json.loads(str(z["meta_json"]))
            if str(z["key"]) == key:
                out = _copy.copy(w)
                out.particles = {k[2:]: np.array(z[k]) for k in z.files if k.startswith("p_")}
                L, _, top = w.meta["box"]
                out.domain = domain((0.0, 0.0, 0.0), (L, L, top), (True, True, False), float(z["cell"]))
                out.meta = meta
                return out
    L, _, top = w.meta["box"]
    vbox = L * L * top
    p = {k: np.array(v, copy=True) for k, v in w.particles.items()}
    vol_of = lambda q: np.array([w.shapes[t].volume for t in q["type"]]) * q["homothety"] ** 3
    phi0 = float(np.sum(vol_of(p))) / vbox
    h_target = (phi / phi0) ** (1.0 / 3.0)
    rb_max = float(np.max(p["radius"]))
    p["homothety"] = p["homothety"] * h0
    p["radius"] = p["radius"] * h0
    p["mass"] = p["mass"] * h0 ** 3
    for a in ("inx", "iny", "inz"):
        p[a] = p[a] * h0 ** 5
    g = _copy.copy(w)
    g.gravity = (0.0, 0.0, 0.0)
    g.dt = dt_grow
    sim = PolySim(g)
    h, stages = h0, 0
    while h < h_target:
        h_new = min(h * growth, h_target)
        f = h_new / h
        p["homothety"] = p["homothety"] * f
        p["radius"] = p["radius"] * f
        p["mass"] = p["mass"] * f ** 3
        for a in ("inx", "iny", "inz"):
            p[a] = p[a] * f ** 5
        g.domain = domain((0.0, 0.0, 0.0), (L, L, top), (True, True, False), 2 * rb_max * h_new + w.rcut_inc)
        for k in ("vx", "vy", "vz", "vrotx", "vroty", "vrotz"):  # each stage starts at rest
            p[k] = np.zeros_like(p[k])
        sim.w = g
        sim.ctx.set_domain(g.domain)
        sim.ctx.upload(p)
        _start_with_accelerations(sim)
        st = sim.run(relax_steps, poll=64)
        p = sim.ctx.download()
        h = h_new
        stages += 1
        if log:
            vmax = float(np.max(np.abs(np.stack([p["vx"], p["vy"], p["vz"]]))))
            log(f"polyhedra_dense: stage {stages}, homothety {h:.4f}, rebuilds {st['rebuilds']}, max |v| {vmax:.3g}")
    for r in range(final_relax):
        for k in ("vx", "vy", "vz", "vrotx", "vroty", "vrotz"):
            p[k] = np.zeros_like(p[k])
        sim.ctx.upload(p)
        _start_with_accelerations(sim)
        st = sim.run(relax_steps, poll=64)
        p = sim.ctx.download()
        if log:
            vmax = float(np.max(np.abs(np.stack([p["vx"], p["vy"], p["vz"]]))))
            log(f"polyhedra_dense: relaxation {r + 1}, rebuilds {st['rebuilds']}, max |v| {vmax:.3g}")
    if settle_steps > 0:
        g.gravity = w.gravity
        for k in ("vx", "vy", "vz", "vrotx", "vroty", "vrotz"):
            p[k] = np.zeros_like(p[k])
        sim.w = g
        sim.ctx.upload(p)
        _start_with_accelerations(sim)
        rs = sim.run(settle_steps, poll=64)
        if log:
            log(f"polyhedra_dense: settled, {rs['rebuilds']} rebuilds")
        p = sim.ctx.download()
    st = sim.ctx.simulation_state()
    if log:
        log(f"polyhedra_dense: {st['n_interactions']} interactions, {st['n_active']} active, min dn {st['min_dn']:.3g}")
    del sim
    for k in ("vx", "vy", "vz", "vrotx", "vroty", "vrotz", "fx", "fy", "fz", "momx", "momy", "momz", "arotx",
              "aroty", "arotz"):
        p[k] = np.zeros_like(p[k])
    out = _copy.copy(w)
    out.particles = p
    out.domain = g.domain
    vp = float(np.sum(vol_of(p)))
    bed = float(np.max(p["rz"] + p["radius"]))
    achieved = vp / vbox
    out.meta = dict(w.meta, solid_fraction=achieved, bed_solid_fraction=vp / (L * L * bed), bed_top=bed,
                    homothety=h, growth_stages=stages, min_dn=st["min_dn"], interactions=st["n_interactions"],
                    active_interactions=st["n_active"],
                    workload=f"cubes + octahedra 50/50, density 1, random orientations, grown from a jittered "
                             f"lattice to solid fraction {achieved:.3f} of the box (homothety {h:.4f})"
                             + (f" and settled under gravity (bed solid fraction {vp / (L * L * bed):.3f})"
                                if settle_steps > 0 else ", relaxed at rest, gravity from the first step")
                             + f", periodic x/y, floor and lid, dt {dt:g}")
    if cache:
        np.savez(cache, key=np.array(key), meta_json=np.array(_json.dumps(out.meta)),
                 cell=np.array(out.domain.cell_size), **{"p_" + k: v for k, v in p.items()})
    return out


def polyhedra_benchmark(n=200_000, log=None, cache=None):
    """The polyhedra benchmark input: n cubes and octahedra 50/50 (edge 1.0, circumradius 0.8, Minkowski
    radius 0.02 before growth, density 1), uniform random orientations (seed 4), grown from a jittered
    lattice of spacing 1.55 to solid fraction 0.45 of the box (polyhedra_dense), periodic in x and y with a
    floor and a lid, gravity -9.81 z, dt 2e-6."""
    return polyhedra_dense(n=n, phi=0.45, seed=4, dt=2e-6, log=log, cache=cache)
