"""Moving-driver inputs shared by the CPU and GPU tests (dt 1e-4, rcut_inc 0.1, so a driver that moves by
0.05 since the last neighbour build triggers a rebuild):

- piston: spheres touching the six walls of a 7.89 x 7.89 x 12 box; the +x wall moves inward
  (LINEAR_MOTION, motion_vector [1, 0, 0], const_vel -8: 8e-4 per step, a rebuild every ~63 steps);
- shaker: the same box with its floor shaking (SHAKER, amplitude 0.01, omega 300, along z);
- ball: spheres between a ball container and a ball obstacle that moves along x (LINEAR_MOTION,
  const_vel 4); start: the piston with motion_start_threshold 5e-3 (the wall starts at step 49);
- drum: a half-filled drum (cylinder along y, spinning) shaking along x (SHAKER, amplitude 0.02, omega 100).
"""
import copy

import numpy as np

from exadem_amd import inputs, native

from helpers import wall_deposition


def lattice_inside(keep, box, spacing=0.98, seed=9, r=0.5):
    """Jittered lattice sites of the box kept by keep(x, y, z), as a sphere set with ids 0..n-1 and
    velocities N(0, 0.3)."""
    g = [np.arange(0.6, box[a] - 0.6, spacing) for a in range(3)]
    x, y, z = np.meshgrid(*g, indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    jit = inputs.uniform(seed, 1, 3 * x.size, -0.03, 0.03).reshape(3, -1)
    x, y, z = x + jit[0], y + jit[1], z + jit[2]
    sel = keep(x, y, z)
    n = int(sel.sum())
    v = inputs.normal(seed + 1, 2, 3 * n, 0.0, 0.3).reshape(3, n)
    return inputs.sphere_set(x[sel], y[sel], z[sel], np.full(n, r), 2500.0, v[0], v[1], v[2])


def settle_onto(p, center, mask, target, band):
    """Moves the spheres whose distance from center (over the coordinates in mask) lies within band of
    target onto target, radially: an overlap of 0.01 with the driver surface there."""
    d = np.stack([(p["rx"] - center[0]) * mask[0], (p["ry"] - center[1]) * mask[1], (p["rz"] - center[2]) * mask[2]])
    r = np.sqrt(np.sum(d * d, axis=0))
    sel = np.abs(r - target) < band
    f = np.where(sel, target / np.maximum(r, 1e-300), 1.0)
    for a, k in enumerate(("rx", "ry", "rz")):
        if mask[a]:
            p[k] = center[a] + d[a] * f
    return int(sel.sum())


def drum():
    box = (14.0, 5.0, 14.0)
    c, R = np.array([7.0, 0.0, 7.0]), 6.5
    # lower half of the drum; the outer layer settled onto the cylinder
    p = lattice_inside(lambda x, y, z: (np.hypot(x - c[0], z - c[2]) < R - 0.49) & (z < c[2] + 0.5), box)
    settle_onto(p, c, (1, 0, 1), R - 0.49, 0.45)
    walls = [{"kind": "cylinder", "radius": R, "axis": (1.0, 0.0, 1.0), "center": tuple(c), "vrot": (0.0, 1.5, 0.0)},
             {"offset": 0.0, "normal": (0.0, 1.0, 0.0)}, {"offset": box[1], "normal": (0.0, 1.0, 0.0)}]
    return workload("drum", p, box, walls)


def balls():
    box = (14.0, 14.0, 14.0)
    c = np.array([7.0, 7.0, 7.0])
    rr = lambda x, y, z: np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    p = lattice_inside(lambda x, y, z: (rr(x, y, z) < 6.0 - 0.49) & (rr(x, y, z) > 2.0 + 0.49), box, seed=10)
    settle_onto(p, c, (1, 1, 1), 6.0 - 0.49, 0.45)
    settle_onto(p, c, (1, 1, 1), 2.0 + 0.49, 0.45)
    walls = [{"kind": "ball", "radius": 6.0, "center": tuple(c)},
             {"kind": "ball", "radius": 2.0, "center": tuple(c), "vrot": (0.0, 0.0, 1.0)}]
    return workload("balls", p, box, walls)


def workload(name, p, box, walls):
    cp = native.contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    dom = native.domain((0.0, 0.0, 0.0), box, (False, False, False), 1.1)
    return inputs.Workload(name, p, dom, walls, cp, cp, 1e-4, 60, 0.1, (0.0, 0.0, -9.81))


DT = 1e-4
STEPS = 150


def _moving(w, k, motion):
    w = copy.copy(w)
    w.walls = [dict(x) for x in w.walls]
    w.walls[k]["motion"] = motion
    w.dt = DT
    return w


def piston(start=0.0):
    w = wall_deposition()
    m = {"motion_type": "LINEAR_MOTION", "motion_vector": (1.0, 0.0, 0.0), "const_vel": -8.0}
    if start:
        m["motion_start_threshold"] = start
    return _moving(w, 1, m)


def shaker():
    w = wall_deposition()
    return _moving(w, 4, {"motion_type": "SHAKER", "amplitude": 0.01, "omega": 300.0, "shaker_dir": (0.0, 0.0, 1.0)})


def moving_ball():
    return _moving(balls(), 1, {"motion_type": "LINEAR_MOTION", "motion_vector": (1.0, 0.0, 0.0), "const_vel": 4.0})


def shaking_drum():
    return _moving(drum(), 0, {"motion_type": "SHAKER", "amplitude": 0.02, "omega": 100.0,
                               "shaker_dir": (1.0, 0.0, 0.0)})


CASES = {"piston": piston, "shaker": shaker, "ball": moving_ball, "drum": shaking_drum,
         "start": lambda: piston(start=5e-3)}
