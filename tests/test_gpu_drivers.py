"""Parity of the cylinder and ball drivers (register_cylinder, register_ball; cylinder.hpp:225-272,
ball.hpp:292-327, contact_sphere.hpp:216-297) between the HIP path and the CPU oracle.

Two contact-rich inputs: a rotating drum (cylinder along y spinning about its axis, closed by two planes)
half filled with spheres under gravity, and a ball container holding spheres around a ball obstacle.
Bars as for the walls: interaction keys and friction history bit-exact, forces and trajectories of the
operator path and the engine equal to the oracle's, the FP32-force form within 1e-5.
"""
import numpy as np
import pytest

from exadem_amd import inputs, native
from exadem_amd.sim import SphereSim, scheme
import pyoracle as po

from helpers import by_id, force, interaction_table, pos, rel_err, vel

pytestmark = pytest.mark.gpu


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


CASES = {"drum": drum, "balls": balls}


def oracle_for(w):
    return po.Oracle(w.particles, w.domain, w.walls)


@pytest.mark.parametrize("case", ["drum", "balls"])
def test_drivers_match_oracle(gpu_ok, case):
    w = CASES[case]()
    assert w.n > 200
    sch = scheme(w)
    o = oracle_for(w)
    o.first_iteration(sch)
    s = SphereSim(w)
    s.first_iteration()
    pg, pq = by_id(s.ctx.download()), by_id(o.particles())
    floor = pq["mass"] * 9.81
    assert rel_err(force(pg), force(pq), floor) <= 1e-12
    gi, gj, gt, gf = s.ctx.interactions()
    oi, oj, ot, of, om = o.interactions()
    kg, fg = interaction_table(gi, gj, gt, gf)
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko) and np.array_equal(fg, fo)
    walls_in_contact = int(np.count_nonzero((ko[:, 2] == 5) & np.any(fo != 0, axis=1)))
    assert walls_in_contact > 10
    # 60 steps: operator path and engine against the oracle
    o.run_steps(sch, w.steps)
    for _ in range(w.steps):
        s.step()
    e = SphereSim(w)
    e.first_iteration()
    e.run(w.steps, poll=7)
    pq = by_id(o.particles())
    for g in (by_id(s.ctx.download()), by_id(e.ctx.download())):
        L = max(1.0, float(np.max(np.abs(pos(pq)))))
        assert float(np.max(np.abs(pos(g) - pos(pq)))) / L <= 1e-10
        assert float(np.max(np.abs(vel(g) - vel(pq)))) / max(1e-3, float(np.max(np.abs(vel(pq))))) <= 1e-10
    kg, fg = interaction_table(*e.ctx.interactions())
    oi, oj, ot, of, om = o.interactions()
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert float(np.max(np.abs(fg - fo))) <= 1e-12 * max(1e-30, float(np.max(np.abs(fo))))
    print(f"{case}: {w.n} spheres, {walls_in_contact} driver contacts at the first pass")


@pytest.mark.parametrize("case", ["drum", "balls"])
def test_drivers_fp32_form(gpu_ok, case):
    w = CASES[case]()
    a = SphereSim(w)
    a.first_iteration()
    b = SphereSim(w, form=native.XDEM_FORM_FP32)
    b.first_iteration()
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    assert rel_err(force(pb), force(pa), pa["mass"] * 9.81) <= 1e-5


def test_driver_errors(gpu_ok):
    """Invalid drivers fail with XDEM_ERR_ARG: an unknown kind, a cylinder or ball without a positive
    radius, a plane without a normal; more than 8 drivers fail with XDEM_ERR_CAPACITY."""
    ctx = native.Context()
    for bad in ({"kind": "cylinder", "radius": 0.0, "axis": (1, 0, 1), "center": (1, 1, 1)},
                {"kind": "ball", "radius": -1.0, "center": (1, 1, 1)},
                {"offset": 1.0, "normal": (0.0, 0.0, 0.0)}):
        with pytest.raises(native.XdemError, match="ARG"):
            ctx.set_surfaces([bad])
    arr = native.surfaces([{"offset": 0.0, "normal": (0, 0, 1)}])
    arr[0].kind = 7
    with pytest.raises(native.XdemError, match="ARG"):
        ctx.call("xdem_set_surfaces", arr, 1)
    with pytest.raises(native.XdemError, match="CAPACITY"):
        ctx.set_surfaces([{"offset": float(k), "normal": (0, 0, 1)} for k in range(9)])


def test_driver_types_in_dump(gpu_ok):
    """Interaction dumps type driver contacts by driver (interaction_enum.hpp:36-38): 4 cylinder,
    5 surface, 6 ball."""
    for case, want, must in (("drum", {4, 5}, 4), ("balls", {6}, 6)):
        w = CASES[case]()
        s = SphereSim(w)
        s.ctx.set_contact_record(True)
        s.first_iteration()
        ii, jj, si, sj, tt, vals = s.ctx.contact_state(subs=True)
        got = set(tt[tt != 0].tolist())
        assert must in got and got <= want, (case, got)
