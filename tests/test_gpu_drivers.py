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
from driver_cases import balls, drum

pytestmark = pytest.mark.gpu


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
