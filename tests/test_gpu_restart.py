"""Restart with interaction history, HIP path against the CPU oracle.

A run is checkpointed (every particle field and the interaction rows with their friction and moment
history, the layout of xdem_interactions_dump / xdem_poly_interactions_dump / xdem_mesh_interactions).
Both the HIP path (particles uploaded, lists rebuilt, xdem_interactions_load) and the oracle (the saved
rows restored as their owners' interaction storage, then its rebuild, whose interaction update hands the
active histories to the new interactions by key, interaction_manager.hpp:88-112,205-217) restart from
that file. Bars: the interactions with their history bit-exact right after the restart, and positions,
velocities, forces and histories bit-exact after 40 more steps. Cases: spheres in a walled box, a
periodic sphere packing, polyhedra with floor and walls, spheres on a triangle-mesh tray.
"""
import numpy as np
import pytest

from exadem_amd import inputs
from exadem_amd.sim import PolySim, SphereSim, scheme
import pyoracle as po

import mesh_cases as mc
from helpers import by_id, interaction_table, periodic_pack, wall_deposition

pytestmark = pytest.mark.gpu

STEPS = 40


def _particles(d):
    return {k[2:]: d[k] for k in d.files if k.startswith("p_")}


def _same_state(g, o, what):
    g, o = by_id(g), by_id(o)
    assert np.array_equal(g["id"], o["id"]), what
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "vrotx", "vroty", "vrotz", "fx", "fy", "fz", "momx", "qw", "qx"):
        assert np.array_equal(g[k], o[k]), (what, k)


def _pairs_equal(s, o, what):
    kg, fg = interaction_table(*s.ctx.interactions())
    oi, oj, ot, of, _ = o.interactions()
    keep = ot < 7
    ko, fo = interaction_table(oi[keep], oj[keep], ot[keep], of[keep])
    assert np.array_equal(kg, ko), what
    assert np.array_equal(fg, fo), what
    return int(np.count_nonzero(np.any(fg != 0, axis=1)))


@pytest.mark.parametrize("case", ["walls", "periodic"])
def test_sphere_restart_matches_oracle(gpu_ok, case, tmp_path):
    w = {"walls": wall_deposition, "periodic": periodic_pack}[case]()
    sch = scheme(w)
    a = SphereSim(w)
    a.first_iteration()
    a.run(60)
    ck = str(tmp_path / "restart.npz")
    a.checkpoint(ck)
    d = np.load(ck)
    b = SphereSim.restart(w, ck)
    o = po.Oracle(_particles(d), w.domain, w.walls)
    assert o.restore_interactions(d["ii"], d["jj"], d["tt"], d["fr"], d["mo"]) == len(d["ii"])
    o.rebuild(w.rcut_inc)
    act = _pairs_equal(b, o, f"{case}: after the restart")
    assert act > 50
    b.run(STEPS)
    o.run_steps(sch, STEPS)
    _pairs_equal(b, o, f"{case}: {STEPS} steps after the restart")
    _same_state(b.ctx.download(), o.particles(), case)
    print(f"{case}: {len(d['ii'])} saved interactions, {act} with history, restart equals the oracle's", flush=True)


def test_poly_restart_matches_oracle(gpu_ok, tmp_path):
    w = inputs.polyhedra(density=2500.0, n=1500, spacing=1.5, jitter=0.05, seed=5, dt=2e-4, sigma_v=0.5, z0=0.6,
                         periodic=(False, False, False))
    sch = scheme(w)
    a = PolySim(w)
    a.first_iteration()
    a.run(60)
    ck = str(tmp_path / "restart.npz")
    a.checkpoint(ck)
    d = np.load(ck)
    assert np.count_nonzero(np.any(d["fr"] != 0, axis=1)) > 50
    b = PolySim.restart(w, ck)
    o = po.Oracle(_particles(d), w.domain, w.walls)
    o.set_shapes(w.shapes)
    assert o.restore_interactions(d["ii"], d["jj"], d["tt"], d["fr"], d["mo"], sw=d["sw"], si=d["si"],
                                  sj=d["sj"]) == len(d["ii"])
    o.rebuild_poly(w.rcut_inc)

    def table(rows):
        ii, jj, tt, sw, si, sj, fr = rows
        keys = np.stack([ii.astype(np.int64), jj.astype(np.int64), tt.astype(np.int64), sw.astype(np.int64),
                         si.astype(np.int64), sj.astype(np.int64)], 1)
        order = np.lexsort(keys.T[::-1])
        return keys[order], fr[order]

    for when in ("after the restart", f"{STEPS} steps after the restart"):
        kg, fg = table(b.ctx.poly_interactions())
        ko, fo = table(o.poly_interactions())
        assert np.array_equal(kg, ko) and np.array_equal(fg, fo), when
        if when == "after the restart":
            assert np.count_nonzero(np.any(fg != 0, axis=1)) > 50
            b.run(STEPS)
            o.run_steps_poly(sch, STEPS)
    _same_state(b.ctx.download(), o.particles(), "polyhedra")


def test_mesh_restart_matches_oracle(gpu_ok, tmp_path):
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    sch = scheme(w)
    a = SphereSim(w, mesh=m)
    a.first_iteration()
    a.run(40)
    ck = str(tmp_path / "restart.npz")
    a.checkpoint(ck)
    d = np.load(ck)
    assert int(np.count_nonzero(np.any(d["m_fr"] != 0, axis=1))) > 100
    b = SphereSim.restart(w, ck, mesh=m)
    o = po.Oracle(_particles(d), w.domain, w.walls)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], m["minkowski"])
    nm = len(d["m_ii"])
    rows = (np.concatenate([d["ii"], d["m_ii"]]), np.concatenate([d["jj"], np.full(nm, m["mesh_id"], np.uint64)]),
            np.concatenate([d["tt"], d["m_tt"]]), np.concatenate([d["fr"], d["m_fr"]]),
            np.concatenate([d["mo"], d["m_mo"]]))
    sj = np.concatenate([np.zeros(len(d["ii"]), np.uint32), d["m_sj"]])
    assert o.restore_interactions(*rows, sj=sj) == len(rows[0])
    o.rebuild(w.rcut_inc)

    def mesh_rows(ii, sj, tt, fr):
        order = np.lexsort((sj, tt, ii))
        return np.stack([ii[order].astype(np.int64), sj[order].astype(np.int64), tt[order].astype(np.int64)], 1), \
            fr[order]

    for when in ("after the restart", f"{STEPS} steps after the restart"):
        _pairs_equal(b, o, when)
        kg, fg = mesh_rows(*b.ctx.mesh_interactions()[:4])
        ko, fo = mesh_rows(*o.mesh_interactions()[:4])
        assert np.array_equal(kg, ko) and np.array_equal(fg, fo), when
        if when == "after the restart":
            assert np.count_nonzero(np.any(fg != 0, axis=1)) > 100
            b.run(STEPS)
            o.run_steps(sch, STEPS)
    _same_state(b.ctx.download(), o.particles(), "mesh")
