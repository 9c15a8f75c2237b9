"""Polyhedra against the triangle-mesh driver: the HIP path (through the C ABI) against the CPU oracle.

Input: 98 cubes and octahedra (density 1, random orientations and spins) falling at 3 m/s onto the tray
with the bump (mesh_cases.tray_polyhedra, Minkowski radius 0.01 on the mesh). Bars: the interaction rows
-- feature pairs, mesh candidates (types 7-12, partner = the mesh id, sub_j = the mesh feature) and their
friction history -- bit-exact per owner in storage order, and positions, velocities, orientations,
forces and moments bit-exact, after first_iteration and after 1500 and 3000 engine steps with
neighbour rebuilds; a restart from a checkpoint equal to the oracle restarted from the same file; the
classifier wave sizes, simulation_state counts and recorded contact state consistent with the rows.
"""
import numpy as np
import pytest

from exadem_amd.sim import PolySim, scheme
import pyoracle as po

import mesh_cases as mc
from helpers import by_id, force, pos, rel_err

pytestmark = pytest.mark.gpu

RM = 0.01


def rows_table(rows):
    ii, jj, tt, sw, si, sj, fr = rows
    order = np.argsort(ii, kind="stable")
    keys = np.stack([ii.astype(np.int64), jj.astype(np.int64), tt.astype(np.int64), sw.astype(np.int64),
                     si.astype(np.int64), sj.astype(np.int64)], axis=1)
    return keys[order], fr[order]


def same_rows(s, o, what):
    kg, fg = rows_table(s.ctx.poly_interactions())
    ko, fo = rows_table(o.poly_interactions())
    assert len(kg) == len(ko) and np.array_equal(kg, ko), what
    assert np.array_equal(fg, fo), what
    return kg, fg


def same_state(g, o, what):
    g, o = by_id(g), by_id(o)
    assert np.array_equal(g["id"], o["id"]), what
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "qw", "qx", "qy", "qz", "vrotx", "vroty", "vrotz", "fx", "fy",
              "fz", "momx", "momy", "momz"):
        assert np.array_equal(g[k], o[k]), (what, k)


def setup():
    w = mc.tray_polyhedra()
    m = dict(mc.tray_mesh(RM), max_candidates=128)  # a polyhedron meets more mesh features than a sphere
    s = PolySim(w, mesh=m)
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.set_shapes(w.shapes)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], RM)
    return w, m, s, o


def test_tray_polyhedra_match_oracle(gpu_ok):
    w, m, s, o = setup()
    sch = scheme(w)
    s.first_iteration()
    o.first_iteration_poly(sch)
    same_rows(s, o, "first_iteration")
    same_state(s.ctx.download(), o.particles(), "first_iteration")
    seen = set()
    for k in range(2):
        st = s.run(1500)
        o.run_steps_poly(sch, 1500)
        keys, fr = same_rows(s, o, f"{1500 * (k + 1)} steps")
        same_state(s.ctx.download(), o.particles(), f"{1500 * (k + 1)} steps")
        act = np.any(fr != 0, axis=1)
        seen |= set(keys[act, 2].tolist())
        print(f"{1500 * (k + 1)} steps: {st['rebuilds']} rebuilds, contacts by type "
              f"{np.bincount(keys[act, 2], minlength=13).tolist()}", flush=True)
    assert {9, 10}.issubset(seen) and seen & {7, 8} and seen & {11, 12} and seen & {0, 1, 2, 3}, seen


def test_tray_polyhedra_restart_matches_oracle(gpu_ok, tmp_path):
    """Checkpoint after 1500 steps; the HIP restart (rows handed to the rebuilt lists by key) and the
    oracle restored from the same rows are equal right after the restart and 300 steps later."""
    w, m, s, _ = setup()
    sch = scheme(w)
    s.first_iteration()
    s.run(1500)
    ck = str(tmp_path / "restart.npz")
    s.checkpoint(ck)
    d = np.load(ck)
    assert np.count_nonzero(np.any(d["fr"][d["tt"] >= 7] != 0, axis=1)) > 5
    b = PolySim.restart(w, ck, mesh=m)
    o = po.Oracle({k[2:]: d[k] for k in d.files if k.startswith("p_")}, w.domain, w.walls)
    o.set_shapes(w.shapes)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], RM)
    assert o.restore_interactions(d["ii"], d["jj"], d["tt"], d["fr"], d["mo"], sw=d["sw"], si=d["si"],
                                  sj=d["sj"]) == len(d["ii"])
    o.rebuild_poly(w.rcut_inc)
    same_rows(b, o, "after the restart")
    b.run(300)
    o.run_steps_poly(sch, 300)
    same_rows(b, o, "300 steps after the restart")
    same_state(b.ctx.download(), o.particles(), "300 steps after the restart")


def test_tray_polyhedra_waves_state_and_dump(gpu_ok):
    """After 1500 steps and one recorded pass: classifier wave sizes 7-12 are the mesh rows by type;
    simulation_state counts feature rows twice and mesh rows once, active ones by their history; the
    recorded contact state has one row per contact, mesh rows with the particle's force f = fn + ft and
    ft equal to the stored friction, and the rows' forces add up to each particle's force."""
    w, m, s, _ = setup()
    s.first_iteration()
    s.run(1500)
    s.ctx.set_contact_record(True)
    s.rebuild()
    s.ctx.reset_force_moment()
    s.compute_force()
    ii, jj, tt, sw, si, sj, fr = s.ctx.poly_interactions()
    sizes = s.ctx.classifier_waves()
    assert np.array_equal(sizes[7:13], np.bincount(tt[tt >= 7].astype(int), minlength=13)[7:13])
    assert np.sum(sizes[7:13]) > 20
    st = s.ctx.simulation_state()
    coef = np.where(tt >= 5, 1, 2)
    assert st["n_interactions"] == int(np.sum(coef))
    assert st["n_active"] == int(np.sum(coef[np.any(fr != 0, axis=1)]))
    ci, cj, csi, csj, ctt, v = s.ctx.contact_state(subs=True)
    assert np.all(np.diff(ctt.astype(np.int64)) >= 0)  # one block per type, as the reference writes
    hist = {(int(a), int(b), int(t), int(x), int(y), int(z)): f for a, b, t, x, y, z, f in zip(ii, jj, tt, sw, si, sj, fr)}
    p = by_id(s.ctx.download())
    F = p["mass"][:, None] * np.asarray(w.gravity)[None, :]
    nmesh = 0
    for k in range(len(ci)):
        a, b, t = int(ci[k]), int(cj[k]), int(ctt[k])
        if t >= 7:
            key = (a, b, t, 0, int(csi[k]), int(csj[k]))
            nmesh += 1
        else:
            key = (a, b, t, 0, int(csi[k]), int(csj[k]))
            if key not in hist:
                key = (b, a, t, 1, int(csj[k]), int(csi[k]))
        assert np.array_equal(v[k, 7:10], hist[key]), key
        f = v[k, 4:7] + v[k, 7:10]
        F[a] += f
        if t < 4:
            F[b] -= f
    assert nmesh > 5
    assert rel_err(F, force(p), p["mass"] * 9.81) <= 1e-12
