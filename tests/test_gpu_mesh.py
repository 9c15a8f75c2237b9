"""Parity of the stationary triangle-mesh driver acting on spheres (register_stl_mesh; candidates
detection_sphere_driver, nbh_sphere.hpp:25-86; contact ContactLawRShapeDriverFunc, contact_sphere.hpp:298-396)
between the HIP path and the CPU oracle.

Input: two jittered layers of spheres resting on a heightfield tray (flat faces, a convex apex vertex,
convex ridge edges), plus one sphere on the apex and one on a ridge edge, with velocities N(0, 0.3) under
gravity. Bars: mesh candidates (owner, feature, type) and friction history bit-exact, forces equal to the
oracle's within 1e-12, trajectories of the operator path (with a neighbour rebuild half-way, which carries
the mesh history over) and of the engine within 1e-10 after 60 steps.
"""
import numpy as np
import pytest

from exadem_amd import native
from exadem_amd.sim import SphereSim, scheme
import pyoracle as po

import mesh_cases as mc
from helpers import by_id, force, interaction_table, pos, rel_err, vel

pytestmark = pytest.mark.gpu


def oracle_for(w, m):
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], m["minkowski"])
    return o


def mesh_table(ii, sj, tt, fr):
    order = np.lexsort((sj, tt, ii))
    return np.stack([ii[order].astype(np.int64), sj[order].astype(np.int64), tt[order].astype(np.int64)], 1), fr[order]


def check_pairs(s, o):
    kg, fg = interaction_table(*s.ctx.interactions())
    oi, oj, ot, of, om = o.interactions()
    keep = ot < 7
    ko, fo = interaction_table(oi[keep], oj[keep], ot[keep], of[keep])
    assert np.array_equal(kg, ko)
    return fg, fo


def check_state(g, pq):
    L = max(1.0, float(np.max(np.abs(pos(pq)))))
    assert float(np.max(np.abs(pos(g) - pos(pq)))) / L <= 1e-10
    assert float(np.max(np.abs(vel(g) - vel(pq)))) / max(1e-3, float(np.max(np.abs(vel(pq))))) <= 1e-10


@pytest.mark.parametrize("rm", [0.0, 0.02])
def test_mesh_matches_oracle(gpu_ok, rm):
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(rm)
    assert w.n > 500
    sch = scheme(w)
    o = oracle_for(w, m)
    o.first_iteration(sch)
    s = SphereSim(w, mesh=m)
    s.first_iteration()
    pg, pq = by_id(s.ctx.download()), by_id(o.particles())
    assert rel_err(force(pg), force(pq), pq["mass"] * 9.81) <= 1e-12
    kg, fg = mesh_table(*s.ctx.mesh_interactions()[:4])
    ko, fo = mesh_table(*o.mesh_interactions()[:4])
    assert np.array_equal(kg, ko) and np.array_equal(fg, fo)
    active = {t: int(np.count_nonzero((ko[:, 2] == t) & np.any(fo != 0, axis=1))) for t in (7, 8, 9)}
    assert active[7] >= 1 and active[8] >= 1 and active[9] > 100, active
    fg2, fo2 = check_pairs(s, o)
    assert np.array_equal(fg2, fo2)
    # operator path with a neighbour rebuild after 30 steps (mesh history carried by key)
    half = w.steps // 2
    o.run_steps(sch, half)
    for _ in range(half):
        s.step()
    o.rebuild(w.rcut_inc)
    s.rebuild()
    o.run_steps(sch, w.steps - half)
    for _ in range(w.steps - half):
        s.step()
    check_state(by_id(s.ctx.download()), by_id(o.particles()))
    kg, fg = mesh_table(*s.ctx.mesh_interactions()[:4])
    ko, fo = mesh_table(*o.mesh_interactions()[:4])
    assert np.array_equal(kg, ko)
    assert float(np.max(np.abs(fg - fo))) <= 1e-12 * max(1e-30, float(np.max(np.abs(fo))))
    assert np.count_nonzero(np.any(fo != 0, axis=1)) > 100
    # engine
    o2 = oracle_for(w, m)
    o2.first_iteration(sch)
    o2.run_steps(sch, w.steps)
    e = SphereSim(w, mesh=m)
    e.first_iteration()
    e.run(w.steps, poll=7)
    check_state(by_id(e.ctx.download()), by_id(o2.particles()))
    kg, fg = mesh_table(*e.ctx.mesh_interactions()[:4])
    ko, fo = mesh_table(*o2.mesh_interactions()[:4])
    assert np.array_equal(kg, ko)
    assert float(np.max(np.abs(fg - fo))) <= 1e-12 * max(1e-30, float(np.max(np.abs(fo))))
    print(f"rm={rm}: {w.n} spheres, active mesh contacts at the first pass {active}")


def test_mesh_errors(gpu_ok):
    """Bad meshes fail with XDEM_ERR_ARG; a sphere with more candidate features than the slots fails the
    neighbour build with XDEM_ERR_CAPACITY."""
    m = mc.tray_mesh()
    ctx = native.Context()
    bad = dict(m, edges=np.array([[0, 10**6]], np.uint32))
    with pytest.raises(native.XdemError, match="ARG"):
        ctx.set_mesh(**bad)
    with pytest.raises(native.XdemError, match="ARG"):
        ctx.set_mesh(**dict(m, minkowski=-1.0))
    w = mc.workload("tray", mc.tray_bed())
    s = SphereSim(w, mesh=dict(m, max_candidates=1))
    with pytest.raises(native.XdemError, match="CAPACITY"):
        s.first_iteration()
    # the same context recovers with enough slots
    s.ctx.set_mesh(**m)
    s.first_iteration()
    assert len(s.ctx.mesh_interactions()[0]) > w.n


def test_mesh_fp32_form(gpu_ok):
    """The FP32-force form evaluates the mesh contacts in FP64, as the walls: same candidates, forces within
    1e-5 of the FP64 form's at the first pass and after 60 engine steps."""
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    a = SphereSim(w, mesh=m)
    a.first_iteration()
    b = SphereSim(w, mesh=m, form=native.XDEM_FORM_FP32)
    b.first_iteration()
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    assert rel_err(force(pb), force(pa), pa["mass"] * 9.81) <= 1e-5
    ka, _ = mesh_table(*a.ctx.mesh_interactions()[:4])
    kb, _ = mesh_table(*b.ctx.mesh_interactions()[:4])
    assert np.array_equal(ka, kb)
    a.run(w.steps, poll=7)
    b.run(w.steps, poll=7)
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    L = max(1.0, float(np.max(np.abs(pos(pa)))))
    assert float(np.max(np.abs(pos(pb) - pos(pa)))) / L <= 1e-9


def test_mesh_checkpoint_restart(gpu_ok, tmp_path):
    """A checkpoint carries the mesh histories: restarting from it (particles uploaded, lists rebuilt,
    pair and mesh histories handed over by key) is, bit for bit, the uninterrupted run with a neighbour
    rebuild at the checkpoint step; restarting without the history gives another state."""
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    a = SphereSim(w, mesh=m)
    a.first_iteration()
    a.run(40)
    ck = str(tmp_path / "restart.npz")
    a.checkpoint(ck)
    d = np.load(ck)
    assert int(np.count_nonzero(np.any(d["m_fr"] != 0, axis=1))) > 100
    r = SphereSim(w, mesh=m)
    r.first_iteration()
    r.run(40)
    r.rebuild()
    r.run(40)
    b = SphereSim.restart(w, ck, mesh=m)
    b.run(40)
    pr, pb = by_id(r.ctx.download()), by_id(b.ctx.download())
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "vrotx", "fx", "momx"):
        assert np.array_equal(pr[k], pb[k]), k
    kr, fr = mesh_table(*r.ctx.mesh_interactions()[:4])
    kb, fb = mesh_table(*b.ctx.mesh_interactions()[:4])
    assert np.array_equal(kr, kb) and np.array_equal(fr, fb)
    c = SphereSim.restart(w, ck, mesh=m, history=False)
    c.run(40)
    assert not np.array_equal(by_id(c.ctx.download())["vx"], pb["vx"])


def mesh_dn(m, rows_t, rows_s, p, R):
    """dn of each mesh row (type 7/8/9, feature) for spheres p [k][3] of radius R, None without contact:
    vertex distance, edge distance for projections strictly inside the edge, face distance for
    projections inside the triangle; minus R + rm."""
    v, e, f, rm = m["vertices"], m["edges"], m["faces"], m["minkowski"]
    out = []
    for t, k, x, r in zip(rows_t, rows_s, p, R):
        if t == 7:
            dn = float(np.linalg.norm(x - v[k])) - (r + rm)
            out.append(dn if dn < 0 else None)
        elif t == 8:
            a, b = v[e[k][0]], v[e[k][1]]
            s = float(np.dot(b - a, x - a) / np.dot(b - a, b - a))
            dn = float(np.linalg.norm(x - a - (b - a) * s)) - (r + rm)
            out.append(dn if 0 < s < 1 and dn <= 0 else None)
        else:
            a, b, c = v[f[k][0]], v[f[k][1]], v[f[k][2]]
            n = np.cross(b - a, c - a)
            n = n / np.linalg.norm(n)
            h = float(np.dot(x - a, n))
            q = x - h * n
            T = np.stack([b - a, c - a], 1)
            uv = np.linalg.lstsq(T, q - a, rcond=None)[0]
            inside = uv[0] > 0 and uv[1] > 0 and uv.sum() < 1
            dn = abs(h) - (r + rm)
            out.append(dn if inside and dn <= 0 else None)
    return out


def test_mesh_simulation_state(gpu_ok):
    """simulation_state counts each mesh candidate once (pairs twice): totals exact, active = overlapping
    or with history, the smallest dn over pairs and mesh contacts from the positions."""
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    s = SphereSim(w, mesh=m)
    s.first_iteration()
    s.run(20)
    st = s.ctx.simulation_state()
    p = by_id(s.ctx.download())
    r, R = pos(p), p["radius"]
    ii, jj, tt, fr = s.ctx.interactions()
    tot = act = 0
    min_dn = 0.0
    for k in range(len(ii)):
        a, b = int(ii[k]), int(jj[k])
        dn = float(np.linalg.norm(r[a] - r[b])) - (R[a] + R[b])
        tot += 2
        if dn < 0 or np.any(fr[k] != 0):
            act += 2
            min_dn = min(min_dn, dn if dn < 0 else 0.0)
    mi, ms, mt, mf, _ = s.ctx.mesh_interactions()
    dns = mesh_dn(m, mt, ms, r[mi.astype(np.int64)], R[mi.astype(np.int64)])
    for k, dn in enumerate(dns):
        tot += 1
        if (dn is not None and dn < 0) or np.any(mf[k] != 0):
            act += 1
            min_dn = min(min_dn, dn if dn is not None and dn < 0 else 0.0)
    print(f"mesh: {st['n_interactions']} interactions ({len(mi)} mesh), {st['n_active']} active, "
          f"min dn {st['min_dn']:.6g}")
    assert st["n_interactions"] == tot and st["n_active"] == act
    assert abs(st["min_dn"] - min_dn) <= 1e-12 and min_dn < 0


def test_mesh_contact_dump(gpu_ok):
    """The recorded contact state (the interaction dump) lists the mesh contacts: partner = the mesh's
    driver index, sub_j = the feature, type 7/8/9; exactly the candidates in contact, with dn from the
    positions and ft equal to the friction history."""
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    s = SphereSim(w, mesh=m)
    s.ctx.set_contact_record(True)
    s.first_iteration()
    ii, jj, si, sj, tt, vals = s.ctx.contact_state(subs=True)
    assert np.all(np.diff(tt.astype(np.int64)) >= 0)  # one block per type, as the reference writes
    sel = tt >= 7
    assert np.all(jj[sel] == m["mesh_id"]) and np.all(si[sel] == 0)
    mi, ms, mt, mf, _ = s.ctx.mesh_interactions()
    hist = {(int(a), int(b), int(c)): k for k, (a, b, c) in enumerate(zip(mi, mt, ms))}
    p = by_id(s.ctx.download())
    r, R = pos(p), p["radius"]
    dns = mesh_dn(m, mt, ms, r[mi.astype(np.int64)], R[mi.astype(np.int64)])
    assert int(sel.sum()) == sum(d is not None for d in dns) > 100
    for a, t, f, v in zip(ii[sel], tt[sel], sj[sel], vals[sel]):
        k = hist[(int(a), int(t), int(f))]
        assert dns[k] is not None and abs(v[0] - dns[k]) <= 1e-12
        assert np.array_equal(v[7:10], mf[k])
        assert np.array_equal(v[10:13], r[int(a)]) and np.all(v[13:16] == 0)


def test_classifier_waves_spheres(gpu_ok):
    """Classifier wave sizes of a sphere context: particle pairs (one per pair), and the mesh candidates by
    kind (vertex 7, edge 8, face 9) equal to the oracle's items of those types; walls by driver kind in a
    walled box."""
    w = mc.workload("tray", mc.tray_bed())
    m = mc.tray_mesh(0.02)
    s = SphereSim(w, mesh=m)
    s.first_iteration()
    o = oracle_for(w, m)
    o.first_iteration(scheme(w))
    sizes = s.ctx.classifier_waves()
    oi, oj, ot, of, om = o.interactions()
    assert sizes[0] == int(np.count_nonzero(ot == 0)) > 0
    for t in (7, 8, 9):
        assert sizes[t] == int(np.count_nonzero(ot == t)) > 0, t
    assert np.all(sizes[1:7] == 0) and np.all(sizes[10:] == 0)
    from helpers import wall_deposition
    ww = wall_deposition()
    b = SphereSim(ww)
    b.first_iteration()
    bs = b.ctx.classifier_waves()
    _, _, bt, _ = b.ctx.interactions()
    assert bs[0] == int(np.count_nonzero(bt == 0)) and bs[5] == int(np.count_nonzero(bt == 5)) > 0


def test_placed_mesh_matches_oracle(gpu_ok):
    """The tray registered with a pose (quaternion of 90 degrees about z, centre (20, 0, 0), which maps the
    tray onto itself with its vertices renumbered): the first pass and 30 engine steps bit-exact with
    the oracle given the same pose."""
    import math
    w = mc.workload("tray", mc.tray_bed())
    m = dict(mc.tray_mesh(0.02), center=(20.0, 0.0, 0.0), quat=(math.cos(math.pi / 4), 0.0, 0.0, math.sin(math.pi / 4)))
    sch = scheme(w)
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.set_mesh(m["mesh_id"], m["vertices"], m["edges"], m["faces"], m["minkowski"], center=m["center"], quat=m["quat"])
    o.first_iteration(sch)
    s = SphereSim(w, mesh=m)
    s.first_iteration()
    kg, fg = mesh_table(*s.ctx.mesh_interactions()[:4])
    ko, fo = mesh_table(*o.mesh_interactions()[:4])
    assert np.array_equal(kg, ko) and np.array_equal(fg, fo) and np.count_nonzero(np.any(fo != 0, axis=1)) > 100
    s.run(30)
    o.run_steps(sch, 30)
    check_state(by_id(s.ctx.download()), by_id(o.particles()))
    kg, fg = mesh_table(*s.ctx.mesh_interactions()[:4])
    ko, fo = mesh_table(*o.mesh_interactions()[:4])
    assert np.array_equal(kg, ko) and np.array_equal(fg, fo)
