"""Parity of the HIP sphere path (through the C ABI) with the CPU oracle, on seeded inputs.

Bars: neighbour streams and interaction lists bit-exact (same order, same keys); friction history
bit-exact; per-particle forces and moments within 1e-12 relative (floor: m|g|, the smallest force
scale of the case); 50-100 step trajectories within 1e-10 relative. The fused multi-step engine
must reproduce the operator-by-operator path bit for bit.
"""
import numpy as np
import pytest

from exadem_amd import native
from exadem_amd.sim import SphereSim, scheme
import pyoracle as po

from exadem_amd import inputs
from helpers import (by_id, dense_deposition, force, interaction_table, periodic_pack, pos, rel_err, vec, vel,
                     wall_deposition)

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-10


def oracle_for(w):
    return po.Oracle(w.particles, w.domain, w.walls)


def gpu_for(w):
    return SphereSim(w)


def compare_state(pg, po_, force_floor, what, ftol=FORCE_TOL, ptol=TRAJ_TOL):
    g, o = by_id(pg), by_id(po_)
    assert np.array_equal(g["id"], o["id"]), what
    L = max(1.0, float(np.max(np.abs(pos(o)))))
    dpos = float(np.max(np.abs(pos(g) - pos(o)))) / L
    vscale = max(1e-3, float(np.max(np.abs(vel(o)))))
    dvel = float(np.max(np.abs(vel(g) - vel(o)))) / vscale
    ef = rel_err(force(g), force(o), force_floor)
    mom_scale = np.maximum(np.linalg.norm(vec(o, "mom"), axis=1), 1e-6 * float(np.max(force_floor)))
    em = float(np.max(np.linalg.norm(vec(g, "mom") - vec(o, "mom"), axis=1) / mom_scale))
    assert dpos <= ptol, f"{what}: positions differ by {dpos:.3e} (relative to box)"
    assert dvel <= ptol, f"{what}: velocities differ by {dvel:.3e}"
    assert ef <= ftol, f"{what}: forces differ by {ef:.3e} relative"
    assert em <= ftol, f"{what}: moments differ by {em:.3e} relative"
    return dpos, dvel, ef, em


def test_upload_download_roundtrip(gpu_ok):
    w = periodic_pack(n=1000)
    ctx = native.Context()
    ctx.set_domain(w.domain)
    ctx.upload(w.particles)
    out = ctx.download()
    for k in ("rx", "ry", "rz", "vx", "radius", "mass", "inx", "qw"):
        assert np.array_equal(out[k], w.particles[k]), k
    assert np.array_equal(out["id"], w.particles["id"])


@pytest.mark.parametrize("case", ["dense", "walls", "periodic"])
def test_streams_bit_exact(gpu_ok, case):
    w = {"dense": dense_deposition, "walls": wall_deposition, "periodic": periodic_pack}[case]()
    o = oracle_for(w)
    o.rebuild(w.rcut_inc)
    s = gpu_for(w)
    s.rebuild()
    a, b, code, own = s.ctx.streams()
    oa, ob, ocode = o.streams()
    assert len(a) == len(oa) and len(a) > 0
    # the device stores particles along a Morton curve of cells, the oracle in cell-index order:
    # compare each particle's stream, whose order (stencil dz,dy,dx then id) must be identical
    g_ord = np.argsort(a, kind="stable")
    o_ord = np.argsort(oa, kind="stable")
    assert np.array_equal(a[g_ord], oa[o_ord])
    assert np.array_equal(b[g_ord], ob[o_ord]) and np.array_equal(code[g_ord], ocode[o_ord])
    # owner rule of nbh_sphere (symetric): lower id owns, periodic images are owned by both sides
    exp_own = (a < b) | (code != 13)
    assert np.array_equal(own.astype(bool), exp_own)
    # positions after the periodic wrap are identical
    pg, pq = by_id(s.ctx.download()), by_id(o.particles())
    assert np.array_equal(pg["id"], pq["id"])
    assert np.array_equal(pg["rx"], pq["rx"]) and np.array_equal(pg["rz"], pq["rz"])


def small_periodic_box(dims, seed):
    """Spheres (radius U[0.4, 0.55]) on a jittered lattice of spacing about 1.1 in a periodic box of dims
    cells of 1.25 (the neighbour distance is at most 1.2): three cells along an axis make every stencil
    wrap, and dims not a multiple of 3 end the neighbour build's row segments early."""
    rng = np.random.default_rng(seed)
    cell, sp = 1.25, 1.1
    L = [d * cell for d in dims]
    g = [max(1, int(round(Lk / sp))) for Lk in L]
    ix, iy, iz = np.meshgrid(*[np.arange(k) for k in g], indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(float)
    pts = (pts + 0.5) * np.array([Lk / k for Lk, k in zip(L, g)]) + rng.uniform(-0.02, 0.02, pts.shape)
    n = len(pts)
    r = rng.uniform(0.4, 0.55, n)
    v = rng.normal(0.0, 0.5, (n, 3))
    p = inputs.sphere_set(pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), r, 2500.0,
                          v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy())
    dom = native.domain((0.0, 0.0, 0.0), tuple(L), (True, True, True), cell)
    cp = native.contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    return inputs.Workload("small periodic", p, dom, [], cp, cp, 1e-3, 200, 0.1, (0.0, 0.0, 0.0))


@pytest.mark.parametrize("dims", [(3, 4, 5), (4, 3, 4), (5, 7, 3), (7, 4, 3), (3, 3, 4), (8, 5, 6)])
def test_small_periodic_boxes_match_oracle(gpu_ok, dims):
    """Streams bit-exact with the oracle after the first build, and 200 engine steps with rebuilds
    (velocities N(0, 0.5), dt 1e-3) equal to the oracle's, in periodic boxes of 3 to 8 cells per axis
    (fewer than 3 on a periodic axis is refused by set_domain)."""
    w = small_periodic_box(dims, seed=sum(dims))
    o = oracle_for(w)
    o.rebuild(w.rcut_inc)
    s = gpu_for(w)
    s.rebuild()
    a, b, code, own = s.ctx.streams()
    oa, ob, ocode = o.streams()
    assert len(a) == len(oa) and len(a) > 0
    g_ord = np.argsort(a, kind="stable")
    o_ord = np.argsort(oa, kind="stable")
    assert np.array_equal(a[g_ord], oa[o_ord])
    assert np.array_equal(b[g_ord], ob[o_ord]) and np.array_equal(code[g_ord], ocode[o_ord])
    sch = scheme(w)
    s2 = gpu_for(w)
    s2.first_iteration()
    st = s2.run(200, poll=8)
    o2 = oracle_for(w)
    o2.first_iteration(sch)
    o2.run_steps(sch, 200)
    floor = by_id(w.particles)["mass"] * 9.81
    compare_state(s2.ctx.download(), o2.particles(), floor, f"dims {dims}, 200 steps")
    print(f"dims {dims}: {w.n} spheres, {len(a)} stream entries, {st['rebuilds']} rebuilds", flush=True)


def test_periodic_axis_needs_three_cells(gpu_ok):
    w = small_periodic_box((2, 4, 4), seed=1)
    with pytest.raises(native.XdemError, match="at least 3 cells"):
        gpu_for(w)


@pytest.mark.parametrize("case", ["dense", "walls", "periodic"])
def test_first_iteration_forces(gpu_ok, case):
    w = {"dense": dense_deposition, "walls": wall_deposition, "periodic": periodic_pack}[case]()
    o = oracle_for(w)
    o.first_iteration(scheme(w))
    s = gpu_for(w)
    s.first_iteration()
    pg, pq = s.ctx.download(), o.particles()
    floor = by_id(pq)["mass"] * 9.81
    compare_state(pg, pq, floor, f"{case} first_iteration", ptol=0.0)
    # interaction lists: same keys, same (zero or carried) history
    gi, gj, gt, gf = s.ctx.interactions()
    oi, oj, ot, of, om = o.interactions()
    kg, fg = interaction_table(gi, gj, gt, gf)
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert np.array_equal(fg, fo)
    if case == "walls":
        assert np.sum(kg[:, 2] == 5) > 0


@pytest.mark.parametrize("case", ["dense", "walls", "periodic"])
def test_operator_steps_match_oracle(gpu_ok, case):
    w = {"dense": dense_deposition, "walls": wall_deposition, "periodic": periodic_pack}[case]()
    sch = scheme(w)
    o = oracle_for(w)
    o.first_iteration(sch)
    s = gpu_for(w)
    s.first_iteration()
    nsteps = 60
    orb = o.run_steps(sch, nsteps)
    for _ in range(nsteps):
        s.step()
    pg, pq = s.ctx.download(), o.particles()
    floor = by_id(pq)["mass"] * 9.81 * 1e-3
    compare_state(pg, pq, floor / by_id(pq)["mass"], f"{case} {nsteps} steps")
    assert s.rebuilds == orb
    gi, gj, gt, gf = s.ctx.interactions()
    oi, oj, ot, of, om = o.interactions()
    kg, fg = interaction_table(gi, gj, gt, gf)
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    scale = max(1e-30, float(np.max(np.abs(fo))))
    assert float(np.max(np.abs(fg - fo))) <= 1e-12 * scale


@pytest.mark.parametrize("spacing", [1.2, 0.98])
def test_deposition_config_1000_steps(gpu_ok, spacing):
    """The CPU-runnable deposition input in full: 10,000 spheres of radius 0.5 on a jittered lattice in a
    40 x 40 x 80 walled box under gravity, kn 1e5, kt 0.8 kn, mu 0.3, damping 0.9, dt 1e-5, 1000 steps,
    through the engine (one kernel per step) against the oracle. At lattice spacing 1.2 the spheres fall
    freely (the fall, 4.9e-4, is far below the gaps); spacing 0.98 is the contact-rich variant (initial
    overlaps). Same rebuilds, interaction keys and history, forces and trajectories as the oracle."""
    w = inputs.deposition(spacing=spacing)
    assert w.n == 10000 and w.steps == 1000
    sch = scheme(w)
    o = oracle_for(w)
    o.first_iteration(sch)
    s = gpu_for(w)
    s.first_iteration()
    orb = o.run_steps(sch, w.steps)
    st = s.run(w.steps, poll=64)
    pg, pq = s.ctx.download(), o.particles()
    floor = by_id(pq)["mass"] * 9.81 * 1e-3
    compare_state(pg, pq, floor / by_id(pq)["mass"], f"deposition spacing {spacing}, {w.steps} steps")
    assert st["rebuilds"] == orb
    gi, gj, gt, gf = s.ctx.interactions()
    oi, oj, ot, of, om = o.interactions()
    kg, fg = interaction_table(gi, gj, gt, gf)
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert np.array_equal(fg, fo)
    contacts = int(np.count_nonzero(np.any(fo != 0, axis=1)))
    print(f"deposition spacing {spacing}: rebuilds {orb}, contacts {contacts}")
    if spacing < 1.0:
        assert contacts > 1000


@pytest.mark.parametrize("case", ["dense", "periodic"])
def test_fused_engine_equals_operator_path(gpu_ok, case):
    w = {"dense": dense_deposition, "periodic": periodic_pack}[case]()
    a = gpu_for(w)
    a.first_iteration()
    b = gpu_for(w)
    b.first_iteration()
    for _ in range(40):
        a.step()
    b.run(40, poll=7)
    pa, pb = a.ctx.download(), b.ctx.download()
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "momx", "vrotx", "vroty", "qw", "qx", "arotx"):
        assert np.array_equal(pa[k], pb[k]), k
    assert a.rebuilds == b.rebuilds


def test_rebuild_history_carry_over(gpu_ok):
    """Small skin: rebuilds fire every few steps; history must survive them bit for bit."""
    w = periodic_pack(n=2000, sigma_v=3.0)
    w.rcut_inc = 0.004
    w.dt = 2e-4
    sch = scheme(w)
    o = oracle_for(w)
    o.first_iteration(sch)
    s = gpu_for(w)
    s.first_iteration()
    orb = o.run_steps(sch, 60)
    st = s.run(60, poll=5)
    assert orb >= 3, "the case must trigger several rebuilds"
    assert st["rebuilds"] == orb
    pg, pq = s.ctx.download(), o.particles()
    compare_state(pg, pq, by_id(pq)["mass"] * 1e-2, "rebuild carry-over", ptol=1e-9, ftol=1e-9)
    gi, gj, gt, gf = s.ctx.interactions()
    oi, oj, ot, of, om = o.interactions()
    kg, fg = interaction_table(gi, gj, gt, gf)
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert np.sum(np.any(fo != 0, axis=1)) > 0, "some interactions must carry history"
    scale = max(1e-30, float(np.max(np.abs(fo))))
    assert float(np.max(np.abs(fg - fo))) <= 1e-9 * scale


def test_deterministic(gpu_ok):
    w = periodic_pack(n=3000)
    runs = []
    for _ in range(2):
        s = gpu_for(w)
        s.first_iteration()
        s.run(30)
        runs.append(s.ctx.download())
    for k in ("rx", "vx", "fx", "momz"):
        assert np.array_equal(runs[0][k], runs[1][k]), k


def test_error_paths(gpu_ok):
    w = periodic_pack(n=500)
    ctx = native.Context()
    ctx.set_domain(w.domain)
    ctx.upload(w.particles)
    with pytest.raises(native.XdemError, match="STATE"):
        ctx.chunk_neighbors_contact(w.rcut_inc)  # before move_particles
    ctx.move_particles()
    ctx.chunk_neighbors_contact(w.rcut_inc)
    with pytest.raises(native.XdemError, match="UNSUPPORTED"):
        ctx.nbh_sphere(w.rcut_inc, symetric=False)
    ctx.nbh_sphere(w.rcut_inc)
    bad = native.contact_params(kn=1e5, fc=0.05, dncut=0.1)
    with pytest.raises(native.XdemError, match="UNSUPPORTED"):
        ctx.contact_sphere(bad, bad, 1e-5)
    dup = dict(w.particles)
    dup["id"] = np.zeros_like(dup["id"])
    with pytest.raises(native.XdemError, match="unique"):
        ctx.upload(dup)


def test_fp32_form_contact_set_and_force_error(gpu_ok):
    """FP32-force form: same contact set as FP64 (FP64 sign test), forces within 1e-5 relative."""
    w = periodic_pack(n=4000)
    a = SphereSim(w, form=native.XDEM_FORM_FP64)
    b = SphereSim(w, form=native.XDEM_FORM_FP32)
    a.first_iteration()
    b.first_iteration()
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    floor = pa["mass"] * 9.81
    err = rel_err(force(pb), force(pa), floor)
    assert err <= 1e-5, err
    merr = moment_rel_err(pb, pa)
    assert merr <= 1e-5, merr
    a.run(50)
    b.run(50)
    ka, _ = interaction_table(*a.ctx.interactions())
    kb, _ = interaction_table(*b.ctx.interactions())
    assert np.array_equal(ka, kb)
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    err = rel_err(force(pb), force(pa), 9.81 * np.ones(len(pa["id"])))
    assert err <= 1e-5, err
    merr = moment_rel_err(pb, pa)
    print(f"fp32 after 50 steps: force error {err:.3g}, moment error {merr:.3g}")
    assert merr <= 1e-5, merr
    assert float(np.max(np.abs(pos(pb) - pos(pa)))) <= 1e-9


def moment_rel_err(pb, pa):
    """max |M32 - M64| / |M64| over the particles whose moment is at least 1% of the largest."""
    ma = np.stack([pa["momx"], pa["momy"], pa["momz"]], axis=1)
    mb = np.stack([pb["momx"], pb["momy"], pb["momz"]], axis=1)
    mn = np.linalg.norm(ma, axis=1)
    sel = mn >= 1e-2 * float(np.max(mn))
    assert np.count_nonzero(sel) > 0
    return float(np.max(np.linalg.norm(mb - ma, axis=1)[sel] / mn[sel]))


def shell_cluster(n_shell=220, seed=5):
    """One sphere of radius 4 with n_shell spheres of radius 0.5 on a Fibonacci shell around it, each
    overlapping it by 0.05: the large sphere's list holds every shell sphere (longer than one 64-slot
    tile of the list), small spheres touch their shell neighbours too. No gravity, no walls."""
    k = np.arange(n_shell) + 0.5
    th = np.arccos(1.0 - 2.0 * k / n_shell)
    ph = np.pi * (1.0 + 5.0 ** 0.5) * k
    R = 4.0 + 0.5 - 0.05
    c = 12.0
    rx = np.concatenate([[c], c + R * np.sin(th) * np.cos(ph)])
    ry = np.concatenate([[c], c + R * np.sin(th) * np.sin(ph)])
    rz = np.concatenate([[c], c + R * np.cos(th)])
    rad = np.concatenate([[4.0], np.full(n_shell, 0.5)])
    v = inputs.normal(seed, 3, 3 * (n_shell + 1), 0.0, 0.05).reshape(3, -1)
    p = inputs.sphere_set(rx, ry, rz, rad, 2500.0, v[0], v[1], v[2])
    dom = native.domain((0.0, 0.0, 0.0), (24.0, 24.0, 24.0), (False, False, False), 8.1)
    cp = native.contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    return inputs.Workload("shell", p, dom, [], cp, cp, 1e-5, 40, 0.1, (0.0, 0.0, 0.0))


def test_long_list_matches_oracle(gpu_ok):
    """A particle with ~220 list entries (several 64-slot tiles): streams, history and 40 steps."""
    w = shell_cluster()
    o = oracle_for(w)
    g = gpu_for(w)
    o.first_iteration(scheme(w))
    g.first_iteration()
    gi, gj, gt, gf = g.ctx.interactions()
    assert int(np.sum(gi == 0)) + int(np.sum(gj == 0)) >= 200  # the large sphere (id 0) takes part in >= 200
    pg, pq = g.ctx.download(), o.particles()
    floor = by_id(pq)["mass"] * 1e-3
    compare_state(pg, pq, floor, "shell first pass")
    orb = o.run_steps(scheme(w), 40)
    for _ in range(40):
        g.step()
    compare_state(g.ctx.download(), o.particles(), floor, "shell 40 steps")
    assert g.rebuilds == orb
    kg, fg = interaction_table(*g.ctx.interactions())
    oi, oj, ot, of, om = o.interactions()
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_inputs(gpu_ok, n):
    """Empty input, a lone sphere and one touching pair: the operators and the fused engine run, and
    the state matches the oracle (free fall under gravity for the lone sphere)."""
    rx = np.array([5.0, 5.9])[:n]
    p = inputs.sphere_set(rx, np.full(n, 5.0), np.full(n, 5.0), np.full(n, 0.5), 2500.0,
                          np.array([0.1, -0.1])[:n], np.zeros(n), np.zeros(n))
    dom = native.domain((0.0, 0.0, 0.0), (10.0, 10.0, 10.0), (False, False, False), 1.1)
    cp = native.contact_params(kn=1e5, mu=0.3, damp_rate=0.9)
    w = inputs.Workload("tiny", p, dom, [], cp, cp, 1e-5, 20, 0.1, (0.0, 0.0, -9.81))
    a, b = gpu_for(w), gpu_for(w)
    a.first_iteration()
    b.first_iteration()
    for _ in range(20):
        a.step()
    b.run(20, poll=4)
    pa, pb = a.ctx.download(), b.ctx.download()
    assert len(pa["id"]) == n and len(pb["id"]) == n
    for k in ("rx", "rz", "vx", "vz", "fx", "fz"):
        assert np.array_equal(pa[k], pb[k]), k
    if n == 0:
        return
    o = oracle_for(w)
    o.first_iteration(scheme(w))
    o.run_steps(scheme(w), 20)
    compare_state(pa, o.particles(), np.full(n, 9.81), f"{n} particles 20 steps")


def test_sampled_kernel_timing(gpu_ok):
    """set_timing(every=k) times the integrate and force launches of every k-th step of the engine."""
    w = periodic_pack(n=2000)
    s = gpu_for(w)
    s.first_iteration()
    s.ctx.set_timing(True, every=4)
    s.ctx.kernel_timing(0, reset=True)
    s.ctx.kernel_timing(1, reset=True)
    s.run(16, poll=8)
    s.ctx.set_timing(False)
    f_ms, f_cnt = s.ctx.kernel_timing(1)
    i_ms, i_cnt = s.ctx.kernel_timing(0)
    # force passes of steps 0, 4, 8, 12 (a step re-issued after a rebuild stall is timed again); a
    # separate integrate launch only starts a chain of integrating force passes (step 0, after a stall)
    assert 4 <= f_cnt <= 8 and 1 <= i_cnt <= f_cnt
    assert f_ms > 0 and i_ms > 0
    with pytest.raises(native.XdemError):
        s.ctx.call("xdem_set_timing", -1)


@pytest.mark.parametrize("poll", [7, 64])
def test_engine_rebuilds_inside_step_chain(gpu_ok, poll):
    """Fast particles (sigma_v 1, dt 2e-4) so that the rebuild trigger fires several times inside the
    engine's chain of integrating force passes: the engine must equal the operator path bit for bit,
    with the same number of rebuilds."""
    w = periodic_pack(n=3000, sigma_v=1.0)
    w.dt = 2e-4
    a = gpu_for(w)
    a.first_iteration()
    b = gpu_for(w)
    b.first_iteration()
    nsteps = 600
    for _ in range(nsteps):
        a.step()
    b.run(nsteps, poll=poll)
    assert a.rebuilds >= 2
    assert a.rebuilds == b.rebuilds
    pa, pb = a.ctx.download(), b.ctx.download()
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "momx", "vrotx", "vroty", "qw", "qx", "arotx"):
        assert np.array_equal(pa[k], pb[k]), k


@pytest.mark.parametrize("form,case,kr", [("fp32", "walls", 0.0), ("fp64", "walls", 2e3), ("fp64", "periodic", 2e3),
                                          ("fp32", "periodic", 2e3)])
def test_engine_variants_equal_operator_path(gpu_ok, form, case, kr):
    """The engine's integrating passes with walls, with a rolling-moment law (kr > 0, rolling history
    slots) and in the FP32-force form reproduce the operator path bit for bit."""
    w = {"walls": wall_deposition, "periodic": periodic_pack}[case]()
    if kr:
        w.cp = native.contact_params(kn=1e5, kt=8e4, kr=kr, mu=0.3, damp_rate=0.9)
        w.cpd = w.cp
    f = native.XDEM_FORM_FP32 if form == "fp32" else native.XDEM_FORM_FP64
    a = SphereSim(w, form=f)
    a.first_iteration()
    b = SphereSim(w, form=f)
    b.first_iteration()
    for _ in range(30):
        a.step()
    b.run(30, poll=8)
    assert a.rebuilds == b.rebuilds
    pa, pb = a.ctx.download(), b.ctx.download()
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "momx", "vrotx", "vroty", "qw", "qx", "arotx"):
        assert np.array_equal(pa[k], pb[k]), k


@pytest.mark.parametrize("case", ["walls", "periodic"])
def test_contact_state_dump(gpu_ok, case, tmp_path):
    """Recorded contact state of a force pass (the reference's analysis tuple dn, contact point, f - ft,
    ft, with the positions the interaction dump writes): dn and the contact point agree with the recorded
    positions and radii, ft is the friction history of the interaction, the rows' forces add up to each
    particle's force (owner rows, the partner side of interior pairs, walls, gravity), the engine records
    what the operator path records, and the dump file reads back."""
    from exadem_amd import itools
    w = {"walls": wall_deposition, "periodic": periodic_pack}[case]()
    s = gpu_for(w)
    with pytest.raises(native.XdemError, match="STATE"):
        s.ctx.contact_state()
    s.ctx.set_contact_record(True)
    s.first_iteration()
    ii, jj, tt, v = s.ctx.contact_state()
    p = by_id(s.ctx.download())
    assert np.array_equal(p["id"], np.arange(len(p["id"])))
    pp = tt == 0
    assert np.count_nonzero(pp) > 0
    if case == "walls":
        assert np.count_nonzero(tt == 5) > 0
    R = p["radius"]
    ri, rj = v[:, 10:13], v[:, 13:16]
    assert np.array_equal(ri, pos(p)[ii.astype(np.int64)])
    d = ri[pp] - rj[pp]
    dist = np.linalg.norm(d, axis=1)
    dn = dist - (R[ii[pp].astype(np.int64)] + R[jj[pp].astype(np.int64)])
    assert np.all(v[pp, 0] < 0)
    assert float(np.max(np.abs(v[pp, 0] - dn))) <= 1e-12
    nrm = d / dist[:, None]
    cp = ri[pp] - nrm * (R[ii[pp].astype(np.int64)] + 0.5 * v[pp, 0])[:, None]
    assert float(np.max(np.abs(v[pp, 1:4] - cp))) <= 1e-12
    hi, hj, ht, hf = s.ctx.interactions()
    hist = {(int(a), int(b), int(t)): f for a, b, t, f in zip(hi, hj, ht, hf)}
    for k in range(len(ii)):
        assert np.array_equal(v[k, 7:10], hist[(int(ii[k]), int(jj[k]), int(tt[k]))])
    # per particle: gravity, + f of its own rows, - f of the interior pairs it is the partner of
    # (a pair across a periodic boundary has a row on each side)
    F = p["mass"][:, None] * np.asarray(w.gravity)[None, :]
    f = v[:, 4:7] + v[:, 7:10]
    keys = set(zip(ii.tolist(), jj.tolist(), tt.tolist()))
    for k in range(len(ii)):
        a = int(ii[k])
        F[a] += f[k]
        if tt[k] == 0 and (int(jj[k]), a, 0) not in keys:
            F[int(jj[k])] -= f[k]
    assert rel_err(F, force(p), p["mass"] * 9.81) <= 1e-12
    # the engine's last step records the same rows as the operator path
    a = gpu_for(w)
    a.ctx.set_contact_record(True)
    a.first_iteration()
    for _ in range(20):
        a.step()
    b = gpu_for(w)
    b.ctx.set_contact_record(True)
    b.first_iteration()
    b.run(20, poll=7)
    for x, y in zip(a.ctx.contact_state(), b.ctx.contact_state()):
        assert np.array_equal(x, y)
    # the dump file in the reference's layout reads back to 6 significant digits
    path, rows = itools.write_dump(b.ctx, str(tmp_path), "InteractionOutputDir-", 20)
    keys2, vals2 = itools.read_dump(path)
    bi, bj, bt, bv = b.ctx.contact_state()
    assert rows == len(bi) > 0
    assert np.array_equal(keys2[:, 0], bi.astype(np.int64)) and np.array_equal(keys2[:, 4], bt.astype(np.int64))
    assert np.allclose(vals2, bv, rtol=1e-5, atol=1e-300)
    # the FP32 form does not record
    c = SphereSim(w, form=native.XDEM_FORM_FP32)
    c.ctx.set_contact_record(True)
    c.first_iteration()
    with pytest.raises(native.XdemError, match="STATE"):
        c.ctx.contact_state()


def packing_stats(p, inter, dom):
    """Coordination number 2 * (active particle pairs) / N, as the reference's interaction counter
    forms it (itools/counter.hpp:145-151; active here: non-zero friction history, which every pair in
    contact with a tangential velocity has), and solid fraction of the bed: sphere volume over the box
    section times the height of the highest sphere top."""
    ii, jj, tt, fr = inter[:4]
    act = np.any(fr != 0, axis=1) & (tt == 0)
    cn = 2.0 * np.count_nonzero(act) / len(p["id"])
    top = float(np.max(p["rz"] + p["radius"]))
    lo, hi = list(dom.lo), list(dom.hi)
    vol = float(np.sum(4.0 / 3.0 * np.pi * p["radius"] ** 3))
    return cn, vol / ((hi[0] - lo[0]) * (hi[1] - lo[1]) * (top - lo[2]))


def test_packing_statistics_match_oracle(gpu_ok):
    """Packing statistics of the contact-rich walled deposition after 100 steps: the FP64 engine's
    coordination number and solid fraction equal the oracle's; the FP32-force form has the same
    coordination number (bit-exact contact set) and its solid fraction within 1e-9."""
    w = dense_deposition()
    sch = scheme(w)
    o = oracle_for(w)
    o.first_iteration(sch)
    o.run_steps(sch, 100)
    s = gpu_for(w)
    s.first_iteration()
    s.run(100)
    b = SphereSim(w, form=native.XDEM_FORM_FP32)
    b.first_iteration()
    b.run(100)
    cn_o, phi_o = packing_stats(o.particles(), o.interactions(), w.domain)
    cn_g, phi_g = packing_stats(s.ctx.download(), s.ctx.interactions(), w.domain)
    cn_b, phi_b = packing_stats(b.ctx.download(), b.ctx.interactions(), w.domain)
    print(f"coordination number {cn_g:.6f} (oracle {cn_o:.6f}, FP32 {cn_b:.6f}), solid fraction {phi_g:.9f} "
          f"(oracle {phi_o:.9f}, FP32 {phi_b:.9f})")
    assert cn_o > 1.0
    assert cn_g == cn_o and phi_g == phi_o
    assert cn_b == cn_g and abs(phi_b - phi_g) <= 1e-9


@pytest.mark.parametrize("case", ["walls", "periodic", "walls_kr"])
def test_checkpoint_restart_with_history(gpu_ok, case, tmp_path):
    """Restart with interaction history: a run checkpointed after 60 steps (particles and interaction
    history) and restarted (particles uploaded, lists rebuilt, history handed over by key) is, bit for
    bit, the uninterrupted run with a neighbour rebuild at the checkpoint step; against the run without
    that rebuild, after 60 more steps: the same contacts (interactions with history), friction history
    within 1e-12 and trajectories within 1e-10 (the lists are built at other positions, so the
    summation order and the rebuild steps may differ). Restarting without the history gives another
    state. walls_kr: with a rolling stiffness, so the rolling-moment history is saved and handed over
    as well."""
    w = {"walls": wall_deposition, "periodic": periodic_pack, "walls_kr": wall_deposition}[case]()
    if case == "walls_kr":
        w.cp = native.contact_params(kn=1e5, kt=8e4, mu=0.3, damp_rate=0.9, kr=2e3)
        w.cpd = w.cp
    a = gpu_for(w)
    a.first_iteration()
    a.run(60)
    ck = str(tmp_path / "restart.npz")
    a.checkpoint(ck)
    nhist = int(np.count_nonzero(np.any(np.load(ck)["fr"] != 0, axis=1)))
    assert nhist > 0
    if case == "walls_kr":
        assert np.load(ck)["mo"].shape == np.load(ck)["fr"].shape
    # what a restart is: the uninterrupted run with a neighbour rebuild at the checkpoint step
    r = gpu_for(w)
    r.first_iteration()
    r.run(60)
    r.rebuild()
    r.run(60)
    a.run(60)
    b = SphereSim.restart(w, ck)
    b.run(60)
    pr, pb = r.ctx.download(), b.ctx.download()
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "vrotx", "fx", "momx"):
        assert np.array_equal(by_id(pr)[k], by_id(pb)[k]), k
    c = SphereSim.restart(w, ck, history=False)
    c.run(60)
    pa, pb, pc = a.ctx.download(), b.ctx.download(), c.ctx.download()
    floor = by_id(pa)["mass"] * 9.81 * 1e-3
    compare_state(pb, pa, floor / by_id(pa)["mass"], f"{case} restart")
    def active(sim):
        k, f = interaction_table(*sim.ctx.interactions())
        act = np.any(f != 0, axis=1)
        return k[act], f[act]

    # interactions with history (the contacts); the lists also hold non-contacting candidate pairs,
    # which depend on the positions the lists were built at
    ka, fa = active(a)
    kb, fb = active(b)
    kc, fc = active(c)
    assert len(ka) > 0 and np.array_equal(ka, kb)
    scale = float(np.max(np.abs(fa)))
    assert float(np.max(np.abs(fb - fa))) <= 1e-12 * scale
    assert not np.array_equal(kc, ka) or float(np.max(np.abs(fc - fa))) > 1e-6 * scale


@pytest.mark.parametrize("case", ["walls", "periodic"])
def test_simulation_state(gpu_ok, case):
    """Thermodynamic state reduced on the device (simulation_state) against the same quantities in
    float64 numpy from the downloaded fields and interactions: energies and mass within 1e-12
    relative, particle and interaction counts exact (pairs twice, walls once; active = overlapping or
    with history), the smallest dn from the positions, the domain volume."""
    w = {"walls": wall_deposition, "periodic": periodic_pack}[case]()
    s = gpu_for(w)
    s.first_iteration()
    s.run(20)
    st = s.ctx.simulation_state()
    p = by_id(s.ctx.download())
    m = p["mass"]
    for a, c in enumerate("xyz"):
        ke = float(np.sum(0.5 * p["v" + c] * p["v" + c] * m))
        re = float(np.sum(0.5 * p["vrot" + c] * p["vrot" + c] * m))
        assert abs(st["kinetic_energy"][a] - ke) <= 1e-12 * max(ke, 1e-300), (c, st["kinetic_energy"][a], ke)
        assert abs(st["rotation_energy"][a] - re) <= 1e-12 * max(re, 1e-300) or re == 0.0
    assert abs(st["mass"] - float(np.sum(m))) <= 1e-12 * float(np.sum(m))
    assert st["n_particles"] == len(m)
    lo, hi = np.array(w.domain.lo), np.array(w.domain.hi)
    assert st["volume"] == pytest.approx(float(np.prod(hi - lo)), rel=1e-15)
    ii, jj, tt, fr = s.ctx.interactions()
    r, R = pos(p), p["radius"]
    per = np.array(list(w.domain.periodic), bool)
    tot = act = 0
    min_dn = 0.0
    for k in range(len(ii)):
        a = int(ii[k])
        if tt[k] == 0:
            d = r[a] - r[int(jj[k])]
            d = np.where(per, d - (hi - lo) * np.round(d / (hi - lo)), d)
            dn = float(np.sqrt(np.dot(d, d))) - (R[a] + R[int(jj[k])])
            coef = 2
        else:
            wall = w.walls[int(jj[k])]
            dn = abs(float(np.dot(r[a], wall["normal"])) - wall["offset"]) - R[a]
            coef = 1
        tot += coef
        if dn < 0 or np.any(fr[k] != 0):
            act += coef
            min_dn = min(min_dn, dn if dn < 0 else 0.0)
    print(f"{case}: {st['n_interactions']} interactions, {st['n_active']} active, min dn {st['min_dn']:.6g}")
    assert st["n_interactions"] == tot and st["n_active"] == act and act > 0
    assert abs(st["min_dn"] - min_dn) <= 1e-12 and min_dn < 0
