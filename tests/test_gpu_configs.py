"""The benchmark workloads at their full size, through the C ABI, against the CPU oracle.

- 1,000,000 polydisperse spheres (radius U[0.5, 1], periodic, solid fraction 0.55, Gaussian velocities
  sigma 0.1; the benchmark input): neighbour streams, interaction keys and friction history bit-exact
  after first_iteration and after 20 fused engine steps; forces and moments within 1e-12.
- 200,000 polyhedra (cubes and octahedra, density 1, random orientations, grown to solid fraction 0.45
  of the box; the benchmark's polyhedra input): feature lists bit-exact, the first force
  pass and 20 engine steps within 1e-12 (forces) and 1e-10 (positions).
- the learned surrogate (K = 10 network) on the 1M input: the drift-checked run without fallback, and
  its accumulated position error against plain DEM below half that of free flight between the same
  checks.
"""
import copy

import numpy as np
import pytest

from exadem_amd import inputs, surrogate
from exadem_amd.sim import PolySim, SphereSim, scheme
import pyoracle as po

from helpers import by_id, force, interaction_table, pos, rel_err, vec, vel

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-10


@pytest.fixture(scope="module")
def spheres_1m():
    return inputs.polydisperse(n=1_000_000)


def _streams_equal(s, o):
    a, b, code, own = s.ctx.streams()
    oa, ob, ocode = o.streams()
    assert len(a) == len(oa) and len(a) > 5_000_000
    g_ord = np.argsort(a, kind="stable")
    o_ord = np.argsort(oa, kind="stable")
    assert np.array_equal(a[g_ord], oa[o_ord])
    assert np.array_equal(b[g_ord], ob[o_ord])
    assert np.array_equal(code[g_ord], ocode[o_ord])
    assert np.array_equal(own.astype(bool), (a < b) | (code != 13))


def _interactions_equal(s, o):
    kg, fg = interaction_table(*s.ctx.interactions())
    oi, oj, ot, of, _ = o.interactions()
    ko, fo = interaction_table(oi, oj, ot, of)
    assert np.array_equal(kg, ko)
    assert np.array_equal(fg, fo)
    return int(np.count_nonzero(np.any(fo != 0, axis=1)))


def _state_close(g, q, floor, what, ptol=TRAJ_TOL):
    g, q = by_id(g), by_id(q)
    assert np.array_equal(g["id"], q["id"])
    L = max(1.0, float(np.max(np.abs(pos(q)))))
    dpos = float(np.max(np.abs(pos(g) - pos(q)))) / L
    dvel = float(np.max(np.abs(vel(g) - vel(q)))) / max(1e-3, float(np.max(np.abs(vel(q)))))
    ef = rel_err(force(g), force(q), floor)
    ms = np.maximum(np.linalg.norm(vec(q, "mom"), axis=1), 1e-6 * float(np.max(floor)))
    em = float(np.max(np.linalg.norm(vec(g, "mom") - vec(q, "mom"), axis=1) / ms))
    print(f"{what}: positions {dpos:.2e}, velocities {dvel:.2e}, forces {ef:.2e}, moments {em:.2e}", flush=True)
    assert dpos <= ptol and dvel <= ptol, what
    assert ef <= FORCE_TOL and em <= FORCE_TOL, what


def test_1m_polydisperse_spheres_match_oracle(gpu_ok, spheres_1m):
    w = spheres_1m
    sch = scheme(w)
    s = SphereSim(w)
    s.first_iteration()
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.first_iteration(sch)
    _streams_equal(s, o)
    act = _interactions_equal(s, o)
    floor = by_id(o.particles())["mass"] * 9.81
    _state_close(s.ctx.download(), o.particles(), floor, "1M spheres, first_iteration", ptol=0.0)
    st = s.run(20)
    o.run_steps(sch, 20)
    act20 = _interactions_equal(s, o)
    _state_close(s.ctx.download(), o.particles(), floor, "1M spheres, 20 engine steps")
    print(f"1M spheres: {act} active pairs after first_iteration, {act20} after 20 steps, "
          f"{st['rebuilds']} rebuilds", flush=True)
    assert act > 1_000_000


def _feature_table(ii, jj, tt, sw, si, sj, fr):
    order = np.argsort(ii, kind="stable")
    keys = np.stack([ii.astype(np.int64), jj.astype(np.int64), tt.astype(np.int64), sw.astype(np.int64),
                     si.astype(np.int64), sj.astype(np.int64)], axis=1)
    return keys[order], fr[order]


def test_200k_polyhedra_match_oracle(gpu_ok):
    """The polyhedra benchmark input (grown to solid fraction 0.45 and relaxed: no overlaps, every pair
    within rVerlet a candidate) bit-exact with the oracle after first_iteration and 20 benchmark steps;
    then, to bring many contacts of every type, the same particles with velocities N(0, 2) per component
    and dt 4e-4 (about 40 steps per contact time), 50 more engine steps with rebuilds, also bit-exact."""
    w = inputs.polyhedra_benchmark(n=200_000)
    assert abs(w.meta["solid_fraction"] - 0.45) < 1e-9 and w.meta["interactions"] > 1_000_000
    assert w.meta["min_dn"] == 0.0  # relaxed: no overlap left
    sch = scheme(w)
    s = PolySim(w)
    s.first_iteration()
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.set_shapes(w.shapes)
    o.first_iteration_poly(sch)
    gk, gf = _feature_table(*s.ctx.poly_interactions())
    ok, of = _feature_table(*o.poly_interactions())
    assert len(gk) == len(ok) and len(gk) > 500_000
    assert np.array_equal(gk, ok) and np.array_equal(gf, of)
    types = np.bincount(gk[:, 2], minlength=6)
    print(f"200k polyhedra: feature candidates by type {types.tolist()}", flush=True)
    assert np.all(types[:4] > 0)
    floor = by_id(o.particles())["mass"] * 9.81
    _state_close(s.ctx.download(), o.particles(), floor, "200k polyhedra, first_iteration", ptol=0.0)
    s.run(20)
    o.run_steps_poly(sch, 20)
    gk, gf = _feature_table(*s.ctx.poly_interactions())
    ok, of = _feature_table(*o.poly_interactions())
    assert np.array_equal(gk, ok) and np.array_equal(gf, of)
    _state_close(s.ctx.download(), o.particles(), floor, "200k polyhedra, 20 engine steps")
    # shaken: contacts of every type
    p = s.ctx.download()
    v = inputs.normal(77, 2, 3 * w.n, 0.0, 2.0).reshape(3, w.n)
    order = np.argsort(p["id"])
    for a, k in enumerate(("vx", "vy", "vz")):
        p[k] = np.empty(w.n)
        p[k][order] = v[a]
    w2 = copy.copy(w)
    w2.particles = p
    w2.dt = 4e-4
    sch2 = scheme(w2)
    s2 = PolySim(w2)
    s2.first_iteration()
    o2 = po.Oracle(p, w2.domain, w2.walls)
    o2.set_shapes(w2.shapes)
    o2.first_iteration_poly(sch2)
    st = s2.run(50)
    o2.run_steps_poly(sch2, 50)
    gk, gf = _feature_table(*s2.ctx.poly_interactions())
    ok, of = _feature_table(*o2.poly_interactions())
    assert np.array_equal(gk, ok) and np.array_equal(gf, of)
    act = np.any(gf != 0, axis=1)
    contacts = np.bincount(gk[act, 2], minlength=6)
    print(f"200k polyhedra shaken, 50 steps of 4e-4: {st['rebuilds']} rebuilds, contacts by type "
          f"{contacts.tolist()}", flush=True)
    assert np.all(contacts[1:4] > 0) and int(np.sum(contacts)) > 10_000
    _state_close(s2.ctx.download(), o2.particles(), floor, "200k polyhedra shaken, 50 engine steps")


def test_1m_surrogate_k10_accumulated_error(gpu_ok, spheres_1m):
    """The shipped K = 10 network on the 1M input: 100 surrogate steps with a DEM check every 20 run
    without fallback, and the accumulated position error against 1000 plain DEM steps from the same
    start is below half that of free flight between the same checks (max and RMS over R printed)."""
    params, scales = surrogate.load(surrogate.WEIGHTS)
    assert scales["substeps"] == 10
    r = surrogate.trajectory_errors(spheres_1m, params, scales, nsteps=100, check_every=20, tol=0.01)
    net, free = r["network"], r["free_flight"]
    print(f"1M surrogate K=10, 100 steps: network max {net['max']:.3e} R rms {net['rms']:.3e} R; "
          f"free flight max {free['max']:.3e} R rms {free['rms']:.3e} R; checks {net['stats']['checks']}, "
          f"fallbacks {net['stats']['fallbacks']}", flush=True)
    assert net["stats"]["fallbacks"] == 0 and net["stats"]["checks"] == 5
    assert net["rms"] < 0.5 * free["rms"]
    assert net["max"] < 0.01
