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
    w = inputs.polyhedra_benchmark(n=200_000)
    assert abs(w.meta["solid_fraction"] - 0.45) < 1e-9 and w.meta["active_interactions"] > 50_000
    assert w.meta["min_dn"] > -0.02  # no vertex deeper than the Minkowski shells
    sch = scheme(w)
    s = PolySim(w)
    s.first_iteration()
    o = po.Oracle(w.particles, w.domain, w.walls)
    o.set_shapes(w.shapes)
    o.first_iteration_poly(sch)
    gk, gf = _feature_table(*s.ctx.poly_interactions())
    ok, of = _feature_table(*o.poly_interactions())
    assert len(gk) == len(ok) and len(gk) > 100_000
    assert np.array_equal(gk, ok) and np.array_equal(gf, of)
    types = np.bincount(gk[:, 2], minlength=6)
    print(f"200k polyhedra: feature interactions by type {types.tolist()}", flush=True)
    assert types[0] + types[1] > 0 and types[2] > 0 and types[3] > 0 and types[5] > 0
    floor = by_id(o.particles())["mass"] * 9.81
    _state_close(s.ctx.download(), o.particles(), floor, "200k polyhedra, first_iteration", ptol=0.0)
    s.run(20)
    o.run_steps_poly(sch, 20)
    gk, gf = _feature_table(*s.ctx.poly_interactions())
    ok, of = _feature_table(*o.poly_interactions())
    assert np.array_equal(gk, ok) and np.array_equal(gf, of)
    _state_close(s.ctx.download(), o.particles(), floor, "200k polyhedra, 20 engine steps")


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
