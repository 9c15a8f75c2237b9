"""xdem_simulation_state_noalloc (the thermodynamic state with its block partials in the library's fixed
per-GPU array, no device allocation) against xdem_simulation_state, through the C ABI: the same result
bit for bit, on two calls in a row with the second on a smaller input, for spheres (periodic packing,
then a walled box) and polyhedra (larger then smaller); and a clear capacity error past its limit.
"""
import numpy as np
import pytest

from exadem_amd import inputs, native
from exadem_amd.sim import PolySim, SphereSim

from helpers import periodic_pack, wall_deposition

pytestmark = pytest.mark.gpu

MAX_PARTICLES = 5_032_704  # 19,659 blocks of 256 partials fit the fixed array with the results


def _same(a, b, what):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (what, k, a[k], b[k])


def test_noalloc_equals_simulation_state_spheres(gpu_ok):
    big = SphereSim(periodic_pack(n=40_000, seed=31))
    big.first_iteration()
    big.run(20)
    ref = big.ctx.simulation_state()
    got = big.ctx.simulation_state(noalloc=True)
    print(f"40k spheres: {got['n_interactions']} interactions, {got['n_active']} active", flush=True)
    _same(ref, got, "40k periodic")
    assert got["n_active"] > 0
    small = SphereSim(wall_deposition(n=600))
    small.first_iteration()
    small.run(10)
    ref2 = small.ctx.simulation_state()
    got2 = small.ctx.simulation_state(noalloc=True)
    _same(ref2, got2, "600 walled")
    assert got2["n_particles"] == 600 and got2["n_active"] > 0
    # the first context again after the smaller call: the shared array holds the last call's values
    _same(ref, big.ctx.simulation_state(noalloc=True), "40k periodic again")


def test_noalloc_equals_simulation_state_polyhedra(gpu_ok):
    for n, seed in ((1500, 5), (300, 6)):
        w = inputs.polyhedra(density=2500.0, n=n, spacing=1.5, jitter=0.05, seed=seed, dt=2e-4, sigma_v=0.5, z0=0.6,
                             periodic=(False, False, False))
        s = PolySim(w)
        s.first_iteration()
        s.run(5)
        ref = s.ctx.simulation_state()
        got = s.ctx.simulation_state(noalloc=True)
        _same(ref, got, f"{n} polyhedra")
        assert got["n_interactions"] > 0


def test_noalloc_rejects_more_than_its_limit(gpu_ok):
    n = MAX_PARTICLES + 1
    side = int(np.ceil(n ** (1.0 / 3.0)))
    k = np.arange(n)
    x, y, z = (k % side) * 2.0 + 1.0, ((k // side) % side) * 2.0 + 1.0, (k // (side * side)) * 2.0 + 1.0
    p = inputs.sphere_set(x, y, z, np.full(n, 0.5), 2500.0)
    dom = native.Domain()
    for a in range(3):
        dom.lo[a] = 0.0
        dom.hi[a] = 2.0 * side
        dom.periodic[a] = 0
    dom.cell_size = 2.0
    ctx = native.Context(0)
    ctx.set_domain(dom)
    ctx.upload(p)
    with pytest.raises(native.XdemError, match="CAPACITY"):
        ctx.simulation_state(noalloc=True)
    st = ctx.simulation_state()
    assert st["n_particles"] == n
    del p
    # the entry point keeps working after the capacity error
    small = SphereSim(periodic_pack(n=2000, seed=41))
    small.first_iteration()
    _same(small.ctx.simulation_state(), small.ctx.simulation_state(noalloc=True), "after the capacity error")
