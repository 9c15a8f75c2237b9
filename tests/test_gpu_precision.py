"""FP32-force form against the FP64 form on the precision input: 4,000,000 spheres with lognormal radii
(median 0.6, sigma 0.2, clipped to [0.3, 1.2]) in a periodic cube at solid fraction 0.58.

Checked, and printed:
- the first force pass from the same state: force error against the contact force and moment error
  against the moment, max and RMS over the particles whose contact force (moment) is at least 1 % of
  the largest; bar 1e-5 on the max;
- the contact set after the first pass and after 500 steps: identical;
- after 500 steps of each form: position difference over the radius (max and RMS) and the
  acceleration and moment difference of the two runs, which then also carry the divergence of the
  trajectories and of the FP32 friction history.
The 500-step bars follow from the force bar: a relative acceleration error of 1e-5 at the largest
accelerations of this input (|a| <= ~50 m/s^2 from the initial overlaps) moves a particle by at most
1e-5 * 50 * t^2 / 2 ~ 1.6e-9 over t = 2.5e-3 s, i.e. ~5e-9 R at the smallest radius; the bar is 1e-7 R.
"""
import time

import numpy as np
import pytest

from exadem_amd import inputs, native
from exadem_amd.sim import SphereSim
from helpers import by_id, force, interaction_table, pos, vec

pytestmark = pytest.mark.gpu

N = 4_000_000
STEPS = 500


def contact_errors(pa, pb, g, accel=False):
    """max / RMS of |F32 - F64| / |F64 - m g| and |M32 - M64| / |M64| over the particles whose contact
    force (moment) is at least 1 % of the largest. accel: the force fields hold accelerations (the state
    after a run, which ends with force_to_accel), so the contact part is a - g."""
    fa, fb = force(pa), force(pb)
    gm = np.asarray(g)[None, :] * (1.0 if accel else pa["mass"][:, None])
    fc = np.linalg.norm(fa - gm, axis=1)
    sel = fc >= 1e-2 * float(np.max(fc))
    ef = np.linalg.norm(fa - fb, axis=1)[sel] / fc[sel]
    ma, mb = vec(pa, "mom"), vec(pb, "mom")
    mn = np.linalg.norm(ma, axis=1)
    msel = mn >= 1e-2 * float(np.max(mn))
    em = np.linalg.norm(ma - mb, axis=1)[msel] / mn[msel]
    return {"force_max": float(np.max(ef)), "force_rms": float(np.sqrt(np.mean(ef ** 2))), "force_n": int(sel.sum()),
            "moment_max": float(np.max(em)), "moment_rms": float(np.sqrt(np.mean(em ** 2))),
            "moment_n": int(msel.sum())}


def contact_keys(sim):
    ii, jj, tt, fr = sim.ctx.interactions()
    keys, hist = interaction_table(ii, jj, tt, fr)
    return keys[np.any(hist != 0, axis=1)]


def test_fp32_against_fp64_at_4m(gpu_ok):
    t0 = time.perf_counter()
    w = inputs.precision_spheres(n=N)
    print(f"\n[precision] {N} spheres, box {w.meta['box'][0]:.2f}, phi 0.58, max initial overlap "
          f"{w.meta['max_overlap']:.3g}, generated in {time.perf_counter() - t0:.1f} s", flush=True)
    a = SphereSim(w, form=native.XDEM_FORM_FP64)
    a.first_iteration()
    b = SphereSim(w, form=native.XDEM_FORM_FP32)
    b.first_iteration()
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    e0 = contact_errors(pa, pb, w.gravity)
    print(f"[precision] first pass: force max {e0['force_max']:.3g} rms {e0['force_rms']:.3g} over {e0['force_n']} "
          f"particles; moment max {e0['moment_max']:.3g} rms {e0['moment_rms']:.3g} over {e0['moment_n']}", flush=True)
    assert e0["force_max"] <= 1e-5, e0
    assert e0["moment_max"] <= 1e-5, e0
    ka, kb = contact_keys(a), contact_keys(b)
    assert ka.shape[0] > 0 and np.array_equal(ka, kb)
    print(f"[precision] contact set identical: {ka.shape[0]} contacts", flush=True)
    del pa, pb
    times = []
    for sim in (a, b):
        sim.ctx.synchronize()
        t1 = time.perf_counter()
        sim.run(STEPS)
        sim.ctx.synchronize()
        times.append(time.perf_counter() - t1)
    print(f"[precision] {STEPS} steps: FP64 {1e3 * times[0] / STEPS:.3f} ms/step, FP32 {1e3 * times[1] / STEPS:.3f} "
          f"ms/step ({times[0] / times[1]:.2f}x)", flush=True)
    ka, kb = contact_keys(a), contact_keys(b)
    assert np.array_equal(ka, kb)
    pa, pb = by_id(a.ctx.download()), by_id(b.ctx.download())
    dx = np.linalg.norm(pos(pa) - pos(pb), axis=1) / pa["radius"]
    e1 = contact_errors(pa, pb, w.gravity, accel=True)
    print(f"[precision] after {STEPS} steps: position difference / R max {np.max(dx):.3g} rms "
          f"{np.sqrt(np.mean(dx ** 2)):.3g}; acceleration max {e1['force_max']:.3g} rms {e1['force_rms']:.3g}; moment max "
          f"{e1['moment_max']:.3g} rms {e1['moment_rms']:.3g}; contacts {ka.shape[0]}", flush=True)
    assert float(np.max(dx)) <= 1e-7
    # accelerations still within the force bar; moments also carry 500 steps of FP32 friction history
    assert e1["force_max"] <= 1e-5, e1
    assert e1["moment_max"] <= 1e-4, e1
