"""Learned surrogate of K DEM steps: training on trajectories of the DEM engine, and the
surrogate-accelerated run with DEM drift checks (host orchestration of the xdem_gnn_* entry points).

Training follows the learned-surrogate workload: polydisperse packings generated like the 1M-sphere
benchmark input (radius uniform in [0.5, 1], solid fraction 0.55, periodic, Gaussian velocities of
sigma 0.1), one DEM run per seed, every window of K = 10 DEM steps a sample: the network sees the
state at the start of the window and learns the mean contact force and torque over the window (in
units of m0 a0 and m0 a0 r0; the surrogate step divides by each particle's mass and inertia).
Forward, backward and the Adam update are the library's HIP kernels.
"""
import math
import os
import time

import numpy as np

from . import inputs, native
from .sim import SphereSim, scheme

HIDDEN, LAYERS, FN, FE, FO = 128, 3, 8, 16, 6
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "surrogate_polydisperse.npz")
# the same network trained on windows of K = 100 DEM steps (fine-tuned from the K = 10 weights)
WEIGHTS_K100 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "surrogate_polydisperse_k100.npz")
# K = 200 DEM steps per surrogate step (fine-tuned from the K = 100 weights)
WEIGHTS_K200 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "surrogate_polydisperse_k200.npz")


def param_shapes():
    out = [("W_in", (HIDDEN, FN)), ("b_in", (HIDDEN,))]
    for l in range(LAYERS):
        out += [(f"W_AB{l}", (2 * HIDDEN, HIDDEN)), (f"W_E{l}", (HIDDEN, FE)), (f"b_E{l}", (HIDDEN,)),
                (f"W_U{l}", (HIDDEN, 2 * HIDDEN)), (f"b_U{l}", (HIDDEN,))]
    out += [("W_out", (FO, HIDDEN)), ("b_out", (FO,))]
    return out


def init_params(seed):
    """He-normal weights (std sqrt(2 / fan_in)), zero biases, output layer scaled by 0.1."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in param_shapes():
        if name.startswith("b_"):
            parts.append(np.zeros(shape).ravel())
        else:
            std = math.sqrt(2.0 / shape[1]) * (0.1 if name == "W_out" else 1.0)
            parts.append(rng.normal(0.0, std, shape).ravel())
    flat = np.concatenate(parts).astype(np.float32)
    assert flat.size == native.GNN_NPARAMS
    return flat


def training_workload(n, seed, sigma_v=0.1):
    return inputs.polydisperse(n=n, seed_radius=1000 + seed, seed_vel=2000 + seed, seed_pos=3000 + seed,
                               sigma_v=sigma_v)


def settle(sim, substeps):
    """first_iteration stores forces without force_to_accel, so the first DEM step after it uses the
    force as an acceleration (the reference's order of operators). That window is not a sample: run
    it before sampling."""
    sim.run(substeps)


def contact_forces(p0, p1, T, g):
    """Mean contact force over a window from two states in the same device order: m (dv/T - g)."""
    return np.stack([p1["mass"] * ((p1[f"v{c}"] - p0[f"v{c}"]) / T - g[k]) for k, c in enumerate("xyz")], 1)


def measure_a0(sim, substeps):
    """Output scales from one window of K DEM steps: m0 = mean mass, a0 = RMS of the mean contact
    force over m0."""
    p0 = sim.ctx.download()
    sim.run(substeps)
    p1 = sim.ctx.download()
    m0 = float(np.mean(p1["mass"]))
    F = contact_forces(p0, p1, sim.w.dt * substeps, np.asarray(sim.w.gravity))
    return (float(np.sqrt(np.mean(F * F))) / m0) or 1.0, m0


def train(n=50_000, runs=64, steps_per_run=20_000, substeps=10, lr=1e-3, lr_end=1e-4, seed=0, time_budget=None,
          log=print, params=None, scales=None):
    """Trains the network on `runs` DEM runs of n spheres. Returns (params, scales, history)."""
    w = training_workload(n, seed)
    sim = SphereSim(w)
    sim.first_iteration()
    settle(sim, substeps)
    if scales is None:
        r0 = float(np.mean(w.particles["radius"]))
        a0, m0 = measure_a0(sim, substeps)
        scales = {"r0": r0, "v0": 0.1, "a0": a0, "m0": m0, "ov_scale": 50.0, "substeps": substeps}
    gnn = attach(sim.ctx, init_params(seed) if params is None else params, scales, train=True)
    windows = steps_per_run // substeps
    total = runs * windows
    history = []
    t0 = time.perf_counter()
    done = 0
    for r in range(runs):
        if r > 0:
            w = training_workload(n, seed + r)
            sim.w = w
            sim.ctx.set_domain(w.domain)
            sim.ctx.upload(w.particles)
            sim.first_iteration()
            settle(sim, substeps)
        sch = scheme(w)
        losses = []
        for k in range(windows):
            frac = done / max(1, total - 1)
            lr_k = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * frac))
            loss, used = gnn.train_window(sch, lr_k)
            done += 1
            if used:
                losses.append(loss)
        history.append({"run": r, "loss": float(np.mean(losses)) if losses else None, "samples": len(losses),
                        "elapsed_s": time.perf_counter() - t0})
        log(f"[surrogate] run {r + 1}/{runs}: mean loss {history[-1]['loss']}, samples {len(losses)}, "
            f"{history[-1]['elapsed_s']:.1f} s")
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            log(f"[surrogate] time budget reached after {r + 1} runs")
            break
    return gnn.get_params(), scales, history


def save(path, params, scales, history=()):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=params.astype(np.float32),
             scales=np.array([scales["r0"], scales["v0"], scales["a0"], scales["m0"], scales["ov_scale"],
                              scales["substeps"]]),
             loss=np.array([h["loss"] if h["loss"] is not None else np.nan for h in history], np.float64))


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        s = z["scales"]
        scales = {"r0": float(s[0]), "v0": float(s[1]), "a0": float(s[2]), "m0": float(s[3]),
                  "ov_scale": float(s[4]), "substeps": int(s[5])}
        return z["params"].astype(np.float32), scales


def attach(ctx, params, scales, train=False):
    g = native.Gnn(ctx, scales["r0"], scales["v0"], scales["a0"], scales["m0"], scales["ov_scale"],
                   scales["substeps"], train)
    g.set_params(params)
    return g


def window_errors(sim, gnn, substeps):
    """One window from the current state: the network's prediction of the mean contact force, the
    method without it (zero contact force: free flight under gravity) and the K DEM steps that follow.
    Returns the RMS force errors of both in units of m0 a0 (None if a rebuild reordered the particles),
    and advances the DEM by K steps."""
    pred = gnn.forward()
    p0 = sim.ctx.download()
    st = sim.run(substeps)
    p1 = sim.ctx.download()
    if st["rebuilds"]:
        return None
    F = contact_forces(p0, p1, sim.w.dt * substeps, np.asarray(sim.w.gravity)) / (gnn.cfg.m0 * gnn.cfg.a0)
    err = float(np.sqrt(np.mean((pred[:, :3] - F) ** 2)))
    base = float(np.sqrt(np.mean(F ** 2)))
    return err, base


def _min_image(d, L, periodic):
    for a in range(3):
        if periodic[a]:
            d[:, a] -= L[a] * np.round(d[:, a] / L[a])
    return d


def trajectory_errors(w, params, scales, nsteps, check_every=20, tol=0.01):
    """Accumulated position error of the drift-checked surrogate run against plain DEM over the same
    simulated time from the same start.

    Three runs of workload w start from the same state (first_iteration, then one settling window of K
    DEM steps): plain DEM for nsteps * K steps; the surrogate-accelerated run (xdem_surrogate_run, a DEM
    check every check_every steps with tolerance tol); and the same run with a network of zero
    parameters, whose output is zero contact force, i.e. free flight under gravity between the same DEM
    checks (the method without the network). Returns a dict with max and RMS |x - x_DEM| / R over the
    particles at the end, for the network and for free flight, and both runs' statistics."""
    K = scales["substeps"]
    sch = scheme(w)
    L = [w.domain.hi[a] - w.domain.lo[a] for a in range(3)]
    per = [bool(w.domain.periodic[a]) for a in range(3)]

    def final_positions(kind):
        s = SphereSim(w)
        s.first_iteration()
        settle(s, K)
        st = None
        if kind == "dem":
            s.run(nsteps * K)
        else:
            g = attach(s.ctx, params if kind == "net" else np.zeros_like(params), scales)
            st = g.surrogate_run(sch, nsteps, check_every=check_every, tol=tol)
            del g
        p = s.ctx.download()
        o = np.argsort(p["id"])
        x = np.stack([p["rx"][o], p["ry"][o], p["rz"][o]], 1)
        r = p["radius"][o]
        del s
        return x, r, st

    xd, r, _ = final_positions("dem")
    out = {"surrogate_steps": nsteps, "substeps": K, "check_every": check_every, "tol": tol}
    for kind in ("net", "free"):
        x, _, st = final_positions(kind)
        e = np.linalg.norm(_min_image(x - xd, L, per), axis=1) / r
        tag = "network" if kind == "net" else "free_flight"
        out[tag] = {"max": float(np.max(e)), "rms": float(np.sqrt(np.mean(e * e))), "stats": st}
    return out


def main():
    import argparse
    ap = argparse.ArgumentParser(description="train the learned DEM surrogate on GPU-DEM trajectories")
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--steps", type=int, default=20_000)
    ap.add_argument("--substeps", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--time-budget", type=float, default=None)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--init", default=None, help="start from the parameters of this weights file")
    args = ap.parse_args()
    init = load(args.init)[0] if args.init else None
    params, scales, hist = train(args.n, args.runs, args.steps, args.substeps, args.lr, seed=args.seed,
                                 time_budget=args.time_budget, params=init)
    save(args.out, params, scales, hist)
    print(f"[surrogate] saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
