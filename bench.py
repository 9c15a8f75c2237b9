#!/usr/bin/env python3
"""Benchmark of the sphere hot path on one MI355X: particle-steps/s at 1M polydisperse spheres.

Workload (default): 1,000,000 spheres, radius U[0.5, 1.0], periodic 200 x 200 x 89.25 box at solid
fraction 0.55 (RSA + growth + relaxation), Gaussian velocities sigma 0.1, Hooke kn=1e5, kt=0.8kn,
mu=0.3, damping 0.9, dt=5e-6, skin 0.1 with rebuild on a 0.05 displacement, gravity -9.81 z.

A step is one numerical_scheme_verlet_macro iteration (prolog, rebuild test and rebuild when it fires,
reset + gravity + contact + force_to_accel, epilog) run by the fused engine. The timed region covers
exactly --steps steps after --warmup untimed steps, with the particle state resident in HBM.
Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def bytes_model(form, n, pairs):
    """Algorithmic HBM bytes per step and for the engine's integrating force pass alone (kr = 0).

    Per step FP64: 412 B/particle + 52 B/pair (operator sequence: integrate 272 B/particle reads x, v,
    a, mom, w, Q, m, I and writes x, v, w, Q; force pass 140 B/particle reads x, v, w, R, m, writes f,
    mom, CSR offset; per half-list pair: neighbour index 4 B + friction history read/write 48 B).
    The engine runs a step as one integrating force pass: 300 B/particle + 52 B/pair (the force pass
    without its f and mom writes, plus Q, I and the rebuild reference read and the 128-byte record
    written). FP32-force: 356 B/particle + 28 B/pair per step; integrating pass 256 B/particle +
    28 B/pair (force pass 44 B read + 24 B write + 4 B offset)."""
    if form == "fp32":
        return 356 * n + 28 * pairs, 256 * n + 28 * pairs
    return 412 * n + 52 * pairs, 300 * n + 52 * pairs


def bytes_model_poly(n, entries, nfeat, contacts, nvert=7.0):
    """Algorithmic HBM bytes per step of the polyhedra path (FP64), per kernel: what each kernel must move
    at least once.

    k_poly_vertices, per particle: the record's position, spin/homothety and orientation (96 B), type (4 B),
    the rebuild reference read (position 24 B + orientation 32 B), the world vertices written (32 B each,
    nvert per particle on average: 8 for a cube, 6 for an octahedron).
    k_poly_feat, per particle: its record (128 B) and vertices (32 B each) read once; per feature: the
    classifier entry (16 B), the contact flag (1 B) and dn (8 B) written; per feature in contact: the
    friction history read and written (48 B) and its force and moment terms written (72 B).
    k_poly_gather, per particle: record (128 B), inertia (24 B), list length (4 B), the accelerations,
    moments and angular accelerations written (72 B); per list entry: word and feature range (12 B); per
    feature: the contact flag (1 B); per feature in contact: its terms read (72 B)."""
    verts = (96 + 4 + 56 + 32 * nvert) * n
    feat = (128 + 32 * nvert) * n + 25 * nfeat + 120 * contacts
    gather = (128 + 24 + 4 + 72) * n + 12 * entries + nfeat + 72 * contacts
    return {"k_poly_vertices": verts, "k_poly_feat": feat, "k_poly_gather": gather,
            "force_pass": feat + gather, "step": verts + feat + gather}


def profiled_traffic(stem):
    """HBM bytes per force launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE, scaled by a copy
    probe), read from the committed summary profiles/<stem>.md; None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{stem}.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--poll", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU-baseline sample length (oracle steps)")
    ap.add_argument("--relax-sweeps", type=int, default=150)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a device copy of 1 GiB (counter calibration for --pmc runs) and exit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-every", type=int, default=8,
                    help="HIP-event timing of the kernels of every k-th timed step (an event pair costs the stream "
                         "a few microseconds; 1 times every step)")
    ap.add_argument("--form", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP32-force and surrogate lines")
    ap.add_argument("--sur-steps", type=int, default=100, help="timed surrogate steps")
    ap.add_argument("--poly-n", type=int, default=200_000)
    ap.add_argument("--poly-steps", type=int, default=1000)
    ap.add_argument("--no-trajectory", action="store_true", help="skip the surrogate's accumulated-error runs")
    ap.add_argument("--poly-contacts", choices=["yes", "no", "only"], default="yes",
                    help="polyhedra: also time the same particles shaken into contact (yes), skip that (no), or "
                         "time only that (only; the relaxed line then carries no timing), e.g. for separate profiles")
    ap.add_argument("--poly-cache", default=None, help="npz file for the grown polyhedra input (read if it "
                                                         "holds the same input, else written)")
    ap.add_argument("--only", default="", help="comma-separated forms to run after the headline line "
                                                "(fp32, surrogate, polyhedra, mesh); default: all")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not provided "
                                   "(spatial slabs with a per-step halo exchange of particle records would shard it)"}))
        sys.exit(2)

    if args.copy_probe:
        import torch  # plumbing: a plain copy kernel of known size
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8,
                          "launches": 5}))
        return

    from exadem_amd import inputs, native
    from exadem_amd.sim import SphereSim, scheme

    t0 = time.time()
    w = inputs.polydisperse(n=args.n, relax_sweeps=args.relax_sweeps)
    t_gen = time.time() - t0
    log(f"generated {w.n} spheres in {t_gen:.1f} s (max overlap {w.meta['max_overlap']:.3g})")
    n = w.n

    def timed(form_name, mesh=None):
        form = native.XDEM_FORM_FP32 if form_name == "fp32" else native.XDEM_FORM_FP64
        sim = SphereSim(w, form=form, mesh=mesh)
        sim.first_iteration()
        sim.run(args.warmup, poll=args.poll)
        sim.ctx.synchronize()
        sim.ctx.set_timing(True, every=args.timing_every)
        sim.ctx.kernel_timing(0, reset=True)
        sim.ctx.kernel_timing(1, reset=True)
        sim.ctx.synchronize()
        t1 = time.perf_counter()
        st = sim.run(args.steps, poll=args.poll)
        sim.ctx.synchronize()
        elapsed = time.perf_counter() - t1
        sim.ctx.set_timing(False)
        f_ms, f_cnt = sim.ctx.kernel_timing(1)
        i_ms, _ = sim.ctx.kernel_timing(0)
        nb = sim.ctx.nbh_stats()
        log(f"{form_name}: timed {args.steps} steps in {elapsed:.3f} s (force {f_ms * 1e3:.1f} us, "
            f"integrate {i_ms * 1e3:.1f} us)")
        return sim, {"elapsed": elapsed, "f_ms": f_ms, "f_cnt": f_cnt, "i_ms": i_ms, "stats": st, "nbh": nb}

    # headline: the plain FP64 form
    sim64, r64 = timed(args.form)
    elapsed = r64["elapsed"]
    f_ms, f_cnt, i_ms, st, st_nb2 = r64["f_ms"], r64["f_cnt"], r64["i_ms"], r64["stats"], r64["nbh"]
    pairs = st_nb2["pairs"]
    step_bytes, force_bytes = bytes_model(args.form, n, pairs)
    value = n * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    achieved = force_bytes / (f_ms * 1e-3) / 1e9 if f_ms > 0 else 0.0
    traffic, traffic_src = profiled_traffic(f"k_force_{args.form}_1M_spheres")
    # one neighbour rebuild (move_particles, chunk_neighbors_contact, nbh_sphere, classify_interactions),
    # forced three times on the state the run left; wall time with the stream synchronised around it
    reb = []
    for _ in range(3):
        sim64.ctx.synchronize()
        tr = time.perf_counter()
        sim64.rebuild()
        sim64.ctx.synchronize()
        reb.append(1e3 * (time.perf_counter() - tr))
    rebuild_ms = float(np.median(reb))
    log(f"{args.form}: rebuild {rebuild_ms:.3f} ms (median of {[round(r, 3) for r in reb]})")
    del sim64

    forms = {}
    only = [k for k in args.only.split(",") if k]
    want = lambda k: not only or k in only
    if args.form == "fp64" and not args.no_forms and want("fp32"):
        # FP32-force form: throughput, and the error of its first force pass against the FP64 pass
        # from the same state (same contacts: the sign test is FP64)
        a = SphereSim(w, form=native.XDEM_FORM_FP64)
        a.first_iteration()
        b = SphereSim(w, form=native.XDEM_FORM_FP32)
        b.first_iteration()
        pa, pb = a.ctx.download(), b.ctx.download()
        oa, ob = np.argsort(pa["id"]), np.argsort(pb["id"])
        fa = np.stack([pa[k][oa] for k in ("fx", "fy", "fz")], 1)
        fb = np.stack([pb[k][ob] for k in ("fx", "fy", "fz")], 1)
        floor = pa["mass"][oa] * 9.81
        ferr = float(np.max(np.linalg.norm(fa - fb, axis=1) / np.maximum(np.linalg.norm(fa, axis=1), floor)))
        # stricter: against the contact part alone (F64 - m g), over particles whose contact force is at
        # least 1% of the largest; moments against the moment, same selection
        mg = pa["mass"][oa, None] * np.asarray(w.gravity)[None, :]
        fc = np.linalg.norm(fa - mg, axis=1)
        sel = fc >= 1e-2 * float(np.max(fc)) if fc.size and np.max(fc) > 0 else np.zeros(len(fc), bool)
        crel = np.linalg.norm(fa - fb, axis=1)[sel] / fc[sel]
        ma = np.stack([pa[k][oa] for k in ("momx", "momy", "momz")], 1)
        mb = np.stack([pb[k][ob] for k in ("momx", "momy", "momz")], 1)
        mn = np.linalg.norm(ma, axis=1)
        msel = mn >= 1e-2 * float(np.max(mn)) if mn.size and np.max(mn) > 0 else np.zeros(len(mn), bool)
        mrel = np.linalg.norm(ma - mb, axis=1)[msel] / mn[msel]
        ia, ib = a.ctx.interactions(), b.ctx.interactions()
        # contact set: the interactions whose friction history is non-zero after the first pass
        act_a, act_b = np.any(ia[3] != 0, axis=1), np.any(ib[3] != 0, axis=1)
        same_set = (len(ia[0]) == len(ib[0]) and all(np.array_equal(x, y) for x, y in zip(ia[:3], ib[:3]))
                    and np.array_equal(act_a, act_b))
        del a, b
        _, r32 = timed("fp32")
        v32 = n * args.steps / r32["elapsed"]
        forms["fp32"] = {"value": v32, "ms_per_step": 1e3 * r32["elapsed"] / args.steps,
                         "force_avg_ms": r32["f_ms"], "speedup_vs_fp64": v32 / value,
                         "max_force_rel_err": ferr, "contact_set_identical": bool(same_set),
                         "contacts": int(np.count_nonzero(act_a)),
                         "error_stated_as": "max over particles of |F32 - F64| / max(|F64|, m g), first force pass",
                         "contact_force_rel_err": {"max": float(np.max(crel)) if crel.size else None,
                                                   "rms": float(np.sqrt(np.mean(crel ** 2))) if crel.size else None,
                                                   "particles": int(crel.size)},
                         "moment_rel_err": {"max": float(np.max(mrel)) if mrel.size else None,
                                            "rms": float(np.sqrt(np.mean(mrel ** 2))) if mrel.size else None,
                                            "particles": int(mrel.size)},
                         "contact_error_stated_as": "|F32 - F64| / |F64 - m g| and |M32 - M64| / |M64| over the "
                                                    "particles whose contact force (moment) is at least 1% of "
                                                    "the largest, first force pass"}
    if args.form == "fp64" and not args.no_forms and want("surrogate"):
        # surrogate-accelerated forms: the trained network with K = 200, 100 and 10 DEM steps per step
        from exadem_amd import surrogate
        for form_key, wpath in (("surrogate", surrogate.WEIGHTS_K200), ("surrogate_k100", surrogate.WEIGHTS_K100),
                                ("surrogate_k10", surrogate.WEIGHTS)):
            if not os.path.exists(wpath):
                forms[form_key] = {"value": None, "status": "no trained weights at " + wpath}
                continue
            params, scales = surrogate.load(wpath)
            K = scales["substeps"]
            s = SphereSim(w)
            s.first_iteration()
            surrogate.settle(s, K)
            g = surrogate.attach(s.ctx, params, scales)
            sch = scheme(w)
            g.surrogate_run(sch, 20, check_every=20, tol=0.01)
            s.ctx.synchronize()
            t1 = time.perf_counter()
            sst = g.surrogate_run(sch, args.sur_steps, check_every=20, tol=0.01)
            s.ctx.synchronize()
            el = time.perf_counter() - t1
            # network against the DEM on two windows of K steps from the state the run left
            werr = [e for e in (surrogate.window_errors(s, g, K) for _ in range(2)) if e is not None]
            eff = n * args.sur_steps * K / el
            forms[form_key] = {"value": eff, "unit": "effective particle-steps/s (simulated time x n / dt)",
                               "ms_per_surrogate_step": 1e3 * el / args.sur_steps, "substeps": K,
                               "speedup_vs_fp64": eff / value,
                               "simulated_time_per_s": args.sur_steps * K * w.dt / el,
                               "checks": sst["checks"], "fallbacks": sst["fallbacks"],
                               "max_pos_err_over_radius": sst["max_err"],
                               "max_pos_err_over_radius_free_flight": sst["max_err_free"],
                               "rms_pos_err_over_radius": sst["rms_err"],
                               "rms_pos_err_over_radius_free_flight": sst["rms_err_free"],
                               "rms_contact_force_err": werr[-1][0] if werr else None,
                               "rms_contact_force_free_flight": werr[-1][1] if werr else None,
                               "error_stated_as": "largest and RMS |x_surrogate - x_DEM| / R over the drift checks "
                                                  "(one surrogate step against K DEM steps from the same state); "
                                                  "RMS error of the predicted mean contact force over a window, "
                                                  "units m0 a0, against zero contact force (free flight)"}
            if form_key == "surrogate_k10" and not args.no_trajectory:
                # accumulated error at the end of a drift-checked run against plain DEM over the same
                # simulated time from the same start, and the same for free flight between the same checks
                tr = surrogate.trajectory_errors(w, params, scales, nsteps=args.sur_steps, check_every=20, tol=0.01)
                forms[form_key]["accumulated_error_over_radius"] = {
                    "network": {"max": tr["network"]["max"], "rms": tr["network"]["rms"]},
                    "free_flight": {"max": tr["free_flight"]["max"], "rms": tr["free_flight"]["rms"]},
                    "dem_steps": args.sur_steps * K,
                    "stated_as": "|x - x_DEM| / R over all particles after the run, against plain FP64 DEM over "
                                 "the same simulated time from the same state"}
                log(f"{form_key}: accumulated error rms {tr['network']['rms']:.3g} R (free flight "
                    f"{tr['free_flight']['rms']:.3g} R), max {tr['network']['max']:.3g} R "
                    f"(free flight {tr['free_flight']['max']:.3g} R)")
            if K == 10:
                forms[form_key]["target_5x_at_k10"] = ("met" if eff / value >= 5.0 else
                                                       f"unmet: {eff / value:.2f}x; a K = 10 network pass costs "
                                                       f"about {1e3 * el / args.sur_steps / ms_step:.0f} DEM steps")
            log(f"{form_key}: {args.sur_steps} steps of K={K} in {el:.3f} s, checks {sst['checks']}, "
                f"max err {sst['max_err']:.3g} R (free flight {sst['max_err_free']:.3g} R), "
                f"rms {sst['rms_err']:.3g} R (free flight {sst['rms_err_free']:.3g} R), "
                f"force err {werr[-1][0] if werr else float('nan'):.3g} vs {werr[-1][1] if werr else float('nan'):.3g}")
            del g, s
    if args.form == "fp64" and not args.no_forms and want("polyhedra"):
        # polyhedra: cubes (edge 1) and octahedra (circumradius 0.8), Minkowski radius 0.02, density 1,
        # random orientations, grown on the GPU from a jittered lattice to solid fraction 0.45 of the box,
        # periodic in x and y with a floor and a lid, gravity, dt 2e-6
        from exadem_amd.sim import PolySim
        tg = time.perf_counter()
        wp = inputs.polyhedra_benchmark(n=args.poly_n, cache=args.poly_cache, log=log)
        gen_s = time.perf_counter() - tg
        log(f"polyhedra: grew {wp.n} to solid fraction {wp.meta['solid_fraction']:.3f} (bed "
            f"{wp.meta['bed_solid_fraction']:.3f}) in {gen_s:.1f} s")
        ps = PolySim(wp)
        ps.first_iteration()
        ps.run(20, poll=args.poll)
        nvert = float(np.mean([wp.shapes[t].nv for t in wp.particles["type"]]))
        if args.poly_contacts != "only":
            ps.ctx.synchronize()
            ps.ctx.set_timing(True, every=args.timing_every)
            ps.ctx.kernel_timing(0, reset=True)
            ps.ctx.kernel_timing(1, reset=True)
            t1 = time.perf_counter()
            pst = ps.run(args.poly_steps, poll=args.poll)
            ps.ctx.synchronize()
            el = time.perf_counter() - t1
            ps.ctx.set_timing(False)
            pf_ms, _ = ps.ctx.kernel_timing(1)
            pi_ms, _ = ps.ctx.kernel_timing(0)
            waves = ps.ctx.classifier_waves()
            nfeat = int(np.sum(waves[:4]))
            pnb = ps.ctx.nbh_stats()
            _, _, ptt, _, _, _, pfr = ps.ctx.poly_interactions()
            contacts = int(np.count_nonzero(np.any(pfr[ptt < 4] != 0, axis=1)))  # features with friction
            pbytes = bytes_model_poly(wp.n, 2 * pnb["pairs"], nfeat, contacts, nvert)  # both directions
            pach = pbytes["force_pass"] / (pf_ms * 1e-3) / 1e9 if pf_ms > 0 else 0.0
            reb = []
            for _ in range(3):
                ps.ctx.synchronize()
                tr = time.perf_counter()
                ps.rebuild()
                ps.ctx.synchronize()
                reb.append(1e3 * (time.perf_counter() - tr))
            forms["polyhedra"] = {"value": wp.n * args.poly_steps / el, "unit": "particle-steps/s",
                                  "n": wp.n, "steps": args.poly_steps, "ms_per_step": 1e3 * el / args.poly_steps,
                                  "force_avg_ms": pf_ms, "integrate_vertices_avg_ms": pi_ms,
                                  "feature_interactions": nfeat, "features_in_contact": contacts,
                                  "classifier_waves": {"VV": int(waves[0]), "VE": int(waves[1]), "VF": int(waves[2]),
                                                       "EE": int(waves[3]), "vertex_wall": int(waves[5])},
                                  "rebuilds": pst["rebuilds"], "rebuild_ms": float(np.median(reb)),
                                  "solid_fraction": wp.meta["solid_fraction"],
                                  "bed_solid_fraction": wp.meta["bed_solid_fraction"],
                                  "generation_s": gen_s,
                                  "roofline": {"bound": "hbm", "achieved": pach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                               "frac": pach / PEAK_HBM_GBS,
                                               "traffic": profiled_traffic("poly_force_200k_polyhedra")[0],
                                               "traffic_source": profiled_traffic("poly_force_200k_polyhedra")[1],
                                               "kernel": "force pass: k_poly_feat + k_poly_gather",
                                               "bytes_per_launch": pbytes["force_pass"], "avg_ms": pf_ms,
                                               "step_bytes": pbytes["step"],
                                               "step_frac": pbytes["step"] / (el / args.poly_steps) / 1e9 / PEAK_HBM_GBS},
                                  "workload": wp.meta.get("workload", "cubes + octahedra 50/50, random orientations, "
                                                                      "jittered lattice spacing 1.55, periodic x/y, "
                                                                      "floor and lid, dt 2e-6")}
            log(f"polyhedra: {args.poly_steps} steps of {wp.n} in {el:.3f} s (force {pf_ms * 1e3:.1f} us, "
                f"rebuild {np.median(reb):.2f} ms)")
        else:
            forms["polyhedra"] = {"n": wp.n, "solid_fraction": wp.meta["solid_fraction"]}
        if args.poly_contacts != "no":
            # The relaxed bed has no contact within the configured 1000 steps (gravity moves a body by ~2e-5 in
            # 2 ms), so the contact law is measured on the same particles shaken into contact: velocities
            # N(0, 2) per component, 50 steps of dt 4e-4 (rebuilds included), then the configured dt again.
            import copy
            p = ps.ctx.download()
            del ps
            vsh = inputs.normal(77, 2, 3 * wp.n, 0.0, 2.0).reshape(3, wp.n)
            order = np.argsort(p["id"])
            for a_, k_ in enumerate(("vx", "vy", "vz")):
                p[k_] = np.empty(wp.n)
                p[k_][order] = vsh[a_]
            wsh = copy.copy(wp)
            wsh.particles = p
            wsh.dt = 4e-4
            ss = PolySim(wsh)
            ss.first_iteration()
            ss.run(50, poll=args.poll)
            wsh.dt = wp.dt
            ss.run(20, poll=args.poll)
            ss.ctx.synchronize()
            ss.ctx.set_timing(True, every=args.timing_every)
            ss.ctx.kernel_timing(1, reset=True)
            t1 = time.perf_counter()
            sst_ = ss.run(args.poly_steps, poll=args.poll)
            ss.ctx.synchronize()
            el_s = time.perf_counter() - t1
            ss.ctx.set_timing(False)
            sf_ms, _ = ss.ctx.kernel_timing(1)
            swaves = ss.ctx.classifier_waves()
            snfeat = int(np.sum(swaves[:4]))
            snb = ss.ctx.nbh_stats()
            _, _, stt, _, _, _, sfr = ss.ctx.poly_interactions()
            sact = np.any(sfr != 0, axis=1)
            scont = int(np.count_nonzero(sact & (stt < 4)))
            sbytes = bytes_model_poly(wp.n, 2 * snb["pairs"], snfeat, scont, nvert)
            sach = sbytes["force_pass"] / (sf_ms * 1e-3) / 1e9 if sf_ms > 0 else 0.0
            forms["polyhedra"]["with_contacts"] = {
                "value": wp.n * args.poly_steps / el_s, "unit": "particle-steps/s",
                "ms_per_step": 1e3 * el_s / args.poly_steps, "force_avg_ms": sf_ms, "rebuilds": sst_["rebuilds"],
                "feature_interactions": snfeat,
                "features_in_contact": {n_: int(np.count_nonzero(sact & (stt == t_)))
                                        for t_, n_ in enumerate(("VV", "VE", "VF", "EE"))},
                "vertex_wall_in_contact": int(np.count_nonzero(sact & (stt == 5))),
                "roofline": {"bound": "hbm", "achieved": sach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": sach / PEAK_HBM_GBS, "bytes_per_launch": sbytes["force_pass"],
                             "traffic": profiled_traffic("poly_force_200k_polyhedra_contacts")[0],
                             "traffic_source": profiled_traffic("poly_force_200k_polyhedra_contacts")[1],
                             "kernel": "force pass: k_poly_feat + k_poly_gather", "avg_ms": sf_ms},
                "workload": "the polyhedra input above with velocities N(0, 2) per component, 50 steps of dt 4e-4 "
                            f"to bring contacts, then {args.poly_steps} timed steps of dt {wp.dt:g}"}
            log(f"polyhedra with contacts: {scont} features in contact, {args.poly_steps} steps in {el_s:.3f} s "
                f"(force {sf_ms * 1e3:.1f} us, {sst_['rebuilds']} rebuilds)")
            del ss
        else:
            del ps
        if not args.no_cpu and args.poly_contacts != "only":
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as po
            op = po.Oracle(wp.particles, wp.domain, wp.walls)
            op.set_shapes(wp.shapes)
            schp = scheme(wp)
            op.first_iteration_poly(schp)
            steps_p = 0
            tc = time.perf_counter()
            while True:
                op.run_steps_poly(schp, 1)
                steps_p += 1
                tcp = time.perf_counter() - tc
                if tcp >= args.cpu_seconds:
                    break
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            forms["polyhedra"]["cpu_baseline"] = {
                "value": wp.n * steps_p / tcp, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                "sample": f"CPU oracle (FP64 restatement of the reference polyhedra path, OpenMP) on the same "
                          f"{wp.n}-polyhedron input: {steps_p} steps after first_iteration, {tcp:.1f} s"}
            log(f"polyhedra cpu baseline: {wp.n * steps_p / tcp:.3g} particle-steps/s")
            del op
    if args.form == "fp64" and not args.no_forms and want("mesh"):
        # STL mesh driver: a stationary triangulated plane through the box at mid-height (grid spacing
        # about 2, two triangles per cell, Minkowski radius 0.01) in the same sphere input, FP64 form
        from exadem_amd import stl
        lo, hi = np.array(w.domain.lo, float), np.array(w.domain.hi, float)
        zc = lo[2] + 0.5 * (hi[2] - lo[2])
        gx = np.linspace(lo[0], hi[0], int((hi[0] - lo[0]) // 2) + 1)
        gy = np.linspace(lo[1], hi[1], int((hi[1] - lo[1]) // 2) + 1)
        tri = []
        for i in range(len(gx) - 1):
            for j in range(len(gy) - 1):
                a, b = (gx[i], gy[j], zc), (gx[i + 1], gy[j], zc)
                c, d = (gx[i + 1], gy[j + 1], zc), (gx[i], gy[j + 1], zc)
                tri += [[a, b, c], [a, c, d]]
        mv, me, mf = stl.build_shape(np.array(tri))
        mesh = {"mesh_id": 0, "vertices": mv, "edges": me, "faces": mf, "minkowski": 0.01}
        msim, rmesh = timed("fp64", mesh=mesh)
        mi, _, mt, mfr, _ = msim.ctx.mesh_interactions()
        vm = n * args.steps / rmesh["elapsed"]
        forms["mesh"] = {"value": vm, "unit": "particle-steps/s", "ms_per_step": 1e3 * rmesh["elapsed"] / args.steps,
                         "force_avg_ms": rmesh["f_ms"], "fp64_without_mesh_ms_per_step": ms_step,
                         "mesh_triangles": int(len(mf)), "mesh_candidates": int(len(mi)),
                         "mesh_contacts": int(np.count_nonzero(np.any(mfr != 0, axis=1))),
                         "spheres_with_candidates": int(len(np.unique(mi))),
                         "workload": "the FP64 input with a stationary triangulated plane at mid-height "
                                     "(spacing ~2, Minkowski radius 0.01); mesh contacts evaluated one "
                                     "candidate per thread in a pass of their own, whose terms the force "
                                     "pass adds in candidate order after the walls"}
        log(f"mesh: {len(mf)} triangles, {len(mi)} candidates on {len(np.unique(mi))} spheres: timed "
            f"{args.steps} steps in {rmesh['elapsed']:.3f} s (force {rmesh['f_ms'] * 1e3:.1f} us)")
        del msim

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        o = po.Oracle(w.particles, w.domain, w.walls)
        sch = scheme(w)
        o.first_iteration(sch)
        log("cpu baseline: oracle first_iteration done")
        steps = 0
        tc = time.perf_counter()
        while True:
            o.run_steps(sch, 2)
            steps += 2
            tcpu = time.perf_counter() - tc
            if tcpu >= args.cpu_seconds:
                break
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": n * steps / tcpu, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (FP64 restatement of the reference sphere path, OpenMP) on the same "
                         f"{n}-sphere input: {steps} steps after first_iteration, {tcpu:.1f} s; "
                         f"{cpu_model()}"}

    out = {
        "metric": "particle-steps/s at 1M spheres: FP64 vs FP32-force vs surrogate, at stated error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic (seeded RSA + relaxation packing, generated in-run)",
        "config": {"workload": f"polydisperse spheres {n}, phi 0.55, periodic {w.meta['box'][0]:.1f}x"
                               f"{w.meta['box'][1]:.1f}x{w.meta['box'][2]:.2f}, dt 5e-6, form {args.form}",
                   "n": n, "pairs": pairs, "rebuilds": st["rebuilds"], "polls": st["polls"],
                   "rebuild_ms": rebuild_ms,
                   "rebuild_ms_per_step_if_every_100_steps": rebuild_ms / 100.0,
                   "list_slots": st_nb2["slots"], "max_initial_overlap": w.meta["max_overlap"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": f"k_force_{args.form} (integrating force pass: force + kick + next prolog)", "avg_ms": f_ms, "launches": f_cnt,
                     "timed_launches": f"HIP events around the launches of every {args.timing_every}-th timed step",
                     "bytes_per_launch": force_bytes, "integrate_avg_ms": i_ms,
                     "step_bytes": step_bytes, "step_frac": step_bytes / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "cpu_baseline": cpu,
        "forms": forms,
        "setup_s": {"generate": t_gen},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
