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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def bytes_model(form, n, pairs):
    """Algorithmic HBM bytes per step and for the force pass alone (kr = 0).

    Per step FP64: 412 B/particle + 52 B/pair (integrate 272 B/particle: reads x, v, a, mom, w, Q, m, I,
    writes x, v, w, Q; force pass 140 B/particle: reads x, v, w, R, m, writes f, mom, CSR offset; per
    half-list pair: neighbour index 4 B + friction history read/write 48 B).
    FP32-force: 356 B/particle + 28 B/pair (force pass 44 B read + 24 B write + 4 B offset)."""
    if form == "fp32":
        return 356 * n + 28 * pairs, 72 * n + 28 * pairs
    return 412 * n + 52 * pairs, 140 * n + 52 * pairs


def profiled_traffic(form):
    """HBM bytes per force launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE, scaled by a copy
    probe), read from the committed summary for this form; None when there is none."""
    path = os.path.join(ROOT, "profiles", f"k_force_{form}_1M_spheres.md")
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
    ap.add_argument("--form", choices=["fp64", "fp32"], default="fp64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not provided "
                                   "(the sharding design is described in DESIGN.md)"}))
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
    form = native.XDEM_FORM_FP32 if args.form == "fp32" else native.XDEM_FORM_FP64
    sim = SphereSim(w, form=form)
    sim.first_iteration()
    sim.run(args.warmup, poll=args.poll)
    sim.ctx.synchronize()
    st_nb = sim.ctx.nbh_stats()
    sim.ctx.set_timing(True)
    sim.ctx.kernel_timing(0, reset=True)
    sim.ctx.kernel_timing(1, reset=True)
    sim.ctx.synchronize()
    t1 = time.perf_counter()
    st = sim.run(args.steps, poll=args.poll)
    sim.ctx.synchronize()
    t2 = time.perf_counter()
    sim.ctx.set_timing(False)
    elapsed = t2 - t1
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    f_ms, f_cnt = sim.ctx.kernel_timing(1)
    i_ms, i_cnt = sim.ctx.kernel_timing(0)
    st_nb2 = sim.ctx.nbh_stats()
    pairs = st_nb2["pairs"]
    n = w.n
    step_bytes, force_bytes = bytes_model(args.form, n, pairs)
    value = n * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    achieved = force_bytes / (f_ms * 1e-3) / 1e9 if f_ms > 0 else 0.0
    traffic, traffic_src = profiled_traffic(args.form)

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
                   "list_slots": st_nb2["slots"], "max_initial_overlap": w.meta["max_overlap"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_force_fp64 (fused force pass)", "avg_ms": f_ms, "launches": f_cnt,
                     "bytes_per_launch": force_bytes, "integrate_avg_ms": i_ms,
                     "step_bytes": step_bytes, "step_frac": step_bytes / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "cpu_baseline": cpu,
        "setup_s": {"generate": t_gen},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
