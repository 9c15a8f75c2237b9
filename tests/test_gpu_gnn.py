"""The learned surrogate's HIP kernels against the CPU oracle (gnn_oracle), and what training buys.

Bars: edge and node inputs within bfloat16 rounding of a float64 evaluation of their formulas;
forward outputs within 1e-2 of the output scale (max) and 1e-3 (mean) of the oracle run on the same
inputs with the same bfloat16 storage points; every parameter gradient within 2e-2 (relative
Frobenius norm) of the oracle's backward; training lowers the loss; the trained network predicts the
mean contact acceleration of the next window better than the method without it (zero contact
acceleration, i.e. free flight under gravity).
"""
import numpy as np
import pytest

import gnn_oracle as G
from exadem_amd import native, surrogate
from exadem_amd.sim import SphereSim, scheme

from helpers import periodic_pack

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    flat = G.init_params(seed) + rng.normal(0, 0.05, G.n_params())
    return flat.astype(np.float32)


@pytest.fixture(scope="module")
def net(gpu_ok):
    w = periodic_pack(n=1500, sigma_v=0.3)
    sim = SphereSim(w)
    sim.first_iteration()
    sim.run(20)
    flat = _params(3)
    gnn = native.Gnn(sim.ctx, r0=0.75, v0=0.1, a0=1.0, m0=4000.0, ov_scale=50.0, substeps=10, train=True)
    gnn.set_params(flat)
    return sim, gnn, flat


def test_params_roundtrip(net):
    _, gnn, flat = net
    assert np.array_equal(gnn.get_params(), flat)


def test_inputs_match_formulas(net):
    sim, gnn, _ = net
    gnn.forward()
    src, dst, e, x = gnn.dump_graph()
    p = sim.ctx.download()
    r = np.stack([p["rx"], p["ry"], p["rz"]], 1)
    v = np.stack([p["vx"], p["vy"], p["vz"]], 1)
    om = np.stack([p["vrotx"], p["vroty"], p["vrotz"]], 1)
    R = p["radius"]
    box = np.array(sim.w.meta["box"])
    d = r[dst] - r[src]
    d -= box * np.round(d / box)
    dist = np.linalg.norm(d, axis=1)
    Rs = R[src] + R[dst]
    nn = d / dist[:, None]
    vr = ((v[dst] + np.cross(om[dst], -R[dst, None] * nn)) - (v[src] + np.cross(om[src], R[src, None] * nn))) / 0.1
    o = (Rs - dist) / Rs * 50.0
    op = np.maximum(o, 0.0)
    cf = (o > 0).astype(float)
    vn = cf * np.sum(vr * nn, axis=1)
    ref = np.concatenate([nn, o[:, None], vr, ((R[dst] - R[src]) / Rs * 4.0)[:, None], op[:, None] * nn,
                          vn[:, None] * nn, op[:, None], cf[:, None]], axis=1)
    assert len(src) > 3 * len(R) and np.count_nonzero(cf) > len(R)
    # contact-gated inputs are compared where the gate is not within rounding of its threshold
    ok = np.abs(o) > 1e-4
    assert np.all(np.abs(e[ok] - ref[ok]) <= 8e-3 * np.abs(ref[ok]) + 1e-3)
    xr = np.concatenate([((R - 0.75) / 0.75 * 4.0)[:, None], om * R[:, None] / 0.1, ((R / 0.75) ** 3)[:, None],
                         np.zeros((len(R), 3))], 1)
    assert np.all(np.abs(x - xr) <= 8e-3 * np.abs(xr) + 1e-3)


def test_forward_matches_oracle(net):
    _, gnn, flat = net
    out = gnn.forward()
    src, dst, e, x = gnn.dump_graph()
    ref, _ = G.Net(flat.astype(np.float64), exact=False).forward(x, src.astype(np.int64), dst.astype(np.int64), e)
    scale = float(np.max(np.abs(ref)))
    assert scale > 0
    diff = np.abs(out - ref)
    assert float(np.max(diff)) <= 1e-2 * scale, float(np.max(diff)) / scale
    assert float(np.mean(diff)) <= 1e-3 * scale


def test_inference_forward_matches_oracle(net):
    """A network created for inference runs the decoder inside the last update product: same outputs
    as the oracle and as the training-mode forward pass (up to FP32 summation order)."""
    sim, gnn, flat = net
    inf = native.Gnn(sim.ctx, r0=0.75, v0=0.1, a0=1.0, m0=4000.0, ov_scale=50.0, substeps=10, train=False)
    inf.set_params(flat)
    out = inf.forward()
    ref_gpu = gnn.forward()
    src, dst, e, x = gnn.dump_graph()
    ref, _ = G.Net(flat.astype(np.float64), exact=False).forward(x, src.astype(np.int64), dst.astype(np.int64), e)
    scale = float(np.max(np.abs(ref)))
    diff = np.abs(out - ref)
    assert float(np.max(diff)) <= 1e-2 * scale, float(np.max(diff)) / scale
    assert float(np.mean(diff)) <= 1e-3 * scale
    assert float(np.max(np.abs(out - ref_gpu))) <= 1e-5 * scale
    inf.close()


def test_backward_matches_oracle(net):
    _, gnn, flat = net
    out = gnn.forward()
    rng = np.random.default_rng(9)
    target = (out + rng.normal(0, 0.5, out.shape)).astype(np.float32)
    gnn.set_target(target)
    loss = gnn.backward()
    grads = gnn.get_grads()
    src, dst, e, x = gnn.dump_graph()
    onet = G.Net(flat.astype(np.float64), exact=False)
    o_out, cache = onet.forward(x, src.astype(np.int64), dst.astype(np.int64), e)
    assert loss == pytest.approx(onet.loss(o_out, target), rel=2e-2)
    og = onet.backward(cache, target)
    off = 0
    for name, shape in G.param_layout():
        k = int(np.prod(shape))
        a = grads[off:off + k].astype(np.float64)
        b = np.asarray(og[name], np.float64).ravel()
        off += k
        nb_ = np.linalg.norm(b)
        assert nb_ > 0, name
        assert np.linalg.norm(a - b) <= 2e-2 * nb_, (name, np.linalg.norm(a - b) / nb_)


def test_training_lowers_loss_and_beats_free_flight(gpu_ok):
    """The library's training loop on 16 packings of 4000 spheres (25 windows of 10 DEM steps each);
    then, on windows of a packing it has not seen, the network's acceleration error against the DEM
    is compared with the method without it (zero contact acceleration)."""
    params, scales, hist = surrogate.train(n=4000, runs=16, steps_per_run=250, substeps=10, lr=1e-3, seed=50,
                                           log=lambda m: None)
    losses = [h["loss"] for h in hist if h["loss"] is not None]
    assert len(losses) == 16
    assert np.mean(losses[-4:]) < 0.5 * losses[0], losses
    w2 = surrogate.training_workload(4000, seed=977)
    sim2 = SphereSim(w2)
    sim2.first_iteration()
    surrogate.settle(sim2, 10)
    g2 = surrogate.attach(sim2.ctx, params, scales)
    errs = [surrogate.window_errors(sim2, g2, 10) for _ in range(5)]
    errs = [e for e in errs if e is not None]
    assert errs
    net_err = np.mean([e[0] for e in errs])
    free_err = np.mean([e[1] for e in errs])
    print(f"trained: run losses {losses[0]:.4g} -> {losses[-1]:.4g}; unseen packing: RMS contact-force "
          f"error / (m0 a0) network {net_err:.4g}, free flight {free_err:.4g}")
    assert net_err < 0.7 * free_err, (net_err, free_err)


def test_surrogate_run_with_drift_checks(gpu_ok):
    """The surrogate-accelerated run: checks happen, the DEM state is kept at checks, errors are
    finite; an impossible tolerance makes every check fail and the run fall back to DEM."""
    w = periodic_pack(n=2000, sigma_v=0.1)
    sim = SphereSim(w)
    sim.first_iteration()
    gnn = native.Gnn(sim.ctx, 0.75, 0.1, 1.0, 4000.0, 50.0, 10, train=False)
    gnn.set_params(surrogate.init_params(2))
    sch = scheme(w)
    st = gnn.surrogate_run(sch, 40, check_every=10, tol=1e9)
    assert st["checks"] == 4 and st["fallbacks"] == 0
    assert st["surrogate_steps"] == 36 and st["dem_steps"] == 40
    assert np.isfinite(st["max_err"]) and np.isfinite(st["max_err_free"])
    st2 = gnn.surrogate_run(sch, 40, check_every=10, tol=0.0)
    assert st2["fallbacks"] >= 1 and st2["dem_steps"] > 40
    p = sim.ctx.download()
    assert np.all(np.isfinite(p["rx"])) and np.all(np.isfinite(p["vx"]))


@pytest.mark.parametrize("weights", ["k10", "k100", "k200"])
def test_shipped_weights_beat_free_flight(gpu_ok, weights):
    """The shipped networks (K = 10, 100 and 200 DEM steps per surrogate step) on a packing they were
    not trained on: drift-checked run without fallback, RMS position error at the checks and RMS
    contact-force error on a window below the free-flight (zero contact force) values."""
    path = {"k10": surrogate.WEIGHTS, "k100": surrogate.WEIGHTS_K100, "k200": surrogate.WEIGHTS_K200}[weights]
    params, scales = surrogate.load(path)
    K = scales["substeps"]
    assert K == int(weights[1:])
    w = surrogate.training_workload(20000, seed=4242)
    sim = SphereSim(w)
    sim.first_iteration()
    surrogate.settle(sim, K)
    g = surrogate.attach(sim.ctx, params, scales)
    st = g.surrogate_run(scheme(w), 20, check_every=10, tol=0.01)
    assert st["checks"] == 2 and st["fallbacks"] == 0
    assert 0 < st["rms_err"] < st["rms_err_free"], st
    errs = [e for e in (surrogate.window_errors(sim, g, K) for _ in range(2)) if e is not None]
    assert errs
    print(f"{weights}: rms position error {st['rms_err']:.3g} R (free flight {st['rms_err_free']:.3g} R), "
          f"contact-force error {errs[-1][0]:.3g} (free flight {errs[-1][1]:.3g}) m0 a0")
    assert errs[-1][0] < 0.3 * errs[-1][1], errs
