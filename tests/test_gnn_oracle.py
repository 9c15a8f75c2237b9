"""The surrogate network's CPU oracle, pinned by mathematics: its backward against central finite
differences of its forward (float64), and its bfloat16 rounding against torch's conversion."""
import numpy as np
import pytest

import gnn_oracle as G


def small_graph(seed, n=14, e=50):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, e)
    dst = (src + rng.integers(1, n, e)) % n
    x = rng.normal(0, 1, (n, G.FN))
    ef = rng.normal(0, 1, (e, G.FE))
    target = rng.normal(0, 1, (n, G.FO))
    flat = G.init_params(seed)
    flat = flat + rng.normal(0, 0.05, flat.shape)  # non-zero biases, output layer not tiny
    return flat, x, src, dst, ef, target


def test_backward_matches_finite_differences():
    flat, x, src, dst, ef, target = small_graph(1)
    net = G.Net(flat, exact=True)
    out, cache = net.forward(x, src, dst, ef)
    grads = G.flatten(net.backward(cache, target))
    rng = np.random.default_rng(2)
    # probe entries of every tensor
    offsets, o = {}, 0
    for name, shape in G.param_layout():
        offsets[name] = (o, int(np.prod(shape)))
        o += int(np.prod(shape))
    probes = []
    for name, (o0, k) in offsets.items():
        probes += list(o0 + rng.choice(k, size=min(k, 6), replace=False))
    eps = 1e-6
    worst = 0.0
    for idx in probes:
        fp = flat.copy()
        fp[idx] += eps
        fm = flat.copy()
        fm[idx] -= eps
        lp = net.loss(G.Net(fp).forward(x, src, dst, ef)[0], target)
        lm = net.loss(G.Net(fm).forward(x, src, dst, ef)[0], target)
        fd = (lp - lm) / (2 * eps)
        scale = max(abs(fd), 1e-6 * float(np.max(np.abs(grads))))
        worst = max(worst, abs(fd - grads[idx]) / scale)
    assert worst < 1e-5, worst


def test_backward_catches_planted_mistake():
    """A dropped term (the residual path of one layer) makes the finite-difference check fail."""
    flat, x, src, dst, ef, target = small_graph(3)
    net = G.Net(flat, exact=True)
    out, cache = net.forward(x, src, dst, ef)
    good = G.flatten(net.backward(cache, target))
    # the same backward with the Q-side scatter sent to the source node instead of the target
    orig = np.add.at
    calls = {"n": 0}

    def wrong_add_at(a, idx, v):
        calls["n"] += 1
        # the dq accumulation is the second add.at of each layer in backward
        if calls["n"] % 2 == 0 and idx is cache["dst"]:
            return orig(a, cache["src"], v)
        return orig(a, idx, v)

    np.add.at = wrong_add_at
    try:
        bad = G.flatten(net.backward(cache, target))
    finally:
        np.add.at = orig
    assert np.max(np.abs(bad - good)) > 1e-3 * np.max(np.abs(good))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1.0, 65504.0, 3.0e38]])
    a = a.astype(np.float32)
    # values exactly halfway between two bfloat16 numbers (ties to even)
    halves = (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    a = np.concatenate([a, halves])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(G.bf16(a), ref)


def test_device_mode_close_to_exact():
    """bfloat16 storage changes the outputs by at most a few percent of their scale."""
    flat, x, src, dst, ef, target = small_graph(7, n=40, e=160)
    a, _ = G.Net(flat, exact=True).forward(x, src, dst, ef)
    b, _ = G.Net(flat, exact=False).forward(x, src, dst, ef)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(a))
