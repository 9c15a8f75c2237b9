"""CPU checks of the boundary and the host side: library exports, generator determinism, packing."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from exadem_amd import inputs, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(xdem_[a-z0-9_]+)\s*\(", txt)))


def test_hip_library_exports_every_declared_symbol():
    lib = C.CDLL(native.lib_path())
    names = declared("xdem.h")
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(native.HIP_SYMBOLS) | set(names)


def test_host_library_exports_every_declared_symbol():
    lib = native.load_host()
    names = declared("xdem_host.h")
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(native.HOST_SYMBOLS)


def test_context_creation_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(native.XdemError, match="ERR_HIP"):
        native.Context()


def test_generators_are_deterministic():
    a = inputs.uniform(7, 1, 1000, 0.5, 1.0)
    b = inputs.uniform(7, 1, 1000, 0.5, 1.0)
    c = inputs.uniform(8, 1, 1000, 0.5, 1.0)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0.5 and a.max() < 1.0
    z = inputs.normal(3, 2, 200000, 0.0, 0.1)
    assert abs(z.mean()) < 2e-3 and z.std() == pytest.approx(0.1, rel=1e-2)


def test_polydisperse_packing():
    w = inputs.polydisperse(n=5000, relax_sweeps=200)
    p = w.particles
    box = np.array(w.meta["box"])
    vol = np.sum(4.0 / 3.0 * np.pi * p["radius"] ** 3)
    assert vol / np.prod(box) == pytest.approx(0.55, rel=1e-12)
    assert w.meta["max_overlap"] < 0.05
    assert np.all(p["rx"] >= 0) and np.all(p["rx"] < box[0])
    w2 = inputs.polydisperse(n=5000, relax_sweeps=200)
    assert np.array_equal(w.particles["rx"], w2.particles["rx"])


def test_deposition_lattice():
    w = inputs.deposition()
    p = w.particles
    assert w.n == 10000
    assert p["mass"][0] == pytest.approx(2500 * 4 / 3 * np.pi * 0.125, rel=1e-15)
    assert p["inx"][0] == pytest.approx(0.4 * p["mass"][0] * 0.25, rel=1e-15)
    assert p["rx"].min() >= 1.45 and p["rx"].max() <= 38.55
    assert len(w.walls) == 6
