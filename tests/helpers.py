"""Shared helpers of the parity tests: small seeded workloads and id-aligned comparisons."""
import math

import numpy as np

from exadem_amd import inputs, native


def by_id(p):
    """Reorders a particle dict by ascending id."""
    order = np.argsort(p["id"], kind="stable")
    return {k: v[order] for k, v in p.items()}


def vec(p, base):
    return np.stack([p[base + "x"], p[base + "y"], p[base + "z"]], axis=1)


def pos(p):
    return np.stack([p["rx"], p["ry"], p["rz"]], axis=1)


def vel(p):
    return np.stack([p["vx"], p["vy"], p["vz"]], axis=1)


def force(p):
    return np.stack([p["fx"], p["fy"], p["fz"]], axis=1)


def rel_err(a, b, floor):
    """max_i |a_i - b_i| / max(|b_i|, floor_i) over rows (vectors)."""
    d = np.linalg.norm(a - b, axis=1)
    s = np.maximum(np.linalg.norm(b, axis=1), floor)
    return float(np.max(d / s)) if len(d) else 0.0


def dense_deposition(n=2000, seed=1):
    """Contact-rich deposition (lattice spacing 0.98 < diameter: initial overlaps), walled box."""
    return inputs.deposition(n=n, spacing=0.98, seed=seed, box=(16.0, 16.0, 80.0), start=0.55)


def wall_deposition(n=600, seed=7):
    """Particles touching the six walls of a small box (wall contacts from the first step)."""
    return inputs.deposition(n=n, spacing=0.99, seed=seed, box=(7.89, 7.89, 12.0), start=0.48, jitter=0.02)


def periodic_pack(n=3000, seed=11, sigma_v=0.3):
    return inputs.polydisperse(n=n, seed_radius=seed, seed_vel=seed + 1, seed_pos=seed + 2, sigma_v=sigma_v,
                               relax_sweeps=60)


def interaction_table(ii, jj, tt, fr):
    """Canonical sort of an interaction dump by (owner id, partner id, type)."""
    keys = np.stack([ii.astype(np.uint64), jj.astype(np.uint64), tt.astype(np.uint64)], axis=1)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    return keys[order], fr[order]
