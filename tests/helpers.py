"""Shared builders for the parity tests: the same mesh, settings and seeded inputs
for the HIP path (dndsr_amd) and the CPU oracle."""
from __future__ import annotations

import numpy as np

from dndsr_amd import mesh as mm

RS = {"Roe": 1, "HLLC": 2, "HLLEP": 3, "HLLEP_V1": 21, "Roe_M1": 11, "Roe_M2": 12, "Roe_M3": 13, "Roe_M4": 14,
      "Roe_M5": 15, "Roe_M6": 16, "Roe_M7": 17, "Roe_M8": 18}


def small_mesh(kind: str, jitter: float = 0.15, seed: int = 7):
    if kind == "2d":
        return mm.periodic_quad_grid(8, 7, 10.0, 9.0, jitter=jitter, seed=seed)
    if kind == "3d":
        return mm.periodic_hex_grid(6, 5, 4, 1.0, 1.2, 0.9, jitter=jitter, seed=seed)
    raise ValueError(kind)


def vr_settings(maxOrder, intOrder=None, hqm=False):
    s = dict(maxOrder=maxOrder, intOrder=intOrder if intOrder is not None else max(maxOrder + 2, 3),
             jacobiRelax=1.0)
    if hqm:
        s.update(dirWeightScheme=1, geomWeightScheme=1, geomWeightBias=1.0, geomWeightPower=0.5)
    else:
        s.update(dirWeightScheme=0, geomWeightScheme=0, geomWeightBias=0.0, geomWeightPower=0.5)
    return s


def random_state(n, seed=0):
    """rho, p in U[0.8,1.2], velocity in U[-0.3,0.3]^3 -> conservative [n, 5]."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.8, 1.2, n)
    p = rng.uniform(0.8, 1.2, n)
    v = rng.uniform(-0.3, 0.3, (n, 3))
    E = p / 0.4 + 0.5 * rho * (v ** 2).sum(1)
    return np.concatenate([rho[:, None], rho[:, None] * v, E[:, None]], 1)


def random_rec(n, M, nVars=5, seed=1, amp=1e-2):
    rng = np.random.default_rng(seed)
    return rng.uniform(-amp, amp, (n, nVars, M))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def colour_scan_order(colors):
    """Serial scan order equivalent to the multicolour sweep: by colour, then index."""
    return np.lexsort((np.arange(colors.size), colors)).astype(np.int64)
