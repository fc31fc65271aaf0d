"""Seeded synthetic point clouds of the benchmark configurations.

* ``gaussian_mixture``: N points from a mixture of 32 isotropic Gaussians, centres
  uniform in [0.1, 0.9]^3, sigma log-uniform in [0.005, 0.05], point i in component
  i mod 32 (the reference's clustered generator, test/testutils.jl:53-56), clipped to
  the unit cube;
* ``uniform``: N points uniform in [0, 1]^3 (test/testutils.jl:52);
* ``sphere``, ``mixed``: the reference's sphere and sphere + cluster + background
  layouts (test/testutils.jl:57-68), the latter forcing level restriction and ghosts.
"""
from __future__ import annotations

import numpy as np


def gaussian_mixture(n: int, seed: int = 0, ncomp: int = 32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centers = 0.1 + 0.8 * rng.random((ncomp, 3))
    sigma = np.exp(rng.uniform(np.log(0.005), np.log(0.05), ncomp))
    comp = np.arange(n) % ncomp
    pts = centers[comp] + sigma[comp, None] * rng.standard_normal((n, 3))
    return np.clip(pts, 0.0, 1.0)


def uniform(n: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3))


def sphere(n: int, seed: int = 0, center=(0.5, 0.5, 0.5), radius: float = 0.4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return np.asarray(center) + radius * v / np.linalg.norm(v, axis=1, keepdims=True)


def mixed(n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    ns, nk = n // 2, n // 4
    v = rng.standard_normal((ns, 3))
    s = 0.5 + 0.3 * v / np.linalg.norm(v, axis=1, keepdims=True)
    k = 0.9 + 0.002 * rng.standard_normal((nk, 3))
    u = rng.random((n - ns - nk, 3))
    return np.clip(np.vstack([s, k, u]), 0.0, 1.0)


def charges(n: int, seed: int = 1) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(n)
