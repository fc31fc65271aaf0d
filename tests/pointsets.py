"""Seeded point layouts of the reference's test suite (test/testutils.jl:50-76), 3D."""
import numpy as np


def uniform(rng, n):
    return rng.random((n, 3))


def clustered(rng, n, nclusters=10, spread=0.01):
    c = 0.1 + 0.8 * rng.random((nclusters, 3))
    return np.clip(c[np.arange(n) % nclusters] + spread * rng.standard_normal((n, 3)), 0, 1)


def sphere(rng, n, center=(0.5, 0.5, 0.5), radius=0.4):
    v = rng.standard_normal((n, 3))
    return np.asarray(center) + radius * v / np.linalg.norm(v, axis=1, keepdims=True)


def helix(rng, n):
    th = 6 * np.pi * rng.random(n)
    p = np.stack([0.5 + 0.3 * np.cos(th), 0.5 + 0.3 * np.sin(th), 0.05 + 0.9 * th / (6 * np.pi)], axis=1)
    return np.clip(p + 1e-3 * rng.standard_normal((n, 3)), 0, 1)


def one_octant(rng, n):
    return 0.3 * rng.random((n, 3))


def mixed(rng, n):
    ns, nk = n // 2, n // 4
    s = sphere(rng, ns, radius=0.3)
    k = 0.9 + 0.002 * rng.standard_normal((nk, 3))
    return np.clip(np.vstack([s, k, uniform(rng, n - ns - nk)]), 0, 1)


DISTRIBUTIONS = dict(uniform=uniform, clustered=clustered, sphere=sphere, helix=helix, one_octant=one_octant,
                     mixed=mixed)
