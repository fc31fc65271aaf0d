"""Seeded synthetic point clouds of the benchmark configurations.

* ``gaussian_mixture``: N points from a mixture of 32 isotropic Gaussians, centres
  uniform in [0.1, 0.9]^3, sigma log-uniform in [0.005, 0.05], point i in component
  i mod 32 (the reference's clustered generator, test/testutils.jl:53-56), clipped to
  the unit cube;
* ``uniform``: N points uniform in [0, 1]^3 (test/testutils.jl:52);
* ``sphere``, ``mixed``: the reference's sphere and sphere + cluster + background
  layouts (test/testutils.jl:57-68), the latter forcing level restriction and ghosts;
* ``ellipsoids``: N points on the surfaces of 64 ellipsoids (semi-axes uniform in
  [0.02, 0.08], random rotations, centres uniform in [0.15, 0.85]^3), area-uniform on
  each, with their outward unit normals (the Stokes surface workload);
* ``stokes_densities``: stokeslet strengths N(0,1)^3 plus stresslets (strength
  N(0,1)^3, the given normals), as the (N, 9) rows [f, g, nu].
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


def _rotation(rng) -> np.ndarray:
    """A Haar-random rotation: QR of a Gaussian matrix with the signs of R's diagonal."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoids(n: int, seed: int = 0, nell: int = 64):
    """(points, normals): n points, split evenly over nell ellipsoids.  On an ellipsoid
    x = c + R diag(a) u (u on the unit sphere) the area element is proportional to
    |diag(1/a) u|, so directions u uniform on the sphere are kept with probability
    min(a) |diag(1/a) u| (at most 1): the kept points are area-uniform."""
    rng = np.random.default_rng(seed)
    centers = 0.15 + 0.7 * rng.random((nell, 3))
    axes = rng.uniform(0.02, 0.08, (nell, 3))
    counts = np.full(nell, n // nell)
    counts[: n - counts.sum()] += 1
    pts, nrm = [], []
    for e in range(nell):
        R, a, m = _rotation(rng), axes[e], int(counts[e])
        got = []
        while sum(len(g) for g in got) < m:
            u = rng.standard_normal((2 * m + 16, 3))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            w = np.linalg.norm(u / a[None, :], axis=1) * a.min()
            got.append(u[rng.random(len(u)) < w])
        u = np.vstack(got)[:m]
        nv = u / a[None, :]
        nv /= np.linalg.norm(nv, axis=1, keepdims=True)
        pts.append(centers[e] + (u * a[None, :]) @ R.T)
        nrm.append(nv @ R.T)
    return np.clip(np.vstack(pts), 0.0, 1.0), np.vstack(nrm)


def stokes_densities(normals: np.ndarray, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = len(normals)
    return np.hstack([rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), normals])


def served_stream(nreq: int, nmin: int = 10_000, nmax: int = 2_000_000, rate: float = 100.0, seed: int = 5,
                  p_laplace: float = 0.7, p_repeat: float = 0.4):
    """The served mixed-size stream: ``nreq`` requests with Poisson arrivals at ``rate``
    per second, point counts log-uniform in [nmin, nmax], a fraction ``p_laplace`` for
    Laplace3D on cube surfaces (p = 8) and the rest for Yukawa3D(5) spherical harmonics
    (p = 10), uniform or Gaussian-mixture layouts (even odds), charges N(0, 1), a fraction
    ``p_repeat`` repeating an earlier geometry of the same kernel with new charges, and
    the 1e-6 and 1e-3 accuracy classes (even odds).

    Returns request dicts: ``t`` (arrival, s), ``lap`` (Laplace or Yukawa), ``fine`` (the
    1e-6 class), ``gid`` (geometry id: a repeat shares its first request's), ``n``,
    ``layout``, ``seed`` (of the points) and ``tau_seed``; ``request_points(r)`` and
    ``request_charges(r)`` make the arrays (the same for every request of a geometry)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs, geoms = [], {True: [], False: []}
    for k in range(nreq):
        lap = bool(rng.random() < p_laplace)
        fine = bool(rng.random() < 0.5)
        if geoms[lap] and rng.random() < p_repeat:
            g = geoms[lap][int(rng.integers(len(geoms[lap])))]
        else:
            n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
            layout = "gaussian_mixture" if rng.random() < 0.5 else "uniform"
            g = dict(gid=sum(len(v) for v in geoms.values()), n=n, layout=layout, seed=int(rng.integers(1 << 30)))
            geoms[lap].append(g)
        reqs.append(dict(t=float(t[k]), lap=lap, fine=fine, gid=g["gid"], n=g["n"], layout=g["layout"],
                         seed=g["seed"], tau_seed=int(rng.integers(1 << 30))))
    return reqs


def request_points(r) -> np.ndarray:
    return gaussian_mixture(r["n"], seed=r["seed"]) if r["layout"] == "gaussian_mixture" else \
        uniform(r["n"], seed=r["seed"])


def request_charges(r) -> np.ndarray:
    return np.random.default_rng(r["tau_seed"]).standard_normal(r["n"])
