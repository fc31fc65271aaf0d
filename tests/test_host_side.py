"""Host-side logic that needs no GPU: the Stokes front end's density rows and error
phrases (src/stokes_api.jl:72-83), the seeded workload generators (the ellipsoid
surfaces of the Stokes workload: points on the surfaces, unit outward normals,
area-uniform sampling, determinism), and bench.py's record of the step's GEMMs."""
import numpy as np
import pytest


def test_stokes_density_rows_and_errors():
    import ffmm
    rng = np.random.default_rng(3)
    n = 7
    f, g = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    nu = rng.standard_normal((n, 3))
    w = rng.random(n)
    d = ffmm.stokes_densities(n, weights=w, stokeslets=f, stresslets=g, stressvecs=nu)
    assert d.shape == (n, 9)
    assert np.array_equal(d[:, :3], f * w[:, None]) and np.array_equal(d[:, 3:6], g * w[:, None])
    assert np.array_equal(d[:, 6:], nu)
    d = ffmm.stokes_densities(n, stokeslets=f)
    assert np.array_equal(d[:, :3], f) and not d[:, 3:].any()
    with pytest.raises(ffmm.ArgumentError, match="given together"):
        ffmm.stokes_densities(n, stresslets=g)
    with pytest.raises(ffmm.ArgumentError, match="at least one"):
        ffmm.stokes_densities(n)
    with pytest.raises(ffmm.DimensionMismatch, match="weights has length 6, expected 7"):
        ffmm.stokes_densities(n, weights=w[:-1], stokeslets=f)
    with pytest.raises(ffmm.ArgumentError, match="doublets"):
        ffmm.stokes_densities(n, doublets=f)


def test_ellipsoid_surfaces():
    from ffmm import workloads as W
    rng = np.random.default_rng(0)
    # one ellipsoid by hand: the generator's sampling on a known shape
    pts, nrm = W.ellipsoids(64 * 500, seed=11)
    assert pts.shape == nrm.shape == (64 * 500, 3)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-12)
    assert pts.min() >= 0.0 and pts.max() <= 1.0
    a, b = W.ellipsoids(1000, seed=11)
    assert np.array_equal(a, W.ellipsoids(1000, seed=11)[0])     # seeded
    # each ellipsoid's points: the normal is the gradient direction of the quadric, and
    # the area element is uniform (a cap test on a flat ellipsoid)
    r = np.random.default_rng(11)
    centers = 0.15 + 0.7 * r.random((64, 3))
    axes = r.uniform(0.02, 0.08, (64, 3))
    R = W._rotation(r)
    e0 = pts[:500]
    y = (e0 - centers[0]) @ R                       # body frame
    q = np.sum((y / axes[0]) ** 2, axis=1)
    assert np.allclose(q, 1.0, atol=1e-9)
    grad = y / axes[0] ** 2
    grad /= np.linalg.norm(grad, axis=1)[:, None]
    assert np.allclose(nrm[:500] @ R, grad, atol=1e-9)
    rot = W._rotation(rng)
    assert np.allclose(rot @ rot.T, np.eye(3)) and np.isclose(np.linalg.det(rot), 1.0)


def test_area_uniform_sampling_on_a_flat_ellipsoid():
    """On a flat ellipsoid (a, a, a/20) almost all the area is on the two faces, which
    the direction-uniform sampling would under-weight: the area-uniform sampler puts the
    share of points with |u_z| > 0.9 at their share of the area (0.19; direction-uniform
    sampling gives 0.1)."""
    rng = np.random.default_rng(5)
    a = np.array([0.05, 0.05, 0.0025])
    m = 200000
    u = rng.standard_normal((3 * m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.linalg.norm(u / a[None, :], axis=1) * a.min()
    kept = u[rng.random(len(u)) < w]
    share = np.mean(np.abs(kept[:, 2]) > 0.9)
    # the faces' share of the area, by quadrature of the area element over the sphere
    v = rng.standard_normal((2_000_000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    dA = np.linalg.norm(v / a[None, :], axis=1)
    ref = np.sum(dA[np.abs(v[:, 2]) > 0.9]) / np.sum(dA)
    assert abs(share - ref) < 0.01
    assert share > 0.15


def test_bench_gemm_record_splits_the_wide_mixed_m2m():
    """bench.py's record of the cube step's GEMMs, in launch order: the mixed form's M2M
    of a level with more than 10,240 sibling groups is recorded as the library calls it
    is made of (at most 10,240 columns each, covering the level's groups once, the same
    flops in total); the FP64 form keeps one call per level."""
    import types
    import bench
    Ns, Nl = 296, 1352
    ops = types.SimpleNamespace(nsmall=Ns, nlarge=Nl, P_up=types.SimpleNamespace(rank=296),
                                P_down=types.SimpleNamespace(rank=296))
    counts = np.array([1, 8, 64, 512, 8 * 25_000, 800])
    plan = types.SimpleNamespace(nlevels=len(counts), level_count=counts, level_s2m=np.full(len(counts), 1000),
                                 level_leaves=np.full(len(counts), 900))
    mixed, plain = bench.gemm_shapes(plan, ops, True), bench.gemm_shapes(plan, ops, False)
    m3 = [g for g in mixed if g["name"].startswith("M2M level 3")]
    assert [g["n"] for g in m3] == [10_240, 10_240, 4_520]
    assert [g["name"] for g in m3] == ["M2M level 3 (columns 0-10239)", "M2M level 3 (columns 10240-20479)",
                                       "M2M level 3 (columns 20480-24999)"]
    assert all(g["library"] and g["type"] == "f32" and g["m"] == Ns and g["k"] == 8 * Ns for g in m3)
    p3 = [g for g in plain if g["name"].startswith("M2M level 3")]
    assert len(p3) == 1 and p3[0]["n"] == 25_000 and p3[0]["type"] == "f64"
    assert sum(g["flop"] for g in m3) == p3[0]["flop"] == 2.0 * Ns * 25_000 * 8 * Ns
    # launch order: M2M from level L-2 up to 2, then L2L from level 3 down
    names = [g["name"].split(" (")[0] for g in mixed]
    assert names == ["UC2E Uh", "UC2E V", "M2M level 4", "M2M level 3", "M2M level 3", "M2M level 3", "M2M level 2",
                     "L2L level 3", "L2L level 4", "L2L level 5", "DC2E Uh", "DC2E V"]
    # narrower levels are one call, named as before
    assert [g["name"] for g in mixed if g["name"].startswith("M2M level 4")] == ["M2M level 4"]
