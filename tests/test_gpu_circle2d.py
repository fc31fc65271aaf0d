"""The 2D circle method (KernelIndependent(Laplace2D(); nequiv)) through the library on
the GPU: the quadtree and lists it builds equal the oracle's field by field; the
device passes match the oracle's FP64 FMM on the same tree (1e-12) and the FP64 direct
sums (rel L2 1e-12 at nequiv = 60, the method's floor: docs/kifmm.md:169, circles
60 at 8e-14); the host API (FMM / build_expansions / evaluate_at_sources / evaluate,
README.md:38-51) returns the same; errors carry the reference's phrases."""
import numpy as np
import pytest

import circle2d as O2

pytestmark = pytest.mark.gpu


def _layout(kind, n, rng):
    if kind == "uniform":
        return rng.random((n, 2))
    a = np.stack([rng.random(n // 2), 0.5 + 1e-3 * rng.standard_normal(n // 2)], 1)
    b = 0.25 + 0.01 * rng.standard_normal((n - n // 2, 2))
    return np.clip(np.concatenate([a, b]), 0, 1)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kind,n,ppl,nequiv", [("uniform", 5000, 50, 60), ("clustered", 4000, 20, 60),
                                                ("uniform", 3000, 30, 24)])
def test_circle_plan_against_oracle(kind, n, ppl, nequiv):
    import ffmm
    rng = np.random.default_rng(60 + n)
    pts = _layout(kind, n, rng)
    tau = rng.standard_normal(n)
    tgt = rng.random((700, 2))
    m = ffmm.KernelIndependent(ffmm.Laplace2D(), nequiv=nequiv)
    cache = ffmm.CircleOperatorCache(m)
    print(f"[circle2d] {kind} n={n} ppl={ppl} nequiv={nequiv}", flush=True)
    plan = ffmm.CirclePlan(m, pts, ppl=ppl, bbox=((0.0, 0.0), 1.0), cache=cache)
    ref = O2.FMM2D(cache.operators, pts, ppl, (0.0, 0.0), 1.0)
    t, to = plan.download_tree(), ref.t
    assert plan.nlevels == to.nlevels and plan.nnodes == to.nnodes
    assert np.array_equal(plan.level_start, to.level_start)
    for f in ("cell", "parent", "child_first", "plo", "phi", "leaf", "ghost", "xlist", "colleagues", "perm"):
        assert np.array_equal(np.asarray(getattr(t, f)), np.asarray(getattr(to, f))), f
    for l in range(plan.nlevels):
        for a, b in zip(t.lists[l], to.lists[l]):
            assert np.array_equal(a, b), l
    u = plan.matvec(tau).cpu().numpy()
    v = plan.evaluate(tgt).cpu().numpy()
    uo, vo = ref.matvec(tau, tgt)
    scale = np.max(np.abs(uo))
    assert np.max(np.abs(u - uo)) < 1e-12 * scale
    assert np.max(np.abs(v - vo)) < 1e-12 * np.max(np.abs(vo))
    gate = 1e-12 if nequiv == 60 else 1e-5
    assert rel(u, O2.direct_self(pts, tau)) < gate
    assert rel(v, O2.direct(pts, tau, tgt)) < gate


def test_circle_host_api_and_errors():
    """The reference's 2D usage (README.md:38-51) through ffmm.FMM; refusals."""
    import ffmm
    rng = np.random.default_rng(7)
    n = 20000
    pts = rng.random((n, 2))
    tau = rng.standard_normal(n)
    tgt = rng.random((5000, 2))
    m = ffmm.KernelIndependent(ffmm.Laplace2D(), nequiv=60)
    print("[circle2d] host API, 20,000 sources, 5,000 targets", flush=True)
    f = ffmm.FMM(m, pts, ppl=50, bbox=((0.0, 0.0), 1.0))
    st = ffmm.build_expansions(f, pts, tau)
    u = ffmm.evaluate_at_sources(f, st)
    v = ffmm.evaluate(f, st, tgt)
    rows = np.sort(rng.choice(n, 2000, replace=False))
    assert rel(u[rows], O2.direct_self(pts, tau, rows=rows)) < 1e-12
    assert rel(v, O2.direct(pts, tau, tgt)) < 1e-12
    # the same charges again (build_expansions on the fixed tree): bitwise the same
    st2 = ffmm.build_expansions(f, pts, tau)
    assert np.array_equal(ffmm.evaluate_at_sources(f, st2), u)
    # linearity
    t2 = rng.standard_normal(n)
    u2 = ffmm.evaluate_at_sources(f, ffmm.build_expansions(f, pts, t2))
    u12 = ffmm.evaluate_at_sources(f, ffmm.build_expansions(f, pts, tau + 2 * t2))
    assert np.max(np.abs(u12 - (u + 2 * u2))) < 1e-12 * np.max(np.abs(u12))
    with pytest.raises(ffmm.ArgumentError, match="outside the requested box"):
        ffmm.FMM(m, pts + 0.5, ppl=50, bbox=((0.0, 0.0), 1.0))
    bad = pts.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ffmm.ArgumentError, match="finite"):
        ffmm.FMM(m, bad, ppl=50, bbox=((0.0, 0.0), 1.0))
    with pytest.raises(ffmm.ArgumentError, match="outside the tree's box"):
        ffmm.evaluate(f, st, tgt + 1.5)
    with pytest.raises(ffmm.ArgumentError, match="max_levels"):
        ffmm.FMM(m, np.full((100, 2), 0.3), ppl=10, bbox=((0.0, 0.0), 1.0), max_levels=6)
    with pytest.raises(ffmm.ArgumentError, match="potential"):
        ffmm.evaluate_at_sources(f, st, output=ffmm.Gradient)
