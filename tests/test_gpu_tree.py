"""The device tree and its lists against the oracle's restatement of the host builder,
field by field (src/tree/builder.jl, src/tree/lists.jl; the reference's own check is
test/test_device_tree.jl:12-38 and test/test_device_passes.jl:111-147): every
structural array bit-exact, the points of every leaf the same set (their order inside
a leaf follows the full key on the device, the input order on the host)."""
import numpy as np
import pytest

import oracle as O
from pointsets import DISTRIBUTIONS, mixed, sphere, uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)
FIELDS = ["level_start", "cell", "level", "parent", "child_first", "plo", "phi", "leaf", "ghost", "xlist",
          "has_source", "colleagues", "point_node"]


def device_plan(pts, ppl, **kw):
    import ffmm
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=4)
    return ffmm.DevicePlan(m, pts, ppl=ppl, bbox=BOX, **kw)


def assert_same_tree(dt, ref, pts):
    for f in FIELDS:
        a, b = np.asarray(getattr(dt, f)), np.asarray(getattr(ref, f))
        assert a.shape == b.shape, f
        assert np.array_equal(a, b), f"field {f} differs"
    # the points of each leaf, as sets
    for i in np.nonzero(ref.leaf)[0]:
        a, b = ref.plo[i], ref.phi[i] + 1
        assert np.array_equal(np.sort(dt.perm[a:b]), np.sort(ref.perm[a:b]))
    assert np.array_equal(np.sort(dt.perm), np.arange(len(pts)))
    assert np.array_equal(dt.points, pts[dt.perm])


CASES = [("mixed", 8000, 40), ("sphere", 8000, 60), ("helix", 8000, 50), ("clustered", 6000, 30),
         ("uniform", 5000, 20), ("one_octant", 3000, 20)]


@pytest.mark.parametrize("name,n,ppl", CASES)
def test_device_tree_equals_host_tree(name, n, ppl):
    print(f"[tree] {name} n={n} ppl={ppl}", flush=True)
    rng = np.random.default_rng(41)
    pts = DISTRIBUTIONS[name](rng, n)
    ref = O.build_tree(pts, ppl, BOX[0], BOX[1])
    plan = device_plan(pts, ppl)
    dt = plan.download_tree()
    assert plan.nlevels == ref.nlevels and plan.nnodes == ref.nnodes
    assert_same_tree(dt, ref, pts)
    for l in range(ref.nlevels):
        ptr, src, far = plan.download_lists(l)
        rptr, rsrc, rfar = ref.lists[l]
        assert np.array_equal(ptr, rptr) and np.array_equal(src, rsrc) and np.array_equal(far, rfar), f"level {l}"
    # deterministic
    assert np.array_equal(device_plan(pts, ppl).download_tree().perm, dt.perm)


def test_spectral_plan_pairs_written_on_demand():
    """A spectral plan writes its far pairs only when they are downloaded (its M2L reads
    the per-target counts alone); they equal the oracle's lists."""
    import ffmm
    rng = np.random.default_rng(43)
    pts = DISTRIBUTIONS["mixed"](rng, 8000)
    print("[tree] spectral plan lists", flush=True)
    ref = O.build_tree(pts, 40, BOX[0], BOX[1])
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, m2l="spectral", precision="mixed")
    for l in range(ref.nlevels):
        ptr, src, far = plan.download_lists(l)
        rptr, rsrc, rfar = ref.lists[l]
        assert np.array_equal(ptr, rptr) and np.array_equal(src, rsrc) and np.array_equal(far, rfar), f"level {l}"
    assert np.array_equal(plan.level_pairs, [len(ref.lists[l][1]) for l in range(ref.nlevels)])


def test_device_tree_gaussian_mixture():
    import ffmm.workloads as W
    pts = W.gaussian_mixture(200_000, seed=7)
    print("[tree] gaussian mixture 200k ppl=64", flush=True)
    ref = O.build_tree(pts, 64, BOX[0], BOX[1])
    plan = device_plan(pts, 64)
    assert plan.nlevels == ref.nlevels >= 8
    assert_same_tree(plan.download_tree(), ref, pts)


def test_device_tree_small_cases_and_guards():
    import ffmm
    rng = np.random.default_rng(5)
    pts = uniform(rng, 10)
    plan = device_plan(pts, 50)
    assert plan.nlevels == 1 and plan.nnodes == 1
    assert_same_tree(plan.download_tree(), O.build_tree(pts, 50, BOX[0], BOX[1]), pts)
    # eight octant clusters: two levels, nine nodes, no ghosts (test/test_3d.jl:101-106)
    oc = np.array([[0.25 + 0.5 * a, 0.25 + 0.5 * b, 0.25 + 0.5 * c] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
    pts8 = np.repeat(oc, 6, axis=0) + 0.01 * rng.random((48, 3))
    p8 = device_plan(pts8, 6)
    t8 = p8.download_tree()
    assert p8.nlevels == 2 and p8.nnodes == 9 and t8.ghost.sum() == 0
    with pytest.raises(ffmm.ArgumentError, match="max_levels"):
        device_plan(np.full((20, 3), 0.3), 5)
    with pytest.raises(ffmm.ArgumentError, match="outside"):
        device_plan(pts + 0.5, 5)
    bad = pts.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ffmm.ArgumentError, match="finite"):
        device_plan(bad, 5)
