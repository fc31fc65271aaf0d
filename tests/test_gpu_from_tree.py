"""Plans from trees built on the host (ffmm_plan_from_tree, the reference's
DevicePlan(f::FMM), src/device_passes.jl:90-96) and the evaluation at targets located
by the caller (ffmm_evaluate_nodes, the reference's evaluate!(out, dev, ds, targets,
nodes), src/device.jl:127-145), on the GPU.

As the reference checks them (test/test_device_tree.jl:40-57): the plan from the host
tree and the plan from the points on the device have the same levels, node lists and
outputs (to 1e-10: the points inside a leaf are in another order); here the host tree
is the oracle's, so the plan from it also matches the oracle's FMM on that very tree
(same point order) to 1e-11."""
import numpy as np
import pytest

import harmonic as HO
import oracle as O
from pointsets import mixed, uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def mrd(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim == 2:
        return float(np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1)))
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def points():
    rng = np.random.default_rng(40)
    pts = mixed(rng, 20000)
    return pts, O.build_tree(pts, 60, *BOX), rng


@pytest.mark.parametrize("p,m2l", [(4, "dense"), (8, "spectral")])
def test_cube_plan_from_host_tree(points, p, m2l):
    import ffmm
    pts, tree, rng = points
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=p)
    cache = ffmm.OperatorCache(m)
    print(f"[from_tree] cube p={p} {m2l}", flush=True)
    plan_d = ffmm.DevicePlan(m, pts, ppl=60, bbox=BOX, cache=cache, m2l=m2l)
    plan_h = ffmm.DevicePlan.from_tree(m, tree, cache=cache, ppl=60, m2l=m2l)
    assert plan_h.nlevels == plan_d.nlevels == tree.nlevels
    assert np.array_equal(plan_h.level_count, plan_d.level_count)
    assert np.array_equal(plan_h.level_s2m, plan_d.level_s2m)
    assert np.array_equal(plan_h.level_leaves, plan_d.level_leaves)
    assert np.array_equal(plan_h.level_pairs, plan_d.level_pairs)
    th = plan_h.download_tree()
    assert np.array_equal(th.perm, tree.perm) and np.array_equal(th.colleagues, tree.colleagues)
    tau = rng.standard_normal(len(pts))
    u_h = plan_h.matvec(tau).cpu().numpy()
    u_d = plan_d.matvec(tau).cpu().numpy()
    assert mrd(u_h, u_d) < 1e-10
    ref = O.FMM(cache.host_operators(), pts, 60, *BOX)
    assert np.array_equal(ref.tree.perm, tree.perm)
    assert mrd(u_h, ref.matvec(tau)) < 1e-11


def test_stokes_plan_from_host_tree(points):
    import ffmm
    pts, tree, rng = points
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=6)
    cache = ffmm.OperatorCache(m)
    print("[from_tree] stokes p=6", flush=True)
    plan_d = ffmm.DevicePlan(m, pts, ppl=60, bbox=BOX, cache=cache)
    plan_h = ffmm.DevicePlan.from_tree(m, tree, cache=cache, ppl=60)
    dens = rng.standard_normal((len(pts), 3))
    u_h = plan_h.matvec(dens, ffmm.Velocity).cpu().numpy()
    u_d = plan_d.matvec(dens, ffmm.Velocity).cpu().numpy()
    assert mrd(u_h, u_d) < 1e-10
    W0 = cache.W0
    f = HO.HarmonicFMM("stokes", lambda l, w: (cache.harmonic_operators(w), W0 / w), pts, 60, *BOX)
    assert mrd(u_h, f.matvec(dens)) < 1e-11


def test_evaluate_at_host_located_targets(points):
    """Targets located by the host locate (the oracle's, src/tree/locate.jl:10-22) give
    the device-located results bitwise (the same leaves, the same order within each);
    a node that is not a real leaf is refused."""
    import ffmm
    pts, tree, rng = points
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=6)
    cache = ffmm.OperatorCache(m)
    plan = ffmm.DevicePlan.from_tree(m, tree, cache=cache, ppl=60)
    plan.build_expansions(rng.standard_normal(len(pts)))
    tgt = uniform(rng, 3000)
    nodes = tree.locate(tgt)
    print("[from_tree] evaluate at host-located targets", flush=True)
    for out in (ffmm.Potential, ffmm.PotentialGradient):
        v_dev = plan.evaluate(tgt, out).cpu().numpy()
        v_host = plan.evaluate(tgt, out, nodes=nodes).cpu().numpy()
        assert np.array_equal(v_dev, v_host)
    inner = int(np.nonzero(tree.leaf == 0)[0][0])
    for badnode in (inner, -1, tree.nnodes + 5):
        bad = nodes.copy()
        bad[17] = badnode
        with pytest.raises(ffmm.ArgumentError, match="not a real leaf"):
            plan.evaluate(tgt, ffmm.Potential, nodes=bad)
    with pytest.raises(ffmm.DimensionMismatch):
        plan.evaluate(tgt, ffmm.Potential, nodes=nodes[:-1])


def test_corrupt_trees_and_plan_data_are_refused_before_upload(points, tmp_path):
    """A host tree or plan data that violates the structure is refused with an
    ArgumentError and no plan; the device keeps working."""
    import ffmm
    pts, tree, rng = points
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=4)
    cache = ffmm.OperatorCache(m)
    bad = O.build_tree(pts, 60, *BOX)
    bad.colleagues = bad.colleagues.copy()
    bad.colleagues[100, 4] = tree.nnodes + 99
    with pytest.raises(ffmm.ArgumentError, match="colleague outside"):
        ffmm.DevicePlan.from_tree(m, bad, cache=cache, ppl=60)
    bad = O.build_tree(pts, 60, *BOX)
    bad.perm = bad.perm.copy()
    bad.perm[0] = bad.perm[1]
    with pytest.raises(ffmm.ArgumentError, match="permutation"):
        ffmm.DevicePlan.from_tree(m, bad, cache=cache, ppl=60)
    plan = ffmm.DevicePlan(m, pts, ppl=60, bbox=BOX, cache=cache)
    path = tmp_path / "plan.bin"
    plan.save(path)
    raw = bytearray(path.read_bytes())
    # the last int32 of the data is the last real leaf's id: point it past the nodes
    raw[-4:] = np.int32(plan.nnodes + 1000).tobytes()
    path.write_bytes(bytes(raw))
    with pytest.raises(ffmm.ArgumentError, match="node lists"):
        ffmm.DevicePlan.load(m, path, cache=cache)
    tau = rng.standard_normal(len(pts))
    assert np.all(np.isfinite(plan.matvec(tau).cpu().numpy()))
