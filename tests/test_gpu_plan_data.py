"""Plan data (``ffmm_plan_save`` / ``ffmm_plan_load``; no reference counterpart): the tree
of a plan written out and a plan made from it without the tree build, for warm restarts
of a serving cache's resident plans.  The restored plan has the same tree (every array
equal) and gives bitwise the same results; the cube method (spectral and dense),
a harmonic method and groups; a cache with other operators and corrupt data are refused."""
import numpy as np
import pytest

from pointsets import mixed

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def _same_tree(a, b):
    ta, tb = a.download_tree(), b.download_tree()
    for k in ta.__dataclass_fields__:
        assert np.array_equal(getattr(ta, k), getattr(tb, k)), k


@pytest.mark.parametrize("form", [("mixed", "spectral"), ("f64", "dense")])
def test_cube_plan_round_trip(tmp_path, form):
    import ffmm
    rng = np.random.default_rng(41)
    pts = mixed(rng, 20000)
    tau = rng.standard_normal(len(pts))
    tgt = rng.random((3000, 3))
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[plan data] cube {form}", flush=True)
    a = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, precision=form[0], m2l=form[1])
    path = tmp_path / "plan.bin"
    a.save(path)
    b = ffmm.DevicePlan.load(m, path, cache=cache)
    assert (b.nlevels, b.nnodes, b.N) == (a.nlevels, a.nnodes, a.N)
    assert np.array_equal(b.level_pairs, a.level_pairs)
    _same_tree(a, b)
    ua, ub = a.matvec(tau).cpu().numpy(), b.matvec(tau).cpu().numpy()
    assert np.array_equal(ua, ub)
    va = a.evaluate(tgt, ffmm.PotentialGradient).cpu().numpy()
    vb = b.evaluate(tgt, ffmm.PotentialGradient).cpu().numpy()
    assert np.array_equal(va, vb)
    # a fresh cache of the same method takes its reference width from the data
    c = ffmm.DevicePlan.load(m, path, cache=ffmm.OperatorCache(m))
    assert np.array_equal(c.matvec(tau).cpu().numpy(), ua)


def test_harmonic_and_groups_round_trip(tmp_path):
    import ffmm
    rng = np.random.default_rng(42)
    xs1, xs2 = mixed(rng, 6000), rng.random((1500, 3))
    d1 = rng.standard_normal((len(xs1), 3))
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=6)
    cache = ffmm.OperatorCache(m)
    print("[plan data] stokes groups", flush=True)
    a = ffmm.DevicePlan(m, (xs1, xs2), ppl=40, bbox=BOX, cache=cache, precision="mixed")
    path = tmp_path / "plan.bin"
    a.save(path)
    b = ffmm.DevicePlan.load(m, path, cache=cache)
    assert b.group_sizes == a.group_sizes
    ra = a.matvec((d1, None), ffmm.Velocity)
    rb = b.matvec((d1, None), ffmm.Velocity)
    for x, y in zip(ra, rb):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())


def test_plan_data_refused(tmp_path):
    import ffmm
    rng = np.random.default_rng(43)
    pts = mixed(rng, 5000)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    a = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    path = tmp_path / "plan.bin"
    a.save(path)
    print("[plan data] refusals", flush=True)
    m6 = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=6)
    with pytest.raises(ffmm.ArgumentError, match="another operator cache"):
        ffmm.DevicePlan.load(m6, path, cache=ffmm.OperatorCache(m6))
    raw = path.read_bytes()
    bad = tmp_path / "bad.bin"
    bad.write_bytes(raw[:-100])
    with pytest.raises(ffmm.ArgumentError, match="truncated|wrong length|trailing"):
        ffmm.DevicePlan.load(m, bad, cache=cache)
    bad.write_bytes(raw + b"\0" * 8)
    with pytest.raises(ffmm.ArgumentError, match="trailing"):
        ffmm.DevicePlan.load(m, bad, cache=cache)
    bad.write_bytes(b"not a plan\n")
    with pytest.raises(ffmm.ArgumentError, match="not a plan file"):
        ffmm.DevicePlan.load(m, bad, cache=cache)


def test_server_warm_restart(tmp_path):
    """A server's resident plan exported (``ffmm_server_export``) and imported into a new
    server (``ffmm_server_import``): the request with that geometry key builds no plan and
    gives bitwise the first server's result; the data equals DevicePlan's plan data of
    the same geometry (the tree build is deterministic); a resident key is not replaced."""
    import ffmm
    from ffmm.serving import Server
    rng = np.random.default_rng(44)
    pts = mixed(rng, 20000)
    tau = rng.standard_normal(len(pts))
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    kw = dict(geometry=5, ppl=40, bbox=BOX, precision="mixed", m2l="spectral")
    print("[plan data] server warm restart", flush=True)
    with Server(nslots=2, hbm_budget=8 << 30) as a:
        u, _, _ = a.wait(a.submit(cache, pts, tau, **kw))
        data = a.export_plan(cache, 5)
        with pytest.raises(ffmm.ArgumentError, match="no resident plan"):
            a.export_plan(cache, 6)
    with Server(nslots=2, hbm_budget=8 << 30) as b:
        b.import_plan(cache, 5, data)
        v, _, _ = b.wait(b.submit(cache, pts, tau, **kw))
        st = b.stats()
        assert st["plans_built"] == 0 and st["plans_reused"] == 1
        assert np.array_equal(u, v)
        with pytest.raises(ffmm.ArgumentError, match="resident"):
            b.import_plan(cache, 5, data)
    p = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, precision="mixed", m2l="spectral")
    path = tmp_path / "plan.bin"
    p.save(path)
    with open(path, "rb") as f:
        f.readline()
        f.readline()
        assert f.read() == data
