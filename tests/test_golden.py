"""The committed fixtures of tests/golden/ (written by tests/golden/make_fixtures.py):
the oracle reproduces them on the CPU; the device tree, its lists and the device FMM
reproduce them on the GPU."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = (np.zeros(3), 1.0)
FIELDS = ["level_start", "cell", "level", "parent", "child_first", "plo", "phi", "leaf", "ghost", "xlist",
          "has_source", "colleagues", "point_node"]
TREES = ["tree_octants.npz", "tree_mixed300.npz"]


def load(name):
    return dict(np.load(os.path.join(HERE, name)))


def assert_tree_matches(t, fx, lists):
    for f in FIELDS:
        assert np.array_equal(np.asarray(getattr(t, f)), fx[f]), f
    for l in range(int(fx["nlevels"])):
        ptr, src, far = lists(l)
        assert np.array_equal(ptr, fx[f"ptr{l}"]) and np.array_equal(src, fx[f"src{l}"])
        assert np.array_equal(far, fx[f"far{l}"])


@pytest.mark.parametrize("name", TREES)
def test_oracle_tree_matches_fixture(name):
    fx = load(name)
    t = O.build_tree(fx["points"], int(fx["ppl"]), *BOX)
    assert_tree_matches(t, fx, lambda l: t.lists[l])


def test_oracle_fmm_matches_direct_fixture():
    from ffmm import operators
    fx = load("fmm_mixed2000.npz")
    f = O.FMM(operators.level_operators(1.0, p=8, with_tables=False), fx["points"], 30, *BOX)
    u, v = f.matvec(fx["tau"], targets=fx["targets"])
    assert np.max(np.abs(u - fx["u_sources"])) / np.max(np.abs(fx["u_sources"])) < 1e-6
    assert np.max(np.abs(v - fx["u_targets"])) / np.max(np.abs(fx["u_targets"])) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", TREES)
def test_device_tree_matches_fixture(name):
    import ffmm
    fx = load(name)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=4)
    plan = ffmm.DevicePlan(m, fx["points"], ppl=int(fx["ppl"]), bbox=BOX)
    assert_tree_matches(plan.download_tree(), fx, plan.download_lists)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,m2l", [("f64", "dense"), ("f64", "spectral"), ("mixed", "spectral"), ("mixed", "dense")])
def test_device_fmm_matches_direct_fixture(precision, m2l):
    import ffmm
    fx = load("fmm_mixed2000.npz")
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    plan = ffmm.DevicePlan(m, fx["points"], ppl=30, bbox=BOX, precision=precision, m2l=m2l)
    u = plan.matvec(fx["tau"]).cpu().numpy()
    v = plan.evaluate(fx["targets"], ffmm.PotentialGradient).cpu().numpy()
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))   # noqa: E731
    assert rel(u, fx["u_sources"]) < 1e-6
    assert rel(v[:, 0], fx["u_targets"]) < 1e-6
    g = fx["g_targets"]
    assert np.max(np.linalg.norm(v[:, 1:] - g, axis=1)) / np.max(np.linalg.norm(g, axis=1)) < 1e-5
