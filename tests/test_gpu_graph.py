"""Captured pass sequences (ffmm_plan_capture / ffmm_plan_replay) on the GPU: a replay
gives the eager matvec's result bitwise, replay after replay, with eager solves of other
densities (and on another plan) between them; for the cube method (spectral mixed,
dense FP64), a Yukawa harmonic plan and the Stokes stresslets.  A new tolerance drops the
graph; a replay without a capture is refused."""
import numpy as np
import pytest

from pointsets import mixed

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def _check_replays(plan, dens_list, output, other=None):
    for k, d in enumerate(dens_list):
        eager = plan.matvec(d, output).cpu().numpy()
        if other is not None:
            other.matvec(np.ones(other.N))       # eager work on another plan between replays
        g = plan.replay(d).cpu().numpy()
        assert np.array_equal(g, eager), k
        g2 = plan.replay(d).cpu().numpy()        # twice in a row, same result
        assert np.array_equal(g2, eager), k


@pytest.mark.parametrize("precision,m2l", [("mixed", "spectral"), ("f64", "dense")])
def test_graph_replay_cube(precision, m2l):
    import ffmm
    rng = np.random.default_rng(50)
    pts = mixed(rng, 20000)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[graph] cube {precision} {m2l}", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, precision=precision, m2l=m2l)
    other = ffmm.DevicePlan(m, mixed(rng, 3000), ppl=40, bbox=BOX, cache=cache)
    plan.capture(ffmm.Potential)
    _check_replays(plan, [rng.standard_normal(len(pts)) for _ in range(3)], ffmm.Potential, other)
    # potential + gradient, captured anew
    plan.capture(ffmm.PotentialGradient)
    _check_replays(plan, [rng.standard_normal(len(pts)) for _ in range(2)], ffmm.PotentialGradient)


def test_graph_replay_harmonic_and_stokes():
    import ffmm
    rng = np.random.default_rng(51)
    pts = mixed(rng, 8000)
    y = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=8)
    print("[graph] yukawa p=8 mixed", flush=True)
    plan = ffmm.DevicePlan(y, pts, ppl=40, bbox=BOX, precision="mixed")
    plan.capture(ffmm.Potential)
    _check_replays(plan, [rng.standard_normal(len(pts)) for _ in range(2)], ffmm.Potential)
    s = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=6)
    print("[graph] stokes p=6 stresslets", flush=True)
    plan = ffmm.DevicePlan(s, pts, ppl=40, bbox=BOX)
    plan.capture(ffmm.Velocity, nd=9)
    dens = [np.concatenate([rng.standard_normal((len(pts), 6)), rng.standard_normal((len(pts), 3))], 1)
            for _ in range(2)]
    _check_replays(plan, dens, ffmm.Velocity)


def test_graph_dropped_and_refused():
    import ffmm
    rng = np.random.default_rng(52)
    pts = mixed(rng, 5000)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=6)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX)
    with pytest.raises(ffmm.ArgumentError, match="capture the plan first"):
        plan.replay(np.ones(len(pts)))
    plan.capture()
    tau = rng.standard_normal(len(pts))
    assert np.array_equal(plan.replay(tau).cpu().numpy(), plan.matvec(tau).cpu().numpy())
    plan.set_tol(1e-6)                       # the buffers may move: the graph is dropped
    with pytest.raises(ffmm.ArgumentError, match="capture the plan first"):
        plan.replay(tau)
    plan.capture()
    assert np.array_equal(plan.replay(tau).cpu().numpy(), plan.matvec(tau).cpu().numpy())
    with pytest.raises(ffmm.DimensionMismatch):
        plan.replay(tau[:-1])
