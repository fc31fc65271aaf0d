"""Groups of points in one plan (``DevicePlan(method, groups::Tuple)``,
src/device_passes.jl:113-134; ``device_matvec!`` with one output per group, 619-640):
one tree over all groups, densities per group (``None``: targets only), outputs per
group in each group's original order.  Checked against the single-group plan of the
concatenated points (the same tree and arithmetic, so bitwise equal) and against FP64
direct sums; sorted inputs in each group's tree order; the error phrases."""
import numpy as np
import pytest

import oracle as O
from pointsets import mixed, uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def mrd(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim == 2:
        return np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1))
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_source_and_target_groups():
    import ffmm
    rng = np.random.default_rng(31)
    xs1, xs2 = mixed(rng, 5000), uniform(rng, 1500)
    tau1 = rng.standard_normal(len(xs1))
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print("[groups] laplace sources + targets", flush=True)
    plan = ffmm.DevicePlan(m, (xs1, xs2), ppl=30, bbox=BOX, cache=cache)
    assert plan.ngroups == 2 and plan.N == len(xs1) + len(xs2)
    u1, g2 = plan.matvec((tau1, None), (ffmm.Potential, ffmm.PotentialGradient))
    u1, g2 = u1.cpu().numpy(), g2.cpu().numpy()
    assert u1.shape == (len(xs1),) and g2.shape == (len(xs2), 4)
    # the concatenated single-group plan with zero densities on the targets
    one = ffmm.DevicePlan(m, np.vstack([xs1, xs2]), ppl=30, bbox=BOX, cache=cache)
    tau = np.concatenate([tau1, np.zeros(len(xs2))])
    assert np.array_equal(one.matvec(tau).cpu().numpy()[:len(xs1)], u1)
    assert np.array_equal(one.matvec(tau, ffmm.PotentialGradient).cpu().numpy()[len(xs1):], g2)
    # FP64 direct sums
    assert mrd(u1, O.direct_self(xs1, tau1)) < 1e-6
    assert mrd(g2[:, 0], O.direct(xs1, tau1, xs2)) < 1e-6
    assert mrd(g2[:, 1:], O.direct(xs1, tau1, xs2, grad=True)) < 1e-5
    # an output not wanted
    r = plan.matvec((tau1, None), (ffmm.Potential, None))
    assert r[1] is None and np.array_equal(r[0].cpu().numpy(), u1)
    # densities in both groups: the sum of the two single-group solves (linearity)
    tau2 = rng.standard_normal(len(xs2))
    a1, a2 = plan.matvec((tau1, tau2))
    b1, b2 = plan.matvec((None, tau2))
    assert mrd(a1.cpu().numpy(), u1 + b1.cpu().numpy()) < 1e-12
    # sorted: each group's densities and outputs in its tree order
    perm = plan.download_tree().perm
    ord1 = perm[perm < len(xs1)]
    ord2 = perm[perm >= len(xs1)] - len(xs1)
    s1, s2 = plan.matvec((tau1[ord1], tau2[ord2]), sorted=True)
    assert np.array_equal(s1.cpu().numpy(), a1.cpu().numpy()[ord1])
    assert np.array_equal(s2.cpu().numpy(), a2.cpu().numpy()[ord2])
    # the error phrases of the reference
    with pytest.raises(ffmm.DimensionMismatch, match="expected 2 groups"):
        plan.matvec((tau1,))
    with pytest.raises(ffmm.DimensionMismatch, match="entries in group 2"):
        plan.matvec((tau1, tau2[:-1]))
    with pytest.raises(ffmm.DimensionMismatch, match="expected 2 outputs"):
        plan.matvec((tau1, None), (ffmm.Potential,) * 3)


def test_host_mirror_groups():
    """FMM / build_expansions / evaluate_at_sources(fmm, state, g) with groups
    (src/fmm.jl:264-298; src/evaluate.jl:104-132)."""
    import ffmm
    rng = np.random.default_rng(32)
    xs1, xs2 = mixed(rng, 3000), uniform(rng, 800)
    tau1 = rng.standard_normal(len(xs1))
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    print("[groups] host mirror", flush=True)
    f = ffmm.FMM(m, (xs1, xs2), ppl=30, bbox=BOX)
    st = ffmm.build_expansions(f, (xs1, xs2), (tau1, None))
    u2 = ffmm.evaluate_at_sources(f, st, 2)
    assert u2.shape == (len(xs2),)
    assert mrd(u2, O.direct(xs1, tau1, xs2)) < 1e-6
    assert mrd(ffmm.evaluate_at_sources(f, st, 1), O.direct_self(xs1, tau1)) < 1e-6
    with pytest.raises(ffmm.ArgumentError, match="group 3 of 2"):
        ffmm.evaluate_at_sources(f, st, 3)
    with pytest.raises(ffmm.DimensionMismatch, match="densities in group 1"):
        ffmm.build_expansions(f, (xs1, xs2), (tau1[:-1], None))


def test_stokes_groups_of_different_densities():
    """A group of stokeslets (N x 3) beside a group of stokeslets with stresslets
    (N x 9): joined with a zero stresslet part on the first, as one plan with both."""
    import ffmm
    rng = np.random.default_rng(33)
    xs1, xs2 = mixed(rng, 2500), mixed(rng, 2000)
    f1 = rng.standard_normal((len(xs1), 3))
    d2 = rng.standard_normal((len(xs2), 9))
    d2[:, 6:] /= np.linalg.norm(d2[:, 6:], axis=1)[:, None]
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=6)
    cache = ffmm.OperatorCache(m)
    print("[groups] stokes 3 + 9", flush=True)
    plan = ffmm.DevicePlan(m, (xs1, xs2), ppl=40, bbox=BOX, cache=cache)
    v1, v2 = plan.matvec((f1, d2), ffmm.Velocity)
    one = ffmm.DevicePlan(m, np.vstack([xs1, xs2]), ppl=40, bbox=BOX, cache=cache)
    d = np.zeros((len(xs1) + len(xs2), 9))
    d[:len(xs1), :3] = f1
    d[len(xs1):] = d2
    w = one.matvec(d, ffmm.Velocity).cpu().numpy()
    assert np.array_equal(v1.cpu().numpy(), w[:len(xs1)]) and np.array_equal(v2.cpu().numpy(), w[len(xs1):])


def test_laplace_groups_of_charges_and_dipoles():
    """A group of charges (N) beside a group of charges and dipoles (N x 4, rows (q, d)):
    joined with a zero dipole on the first, as one plan with both (the reference's
    groups of different density types, src/device_passes.jl:113-134)."""
    import ffmm
    rng = np.random.default_rng(34)
    xs1, xs2 = mixed(rng, 2500), mixed(rng, 2000)
    q1 = rng.standard_normal(len(xs1))
    d2 = rng.standard_normal((len(xs2), 4))
    m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=6)
    cache = ffmm.OperatorCache(m)
    print("[groups] laplace 1 + 4", flush=True)
    plan = ffmm.DevicePlan(m, (xs1, xs2), ppl=40, bbox=BOX, cache=cache)
    v1, v2 = plan.matvec((q1, d2), ffmm.PotentialGradient)
    one = ffmm.DevicePlan(m, np.vstack([xs1, xs2]), ppl=40, bbox=BOX, cache=cache)
    d = np.zeros((len(xs1) + len(xs2), 4))
    d[:len(xs1), 0] = q1
    d[len(xs1):] = d2
    w = one.matvec(d, ffmm.PotentialGradient).cpu().numpy()
    assert np.array_equal(v1.cpu().numpy(), w[:len(xs1)]) and np.array_equal(v2.cpu().numpy(), w[len(xs1):])
