"""The dense class M2L with low-rank class operators (DevicePlan(...; tol),
src/device_passes.jl:14-27, 570-576) on the GPU, as the reference tests it
(test/test_device_passes.jl:33-50): Laplace, p = 8, tol = 1e-8, 4,000 uniform points,
ppl = 40; the ranks lie below n/2, the result is within 1e-7 (max-relative) of the
FP64 FMM with the dense classes (the oracle) and within 1e-6 of the direct sum."""
import numpy as np
import pytest

import oracle as O
from pointsets import uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def mrd(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def case():
    import ffmm
    rng = np.random.default_rng(33)
    pts = uniform(rng, 4000)
    tau = rng.standard_normal(4000)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    ref = O.FMM(cache.host_operators(), pts, 40, *BOX, perm=plan.download_tree().perm).matvec(tau)
    return ffmm, m, cache, pts, tau, plan, ref


@pytest.mark.parametrize("precision,tol_oracle", [("f64", 1e-7), ("mixed", 2e-6)])
def test_lowrank_class_m2l(case, precision, tol_oracle):
    ffmm, m, cache, pts, tau, _, ref = case
    print(f"[lowrank] {precision} tol=1e-8", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, precision=precision, tol=1e-8)
    assert plan.class_ranks is not None and np.all(plan.class_ranks < m.nsurface // 2)
    u = plan.matvec(tau).cpu().numpy()
    assert mrd(u, ref) < tol_oracle
    rows = np.arange(0, len(pts), 9)
    ud = O.direct_self(pts, tau, rows=rows)
    assert mrd(u[rows], ud) < 1e-6


def test_lowrank_switch_and_refusals(case):
    """set_tol(0) gives the dense representatives back (bitwise the plain plan's result);
    tol > 0 with the spectral M2L or a harmonic method, and the fast/eltype keywords the
    library does not offer, are ArgumentErrors."""
    ffmm, m, cache, pts, tau, plan, _ = case
    u_dense = plan.matvec(tau).cpu().numpy()
    plan.set_tol(1e-6)
    u_lr = plan.matvec(tau).cpu().numpy()
    assert 0 < mrd(u_lr, u_dense) < 1e-5
    plan.set_tol(0.0)
    assert np.array_equal(plan.matvec(tau).cpu().numpy(), u_dense)
    with pytest.raises(ffmm.ArgumentError, match="dense class M2L"):
        ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, m2l="spectral", tol=1e-8)
    with pytest.raises(ffmm.ArgumentError, match="fast"):
        ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, fast=True)
    with pytest.raises(ffmm.ArgumentError, match="eltype must hold"):
        ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, eltype=np.float32)
    h = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=6)
    with pytest.raises(ffmm.ArgumentError, match="dense class M2L"):
        ffmm.DevicePlan(h, pts, ppl=40, bbox=BOX, tol=1e-8)
