"""The Stokes front end (src/stokes_api.jl): StokesFMMPlan for a fixed geometry,
stokes_fmm_eval! at the sources (self) or at targets, stokes_fmm_eval_src_trg!, the
one-shot StokesFMMKernel, weights, and the error phrases; against the device FP64
direct Stokes sum (itself checked against the oracle) and against the plan's own
device path."""
import numpy as np
import pytest

import harmonic as HO
from pointsets import mixed, uniform

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def test_stokes_front_end():
    import ffmm
    rng = np.random.default_rng(51)
    n = 4000
    pts, tgt = mixed(rng, n), uniform(rng, 500)
    f, g = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    nu = rng.standard_normal((n, 3))
    nu /= np.linalg.norm(nu, axis=1)[:, None]
    w = rng.random(n) + 0.5
    K = ffmm.Stokeslet3D()
    dens = np.hstack([w[:, None] * f, w[:, None] * g, nu])
    print("[stokes api] self plan p=10", flush=True)
    plan = ffmm.StokesFMMPlan(pts, p=10, ppl=40)
    v = np.empty((n, 3))
    ffmm.stokes_fmm_eval_(v, plan, stokeslets=f, stresslets=g, stressvecs=nu, weights=w)
    rows = np.arange(0, n, 13)
    assert rel(v[rows], ffmm.direct_self(K, pts, dens, rows=rows)) < 1e-3
    # the same through the device plan directly: the densities are the (N, 9) rows
    assert np.array_equal(v, plan.plan.matvec(dens, ffmm.Velocity).cpu().numpy())
    # a plan with targets; one build, two evaluations
    print("[stokes api] plan with targets", flush=True)
    pt = ffmm.StokesFMMPlan(pts, targets=tgt, p=10, ppl=40)
    vt = ffmm.stokes_fmm_eval_(None, pt, stokeslets=f)
    assert rel(vt, ffmm.direct(K, pts, f, tgt)) < 1e-3
    vs, vt2 = ffmm.stokes_fmm_eval_src_trg_(None, None, pt, stokeslets=f)
    assert np.array_equal(vt2, vt)
    assert rel(vs[rows], ffmm.direct_self(K, pts, f, rows=rows)) < 1e-3
    # the pressure (output=Pressure, src/stokes_api.jl:100-104), against the device plan's
    # and the oracle's direct sum
    pr = ffmm.stokes_fmm_eval_(None, plan, stokeslets=f, stresslets=g, stressvecs=nu, weights=w,
                               output=ffmm.Pressure)
    assert pr.shape == (n,)
    assert np.array_equal(pr, plan.plan.matvec(dens, ffmm.Pressure).cpu().numpy())
    assert rel(pr, HO.direct("stokes", pts, dens, pts, 3, skip=np.arange(n))) < 1e-4
    # one-shot kernel, D x N input
    u = ffmm.StokesFMMKernel(pts.T.copy(), stokeslets=f, p=10, ppl=40)
    assert rel(u[rows], ffmm.direct_self(K, pts, f, rows=rows)) < 1e-3
    # errors
    with pytest.raises(ffmm.ArgumentError, match="given together"):
        ffmm.stokes_fmm_eval_(None, plan, stresslets=g)
    with pytest.raises(ffmm.ArgumentError, match="at least one"):
        ffmm.stokes_fmm_eval_(None, plan)
    with pytest.raises(ffmm.DimensionMismatch, match="expected 4000"):
        ffmm.stokes_fmm_eval_(None, plan, stokeslets=f[:-1])
    with pytest.raises(ffmm.ArgumentError, match="no targets"):
        ffmm.stokes_fmm_eval_src_trg_(None, None, plan, stokeslets=f)
    with pytest.raises(ffmm.ArgumentError, match="velocity"):
        ffmm.stokes_fmm_eval_(None, plan, stokeslets=f, output=ffmm.Gradient)
    with pytest.raises(ffmm.ArgumentError, match="not built"):
        ffmm.StokesFMMPlan(pts, method="cube")
