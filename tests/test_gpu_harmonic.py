"""The harmonic methods on the GPU through the C ABI: SphericalHarmonics (Laplace3D,
Yukawa3D) and StokesHarmonics against the harmonic oracle (same tree, same operators,
FP64 throughout: agreement to rounding) and against FP64 direct sums (the method's
truncation error)."""
import numpy as np
import pytest

import harmonic as HO
from pointsets import mixed, uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_for(method, cache, pts, ppl, perm):
    from ffmm import Yukawa3D, StokesHarmonics
    kind = "stokes" if isinstance(method, StokesHarmonics) else "yukawa" if isinstance(method.kernel, Yukawa3D) else "laplace"
    kappa = method.kernel.k if kind == "yukawa" else 0.0
    W0 = cache.W0

    def level_ops(l, w):
        ops = cache.harmonic_operators(w)
        return ops, (1.0 if kind == "yukawa" else W0 / w)

    return HO.HarmonicFMM(kind, level_ops, pts, ppl, *BOX, kappa=kappa, perm=perm), kind, kappa


@pytest.mark.parametrize("kernel", ["laplace", "yukawa"])
def test_spherical_harmonics_vs_oracle_and_direct(kernel):
    import ffmm
    rng = np.random.default_rng(11)
    n = 5000
    pts = mixed(rng, n)
    tau = rng.standard_normal(n)
    tgt = uniform(rng, 400)
    K = ffmm.Laplace3D() if kernel == "laplace" else ffmm.Yukawa3D(5.0)
    m = ffmm.SphericalHarmonics(K, p=10)
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] {kernel} n={n} p=10", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    u = plan.matvec(tau, ffmm.PotentialGradient).cpu().numpy()
    v = plan.evaluate(tgt, ffmm.PotentialGradient).cpu().numpy()
    up = plan.evaluate_at_sources(ffmm.Potential).cpu().numpy()
    ug = plan.evaluate_at_sources(ffmm.Gradient).cpu().numpy()
    tree = plan.download_tree()
    f, kind, kappa = oracle_for(m, cache, pts, 40, tree.perm)
    assert f.tree.nnodes == plan.nnodes
    uo, vo = f.matvec(tau, output=2, targets=tgt)
    # device and oracle: the same algorithm in FP64
    assert rel(u, uo) < 1e-11 and rel(v, vo) < 1e-11
    assert np.allclose(up, u[:, 0], rtol=1e-13, atol=1e-13 * np.max(np.abs(up)))
    assert np.allclose(ug, u[:, 1:], rtol=1e-13, atol=1e-13 * np.max(np.abs(ug)))
    # multipoles and the leaves' locals against the oracle's
    M, Pl, Lc = plan.download_expansions()
    assert np.max(np.abs(M.T - f.multipoles)) <= 1e-11 * np.max(np.abs(f.multipoles))
    leaves = np.nonzero(tree.leaf)[0]
    lv = tree.level[leaves]
    order = leaves[np.argsort(lv, kind="stable")]
    assert np.max(np.abs(Lc.T - f.locals[order])) <= 1e-11 * np.max(np.abs(f.locals))
    # the truncation error of p = 10
    ud = HO.direct(kind, pts, tau, pts, 2, kappa=kappa, skip=np.arange(n))
    vd = HO.direct(kind, pts, tau, tgt, 2, kappa=kappa)
    assert rel(u[:, 0], ud[:, 0]) < 2e-5 and rel(v[:, 0], vd[:, 0]) < 2e-5
    assert rel(u[:, 1:], ud[:, 1:]) < 1e-4 and rel(v[:, 1:], vd[:, 1:]) < 1e-4


@pytest.mark.parametrize("nd", [3, 9])
def test_stokes_harmonics_vs_oracle_and_direct(nd):
    import ffmm
    rng = np.random.default_rng(12)
    n = 4000
    pts = mixed(rng, n)
    dens = rng.standard_normal((n, nd))
    if nd == 9:
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
    tgt = uniform(rng, 300)
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] stokes nd={nd} n={n} p=8", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    u = plan.matvec(dens, ffmm.Velocity).cpu().numpy()
    v = plan.evaluate(tgt, ffmm.Velocity).cpu().numpy()
    assert u.shape == (n, 3) and v.shape == (300, 3)
    f, kind, _ = oracle_for(m, cache, pts, 40, plan.download_tree().perm)
    uo, vo = f.matvec(dens, targets=tgt)
    assert rel(u, uo) < 1e-11 and rel(v, vo) < 1e-11
    assert rel(u, HO.direct("stokes", pts, dens, pts, skip=np.arange(n))) < 1e-3
    assert rel(v, HO.direct("stokes", pts, dens, tgt)) < 1e-3


@pytest.mark.parametrize("case", ["laplace-p3", "yukawa-p13", "stokes9-p16"])
def test_s2m_tiles_vs_oracle(case):
    """The S2M runs over tiles of S sources with 64 / S lanes per source, S set by p (the
    harmonics of a tile fit 8 KiB): p = 3 gives 32 sources of 2 lanes, p = 13 (odd: the
    order pairs (m, p - m) end without a middle order) 4 sources of 16 lanes with the
    Yukawa radial weights, p = 16 2 sources of 32 lanes with stresslets.  The multipoles
    (S2M, then M2M) against the oracle's, FP64 both."""
    import ffmm
    kind, p = case.split("-p")
    p = int(p)
    rng = np.random.default_rng(13)
    n = 3000
    pts = mixed(rng, n)
    if kind == "laplace":
        m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=p)
        dens = rng.standard_normal(n)
    elif kind == "yukawa":
        m = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=p)
        dens = rng.standard_normal(n)
    else:
        m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=p)
        dens = rng.standard_normal((n, 9))
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] s2m tiles {case} n={n}", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    plan.matvec(dens, ffmm.Velocity if kind == "stokes" else ffmm.Potential)
    M, _, _ = plan.download_expansions()
    tree = plan.download_tree()
    f, _, _ = oracle_for(m, cache, pts, 40, tree.perm)
    assert f.tree.nnodes == plan.nnodes
    Mo = f.upward(np.ascontiguousarray(dens[f.tree.perm]))
    assert np.max(np.abs(M.T - Mo)) <= 1e-11 * np.max(np.abs(Mo))


def test_laplace_dipoles_vs_oracle_and_direct():
    """SphericalHarmonics(Laplace3D()) with a charge and a dipole per point (rows (q, d),
    Dipole{3} densities, src/sources.jl:17-29): potential and gradient at the sources and
    at targets, and the multipoles, against the harmonic oracle (FP64 both: agreement to
    rounding) and against FP64 direct sums (the truncation of p = 10); a dipole-only
    density is the dipole part of the sum; Yukawa refuses dipoles."""
    import ffmm
    rng = np.random.default_rng(14)
    n = 4000
    pts = mixed(rng, n)
    dens = rng.standard_normal((n, 4))
    tgt = uniform(rng, 300)
    m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=10)
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] laplace dipoles n={n} p=10", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    u = plan.matvec(dens, ffmm.PotentialGradient).cpu().numpy()
    v = plan.evaluate(tgt, ffmm.PotentialGradient).cpu().numpy()
    M, _, _ = plan.download_expansions()
    f, _, _ = oracle_for(m, cache, pts, 40, plan.download_tree().perm)
    uo, vo = f.matvec(dens, output=2, targets=tgt)
    assert rel(u, uo) < 1e-11 and rel(v, vo) < 1e-11
    assert np.max(np.abs(M.T - f.multipoles)) <= 1e-11 * np.max(np.abs(f.multipoles))
    ud = HO.direct("laplace", pts, dens, pts, 2, skip=np.arange(n))
    vd = HO.direct("laplace", pts, dens, tgt, 2)
    assert rel(u, ud) < 3e-5 and rel(v, vd) < 3e-5
    # linearity in the density kind: the charge part plus the dipole part
    dq = dens.copy()
    dq[:, 1:] = 0.0
    dd = dens.copy()
    dd[:, 0] = 0.0
    uq = plan.matvec(dq, ffmm.PotentialGradient).cpu().numpy()
    ud2 = plan.matvec(dd, ffmm.PotentialGradient).cpu().numpy()
    uc = plan.matvec(dens[:, 0].copy(), ffmm.PotentialGradient).cpu().numpy()
    assert rel(uq + ud2, u) < 1e-13 and rel(uq, uc) < 1e-13
    my = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=6)
    py = ffmm.DevicePlan(my, pts, ppl=40, bbox=BOX, cache=ffmm.OperatorCache(my))
    with pytest.raises(ffmm.ArgumentError, match="one charge per point"):
        py.matvec(dens, ffmm.Potential)


@pytest.mark.parametrize("nd", [3, 9])
def test_stokes_pressure_vs_oracle_and_direct(nd):
    """The Stokes pressure (Pressure(), src/sources.jl:34-42): far field -s tr(grad Phi)
    (src/operators/harmonics_stokes.jl:348-351), near field the stokeslet and stresslet
    pressure kernels, at the sources and at targets, against the harmonic oracle (FP64
    both) and the FP64 direct sum; the velocity of the same plan is unchanged."""
    import ffmm
    rng = np.random.default_rng(15)
    n = 4000
    pts = mixed(rng, n)
    dens = rng.standard_normal((n, nd))
    if nd == 9:
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
    tgt = uniform(rng, 300)
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] stokes pressure nd={nd} n={n} p=8", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    u = plan.matvec(dens, ffmm.Pressure).cpu().numpy()
    v = plan.evaluate(tgt, ffmm.Pressure).cpu().numpy()
    assert u.shape == (n,) and v.shape == (300,)
    f, _, _ = oracle_for(m, cache, pts, 40, plan.download_tree().perm)
    uo, vo = f.matvec(dens, output=3, targets=tgt)
    assert rel(u, uo) < 1e-11 and rel(v, vo) < 1e-11
    assert rel(u, HO.direct("stokes", pts, dens, pts, 3, skip=np.arange(n))) < 1e-4
    assert rel(v, HO.direct("stokes", pts, dens, tgt, 3)) < 1e-4
    w = plan.evaluate_at_sources(ffmm.Velocity).cpu().numpy()
    assert rel(w, f.matvec(dens)) < 1e-11
    ml = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=4)
    pl = ffmm.DevicePlan(ml, pts, ppl=40, bbox=BOX, cache=ffmm.OperatorCache(ml))
    pl.build_expansions(dens[:, 0].copy())
    with pytest.raises(ffmm.ArgumentError, match="Stokes"):
        pl.evaluate_at_sources(ffmm.Pressure)


def test_harmonic_sorted_densities_linearity_and_errors():
    import ffmm
    rng = np.random.default_rng(13)
    n = 3000
    pts = uniform(rng, n)
    m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=6)
    plan = ffmm.DevicePlan(m, pts, ppl=30, bbox=BOX)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    ua = plan.matvec(a).cpu().numpy()
    ub = plan.matvec(b).cpu().numpy()
    uab = plan.matvec(2 * a - 3 * b).cpu().numpy()
    assert np.allclose(uab, 2 * ua - 3 * ub, rtol=1e-12, atol=1e-12 * np.max(np.abs(uab)))
    perm = plan.download_tree().perm
    us = plan.matvec(a[perm], sorted=True).cpu().numpy()
    assert np.array_equal(us, ua)
    with pytest.raises(ffmm.ArgumentError, match="dense class M2L"):
        ffmm.DevicePlan(m, pts, ppl=30, bbox=BOX, m2l="spectral")
    ms = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=4)
    ps = ffmm.DevicePlan(ms, pts, ppl=30, bbox=BOX)
    with pytest.raises(ffmm.ArgumentError, match="stokeslets"):
        ps.build_expansions(a)
    ps.build_expansions(rng.standard_normal((n, 3)))
    with pytest.raises(ffmm.ArgumentError, match="velocity"):
        ps.evaluate_at_sources(ffmm.Gradient)


@pytest.mark.parametrize("kind", ["yukawa", "stokes", "laplace-dipoles"])
def test_harmonic_mixed_form(kind):
    """The mixed form of the harmonic methods: the class M2L on FP32 MFMA, and the
    evaluation in FP32 (the folds of the local expansion, the near field with FP64 pair
    differences and Kahan-compensated FP32 sums folded into FP64 per tile); S2M, M2M and
    L2L FP64. Against the FP64 plan on the same tree (1e-6), at the sources and at
    targets, and against direct sums (the truncation of p = 10)."""
    import ffmm
    rng = np.random.default_rng(14)
    n = 5000
    pts = mixed(rng, n)
    tgt = uniform(rng, 400)
    if kind == "yukawa":
        m = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=10)
        dens, outs = rng.standard_normal(n), [ffmm.PotentialGradient, ffmm.Potential, ffmm.Gradient]
    elif kind == "stokes":
        m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=10)
        dens = rng.standard_normal((n, 9))
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
        outs = [ffmm.Velocity, ffmm.Pressure]
    else:
        m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=10)
        dens, outs = rng.standard_normal((n, 4)), [ffmm.PotentialGradient, ffmm.Potential, ffmm.Gradient]
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] mixed form {kind} p=10", flush=True)
    p64 = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache)
    p32 = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=cache, precision="mixed")
    for out in outs:
        u64 = p64.matvec(dens, out).cpu().numpy()
        u32 = p32.matvec(dens, out).cpu().numpy()
        v64 = p64.evaluate(tgt, out).cpu().numpy()
        v32 = p32.evaluate(tgt, out).cpu().numpy()
        print(f"[harmonic] mixed {kind} output {out}: {rel(u32, u64):.2e} {rel(v32, v64):.2e}", flush=True)
        assert rel(u32, u64) < 1e-6 and rel(v32, v64) < 1e-6
    rows = rng.choice(n, 300, replace=False)
    u32 = p32.matvec(dens, outs[0]).cpu().numpy()
    if kind == "stokes":
        ud = HO.direct("stokes", pts, dens, pts[rows], skip=rows)
        assert rel(u32[rows], ud) < 1e-4
        pd = HO.direct("stokes", pts, dens, pts[rows], 3, skip=rows)
        assert rel(p32.matvec(dens, ffmm.Pressure).cpu().numpy()[rows], pd) < 1e-3
    elif kind == "yukawa":
        ud = HO.direct("yukawa", pts, dens, pts[rows], skip=rows, kappa=5.0)
        assert rel(u32[rows, 0], np.ravel(ud)) < 2e-5     # the truncation error of p = 10
    else:
        ud = HO.direct("laplace", pts, dens, pts[rows], 2, skip=rows)
        assert rel(u32[rows], ud) < 3e-5


@pytest.mark.parametrize("nd", [3, 9])
def test_device_stokes_direct_sum(nd):
    """ffmm_direct_stokes (direct! for Stokeslet3D, stresslets as harmonics_stokes.jl)
    against the oracle's FP64 sum, with and without the self term."""
    import ffmm
    rng = np.random.default_rng(15)
    s, t = rng.random((6000, 3)), rng.random((400, 3))
    d = rng.standard_normal((6000, nd))
    if nd == 9:
        d[:, 6:] /= np.linalg.norm(d[:, 6:], axis=1)[:, None]
    K = ffmm.Stokeslet3D()
    print(f"[harmonic] stokes direct nd={nd}", flush=True)
    assert rel(ffmm.direct(K, s, d, t), HO.direct("stokes", s, d, t)) < 1e-13
    rows = np.arange(0, 6000, 17)
    assert rel(ffmm.direct_self(K, s, d, rows=rows), HO.direct("stokes", s, d, s[rows], skip=rows)) < 1e-13


@pytest.mark.parametrize("case", ["stokes-mixed", "yukawa-mixed", "laplace-f64", "cube-mixed"])
def test_class_m2l_chunks(case, monkeypatch):
    """The dense class M2L runs by target chunks (all classes of a chunk at once; in the
    mixed harmonic form, the fused gather + product kernel in tiles of 128 columns).
    A plan whose chunks are capped at 100 pairs (FFMM_M2L_CHUNK_PAIRS, read at plan
    creation) gives the result of the single-chunk plan.  In FP64 each pair's product is
    formed alone and each target sums its pairs in the same order (rel 1e-12).  The
    fused mixed kernel sums a target's consecutive products within a tile in FP32, and
    the tiles start elsewhere in other chunks; the cube method's mixed form runs its FP32
    class GEMMs in the library, which picks other kernels (another K order) for other
    column counts, and the DC2E pseudo-inverse amplifies that FP32 rounding: the mixed
    cases are held to the mixed form's 1e-6."""
    import ffmm
    rng = np.random.default_rng(16)
    n = 6000
    pts = mixed(rng, n)
    kw = dict(ppl=40, bbox=BOX)
    if case == "stokes-mixed":
        m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=10)
        dens = rng.standard_normal((n, 9))
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
        out, kw["precision"] = ffmm.Velocity, "mixed"
    elif case == "yukawa-mixed":
        m = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=10)
        dens, out, kw["precision"] = rng.standard_normal(n), ffmm.PotentialGradient, "mixed"
    elif case == "laplace-f64":
        m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=8)
        dens, out = rng.standard_normal(n), ffmm.PotentialGradient
    else:
        m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
        dens, out = rng.standard_normal(n), ffmm.PotentialGradient
        kw.update(precision="mixed", m2l="dense")
    cache = ffmm.OperatorCache(m)
    print(f"[harmonic] class M2L chunks {case}", flush=True)
    one = ffmm.DevicePlan(m, pts, cache=cache, **kw).matvec(dens, out).cpu().numpy()
    monkeypatch.setenv("FFMM_M2L_CHUNK_PAIRS", "100")
    many = ffmm.DevicePlan(m, pts, cache=cache, **kw).matvec(dens, out).cpu().numpy()
    assert rel(many, one) < (1e-12 if case == "laplace-f64" else 1e-6)


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("n", [1, 2, 37])
def test_dipoles_and_pressure_small_cases(n, precision):
    """Laplace dipoles and the Stokes pressure with 1, 2 and 37 points (one leaf: every
    interaction is in the near field) against the FP64 direct sums, at targets and at
    the sources (a point skips itself; one point alone gets exact zeros). The mixed
    form's FP32 near field is held to 1e-6."""
    import ffmm
    rng = np.random.default_rng(16 + n)
    pts = uniform(rng, n)
    tgt = uniform(rng, 20)
    tol = 1e-12 if precision == "f64" else 1e-6
    print(f"[harmonic] small dipole / pressure n={n} {precision}", flush=True)
    m = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=6)
    plan = ffmm.DevicePlan(m, pts, ppl=40, bbox=BOX, cache=ffmm.OperatorCache(m), precision=precision)
    dens = rng.standard_normal((n, 4))
    u = plan.matvec(dens, ffmm.PotentialGradient).cpu().numpy().reshape(n, 4)
    v = plan.evaluate(tgt, ffmm.PotentialGradient).cpu().numpy()
    vd = HO.direct("laplace", pts, dens, tgt, 2)
    assert rel(v, vd) < tol
    if n > 1:
        assert rel(u, HO.direct("laplace", pts, dens, pts, 2, skip=np.arange(n))) < tol
    else:
        assert np.all(u == 0.0)
    ms = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=4)
    ps = ffmm.DevicePlan(ms, pts, ppl=40, bbox=BOX, cache=ffmm.OperatorCache(ms), precision=precision)
    ds = rng.standard_normal((n, 9))
    ds[:, 6:] /= np.linalg.norm(ds[:, 6:], axis=1)[:, None]
    pu = ps.matvec(ds, ffmm.Pressure).cpu().numpy().reshape(n)
    pv = ps.evaluate(tgt, ffmm.Pressure).cpu().numpy()
    assert rel(pv, HO.direct("stokes", pts, ds, tgt, 3)) < tol
    if n > 1:
        assert rel(pu, HO.direct("stokes", pts, ds, pts, 3, skip=np.arange(n))) < tol
    else:
        assert np.all(pu == 0.0)
    vu = ps.matvec(ds, ffmm.Velocity).cpu().numpy().reshape(n, 3)
    vv = ps.evaluate(tgt, ffmm.Velocity).cpu().numpy()
    assert rel(vv, HO.direct("stokes", pts, ds, tgt)) < tol
    if n > 1:
        assert rel(vu, HO.direct("stokes", pts, ds, pts, skip=np.arange(n))) < tol
    else:
        assert np.all(vu == 0.0)
