"""The device passes and evaluation against the oracle (the restated host FMM on the
same tree and operators) and against the FP64 direct sum.  Tolerances follow the
reference's device-vs-CPU checks (test/test_device_passes.jl:6-31: matvec 1e-11,
locals 100x that; vs direct 1e-6 at p = 8, test/test_3d.jl:155-163)."""
import numpy as np
import pytest

import oracle as O
from pointsets import mixed, uniform, sphere

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def mrd(a, b):
    """max-norm relative difference (test/testutils.jl:9)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim == 2:
        return np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1))
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def setup():
    import ffmm
    rng = np.random.default_rng(45)
    N = 6000
    pts = mixed(rng, N)
    tau = rng.standard_normal(N)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    plan = ffmm.DevicePlan(m, pts, ppl=30, bbox=BOX, cache=cache)
    ops = cache.host_operators()
    ref = O.FMM(ops, pts, 30, BOX[0], BOX[1])
    u_ref = ref.matvec(tau)
    return dict(ffmm=ffmm, pts=pts, tau=tau, plan=plan, ref=ref, u_ref=u_ref, rng=rng, m=m, cache=cache)


def test_matvec_f64_dense_vs_oracle_and_direct(setup):
    s = setup
    print(f"[fmm] f64 dense N={len(s['pts'])} levels={s['plan'].nlevels}", flush=True)
    assert s["plan"].nlevels >= 5
    u = s["plan"].matvec(s["tau"]).cpu().numpy()
    assert mrd(u, s["u_ref"]) < 1e-11
    rows = s["rng"].choice(len(u), 400, replace=False)
    ud = O.direct_self(s["pts"], s["tau"], rows=rows)
    assert mrd(u[rows], ud) < 1e-6


@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("mixed", 2e-6)])
def test_s2m_uc2e_nodes_beyond_the_lds_stage(precision, tol):
    """ppl = 300 on 12,000 uniform points: the 64 boxes of level 2 are leaves of ~190
    sources, one group of 64 childless nodes holding more sources than the fused S2M +
    UC2E kernel stages in LDS (4,096 in FP32, 2,048 in FP64), so it reads them from global
    memory; against the oracle."""
    import ffmm
    rng = np.random.default_rng(46)
    N, ppl = 12_000, 300
    pts = uniform(rng, N)
    tau = rng.standard_normal(N)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[fmm] fused S2M beyond the LDS stage, {precision}, ppl={ppl}", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=ppl, bbox=BOX, cache=cache, precision=precision)
    # the childless sourced nodes of levels 2 and below, in node order, by 64: some group
    # holds more sources than either stage
    t = plan.download_tree()
    idx = np.nonzero((t.child_first < 0) & (t.phi >= t.plo))[0]
    cnt = (t.phi - t.plo + 1)[idx[idx >= t.level_start[2]]]
    assert np.add.reduceat(cnt, np.arange(0, len(cnt), 64)).max() > 4096
    u = plan.matvec(tau).cpu().numpy()
    ref = O.FMM(cache.host_operators(), pts, ppl, BOX[0], BOX[1])
    assert mrd(u, ref.matvec(tau)) < tol


@pytest.mark.parametrize("precision,tol_oracle,tol_direct", [("f64", 1e-11, 1e-6), ("mixed", 2e-6, 2e-6)])
def test_matvec_spectral_vs_oracle_and_direct(setup, precision, tol_oracle, tol_direct):
    """The spectral M2L (batched DFTs + per-pair Hadamard products, the reference CPU's
    form for the cube method, src/operators/spectral_m2l.jl:328-451) against the
    oracle's dense class translations: equal to rounding in FP64 (spectral equals dense
    to rtol 1e-12, test/test_cubeki.jl:43-68); FP32 spectra in the mixed form."""
    s = setup
    ffmm = s["ffmm"]
    print(f"[fmm] spectral {precision}", flush=True)
    plan = ffmm.DevicePlan(s["m"], s["pts"], ppl=30, bbox=BOX, cache=s["cache"], m2l="spectral", precision=precision)
    u = plan.matvec(s["tau"]).cpu().numpy()
    assert mrd(u, s["u_ref"]) < tol_oracle
    rows = s["rng"].choice(len(u), 400, replace=False)
    assert mrd(u[rows], O.direct_self(s["pts"], s["tau"], rows=rows)) < tol_direct
    # the plocals of the spectral and dense device plans
    s["plan"].build_expansions(s["tau"])
    _, Pd, _ = s["plan"].download_expansions()
    _, Ps, _ = plan.download_expansions()
    assert mrd(Ps.ravel(), Pd.ravel()) < (1e-12 if precision == "f64" else 1e-6)


def test_matvec_dense_fp32_classes_vs_oracle_and_direct(setup):
    """The mixed form of the dense class M2L: FP32 gather, FP32 GEMM (sgemm on FP32
    MFMA), scatter into the FP64 partial locals; pseudo-inverses and L2L FP64."""
    s = setup
    ffmm = s["ffmm"]
    print("[fmm] dense mixed", flush=True)
    plan = ffmm.DevicePlan(s["m"], s["pts"], ppl=30, bbox=BOX, cache=s["cache"], m2l="dense", precision="mixed")
    u = plan.matvec(s["tau"]).cpu().numpy()
    assert mrd(u, s["u_ref"]) < 2e-6
    rows = s["rng"].choice(len(u), 400, replace=False)
    assert mrd(u[rows], O.direct_self(s["pts"], s["tau"], rows=rows)) < 2e-6
    s["plan"].build_expansions(s["tau"])
    _, Pd, _ = s["plan"].download_expansions()
    _, Pm, _ = plan.download_expansions()
    assert mrd(Pm.ravel(), Pd.ravel()) < 1e-5


def test_expansions_vs_oracle(setup):
    """Expansions against the oracle summing in the device's point order.  The
    equivalent densities (multipoles, locals) come out of pseudo-inverses that keep
    modes down to 50 eps of the largest singular value, so two GEMM libraries leave
    them apart by that rounding amplified in the near-null modes (1e-9 at p = 8);
    what they represent is pinned: the multipoles through their potentials on the
    upward check surface (K_up M), the locals through theirs on the downward check
    surface (K_down L), the partial locals (check data) directly.  The reference's own
    comparison, at p = 6, is test/test_device_passes.jl:91-109 (and 6-31)."""
    from ffmm import operators as fo
    s = setup
    plan = s["plan"]
    ops = s["cache"].host_operators()
    ref = O.FMM(ops, s["pts"], 30, BOX[0], BOX[1], perm=plan.download_tree().perm)
    u_ref = ref.matvec(s["tau"])
    assert mrd(plan.matvec(s["tau"]).cpu().numpy(), u_ref) < 1e-12
    M, Pl, Lc = plan.download_expansions()
    K_up = fo.laplace3d_matrix(ops.small, ops.large)        # multipole -> upward check potentials
    K_dn = fo.laplace3d_matrix(ops.large, ops.small)        # local density -> downward check potentials
    assert mrd((K_up @ M).ravel(), (K_up @ ref.multipoles).ravel()) < 1e-11
    assert mrd(Pl.ravel(), ref.plocals.ravel()) < 1e-11
    leaves = np.concatenate(ref.leaf_nodes)
    assert mrd((K_dn @ Lc).ravel(), (K_dn @ ref.locals[:, leaves]).ravel()) < 1e-11
    nonleaf = np.ones(ref.tree.nnodes, bool)
    nonleaf[leaves] = False
    assert np.all(ref.locals[:, nonleaf] == 0)


@pytest.mark.parametrize("out", ["potential", "gradient", "both"])
def test_targets_vs_oracle_and_direct(setup, out):
    s = setup
    ffmm = s["ffmm"]
    plan, ref = s["plan"], s["ref"]
    targets = uniform(np.random.default_rng(3), 3000)
    plan.build_expansions(s["tau"])
    output = dict(potential=ffmm.Potential, gradient=ffmm.Gradient, both=ffmm.PotentialGradient)[out]
    v = plan.evaluate(targets, output).cpu().numpy()
    tau_sorted = s["tau"][ref.tree.perm]
    ref.matvec(s["tau"])
    if out in ("potential", "both"):
        vr = ref.evaluate(ref.Lc, tau_sorted, targets, grad=False)
        vp = v if out == "potential" else v[:, 0]
        assert mrd(vp, vr) < 1e-11
        assert mrd(vp, O.direct(s["pts"], s["tau"], targets)) < 1e-6
    if out in ("gradient", "both"):
        gr = ref.evaluate(ref.Lc, tau_sorted, targets, grad=True)
        vg = v if out == "gradient" else v[:, 1:]
        assert mrd(vg, gr) < 1e-11
        assert mrd(vg, O.direct(s["pts"], s["tau"], targets, grad=True)) < 1e-5


def test_mixed_evaluation_vs_direct_and_f64(setup):
    """The mixed form's FP32 stages: S2M (source offsets from the node centre formed in
    FP64, then rounded), M2M (sgemm), and the evaluation (near-field pair differences in
    FP64, then FP32 with v_rsq_f32 and a Kahan-compensated potential sum; L2T in
    leaf-relative FP32 coordinates).  Potential and gradient at the sources and at
    uniform targets against FP64 direct sums, and against the FP64 plan."""
    s = setup
    ffmm = s["ffmm"]
    print("[fmm] mixed evaluation", flush=True)
    plan = ffmm.DevicePlan(s["m"], s["pts"], ppl=30, bbox=BOX, cache=s["cache"], m2l="spectral", precision="mixed")
    targets = uniform(np.random.default_rng(5), 3000)
    plan.build_expansions(s["tau"])
    v = plan.evaluate(targets, ffmm.PotentialGradient).cpu().numpy()
    assert mrd(v[:, 0], O.direct(s["pts"], s["tau"], targets)) < 2e-6
    assert mrd(v[:, 1:], O.direct(s["pts"], s["tau"], targets, grad=True)) < 2e-5
    s["plan"].build_expansions(s["tau"])
    v64 = s["plan"].evaluate(targets, ffmm.PotentialGradient).cpu().numpy()
    assert mrd(v[:, 0], v64[:, 0]) < 2e-6 and mrd(v[:, 1:], v64[:, 1:]) < 1e-5
    u = plan.matvec(s["tau"], ffmm.PotentialGradient).cpu().numpy()
    rows = s["rng"].choice(len(u), 400, replace=False)
    assert mrd(u[rows, 0], O.direct_self(s["pts"], s["tau"], rows=rows)) < 2e-6
    assert mrd(u[rows, 1:], O.direct_self(s["pts"], s["tau"], grad=True, rows=rows)) < 2e-5
    assert np.all(np.isfinite(u))


def test_gradient_at_sources(setup):
    s = setup
    ffmm = s["ffmm"]
    g = s["plan"].matvec(s["tau"], ffmm.Gradient).cpu().numpy()
    gr = s["ref"].evaluate_at_sources(s["ref"].Lc, s["tau"][s["ref"].tree.perm], grad=True)
    assert mrd(g, gr) < 1e-11


def test_determinism_sorted_inputs_linearity_nan_recovery(setup):
    s = setup
    plan, tau = s["plan"], s["tau"]
    u0 = plan.matvec(tau).cpu().numpy()
    assert np.array_equal(plan.matvec(tau).cpu().numpy(), u0)              # bitwise repeatable
    perm = plan.download_tree().perm
    assert np.array_equal(plan.matvec(tau[perm], sorted=True).cpu().numpy(), u0)
    t2 = np.random.default_rng(9).standard_normal(len(tau))
    u2 = plan.matvec(t2).cpu().numpy()
    u12 = plan.matvec(tau + 2 * t2).cpu().numpy()
    assert mrd(u12, u0 + 2 * u2) < 1e-12
    bad = tau.copy()
    bad[5] = np.nan
    assert np.any(np.isnan(plan.matvec(bad).cpu().numpy()))
    assert np.array_equal(plan.matvec(tau).cpu().numpy(), u0)              # clean again, bitwise


def test_host_api_mirror(setup):
    """FMM / build_expansions / evaluate_at_sources / evaluate, as the reference's
    host driver is called (src/fmm.jl:31-49, 300-314; src/evaluate.jl:98-132)."""
    s = setup
    ffmm = s["ffmm"]
    f = ffmm.FMM(s["m"], s["pts"], ppl=30, bbox=BOX, cache=s["cache"])
    st = ffmm.build_expansions(f, s["pts"], s["tau"])
    u = ffmm.evaluate_at_sources(f, st)
    assert mrd(u, s["u_ref"]) < 1e-11
    targets = uniform(np.random.default_rng(4), 500)
    v = ffmm.evaluate(f, st, targets)
    assert mrd(v, O.direct(s["pts"], s["tau"], targets)) < 1e-6
    with pytest.raises(ffmm.DimensionMismatch):
        ffmm.build_expansions(f, s["pts"], s["tau"][:-1])
    with pytest.raises(ffmm.ArgumentError, match="outside"):
        ffmm.evaluate(f, st, targets + 2.0)


def test_device_direct_sums_against_oracle():
    """ffmm_direct (direct! / direct_self!, src/kernels.jl:442-487) against the oracle's
    FP64 sums: potential, gradient and both, with and without the self term."""
    import ffmm
    rng = np.random.default_rng(21)
    s, t = rng.random((20000, 3)), rng.random((700, 3))
    tau = rng.standard_normal(20000)
    K = ffmm.Laplace3D()
    assert mrd(ffmm.direct(K, s, tau, t), O.direct(s, tau, t)) < 1e-13
    g = ffmm.direct(K, s, tau, t, ffmm.Gradient)
    assert mrd(g, O.direct(s, tau, t, grad=True)) < 1e-13
    b = ffmm.direct(K, s, tau, t, ffmm.PotentialGradient)
    assert np.allclose(b[:, 0], ffmm.direct(K, s, tau, t), rtol=1e-14, atol=0) and np.allclose(b[:, 1:], g, rtol=1e-14)
    rows = np.arange(0, 20000, 37)
    assert mrd(ffmm.direct_self(K, s, tau, rows=rows), O.direct_self(s, tau, rows=rows)) < 1e-13


def test_analytic_charged_sphere():
    """Unit surface charge on a sphere of radius a: a^2/r outside, a inside
    (test/test_3d.jl:240-251)."""
    import ffmm
    c, a = np.array([0.5, 0.5, 0.5]), 0.3
    nth, nph = 40, 80
    x, wg = np.polynomial.legendre.leggauss(nth)
    pts, wts = [], []
    for i in range(nth):
        st = np.sqrt(1 - x[i] ** 2)
        for j in range(nph):
            ph = 2 * np.pi * j / nph
            nrm = np.array([st * np.cos(ph), st * np.sin(ph), x[i]])
            pts.append(c + a * nrm)
            wts.append(a * a * wg[i] * 2 * np.pi / nph)
    pts, wts = np.array(pts), np.array(wts)
    assert abs(wts.sum() - 4 * np.pi * a * a) < 1e-12
    rng = np.random.default_rng(11)
    cand = rng.random((3000, 3))
    ext = cand[np.linalg.norm(cand - c, axis=1) > 0.45][:300]
    cand2 = c + 0.4 * (rng.random((3000, 3)) - 0.5)
    inn = cand2[np.linalg.norm(cand2 - c, axis=1) < 0.2][:300]
    f = ffmm.FMM(ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8), pts, ppl=40, bbox=BOX)
    assert f.nlevels >= 4
    st = ffmm.build_expansions(f, pts, wts)
    ue = ffmm.evaluate(f, st, ext)
    ui = ffmm.evaluate(f, st, inn)
    assert mrd(ue, a * a / np.linalg.norm(ext - c, axis=1)) < 1e-6
    assert mrd(ui, np.full(len(inn), a)) < 1e-6
