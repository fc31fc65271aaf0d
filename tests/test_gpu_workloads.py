"""The benchmark workloads on the GPU, each at a size the oracle covers in seconds,
with the generators of the full-size runs (ffmm.workloads):

* Gaussian mixture, plain FP64 (the 1M-source configuration): dense and spectral M2L
  against the oracle (1e-11) and the FP64 direct sum (rel L2 <= 1e-6);
* Gaussian mixture + uniform targets, mixed precision with the spectral M2L (the
  4M-source configuration): potential at the sources, potential and gradient at the
  targets, against the oracle and, on 10,000 sampled points, the direct sum (the 1e-6
  gate); at 400,000 sources, where both Hadamard kernels of the spectral M2L run; at
  3,000,000 sources against the plain FP64 plan, where the mixed M2M runs in column
  chunks;
* StokesHarmonics p = 10 on ellipsoid surfaces, stokeslets + stresslets with the
  normals (the 2M-point configuration): the mixed and FP64 forms against the harmonic
  oracle and the FP64 Stokes direct sum;
* the served stream (70% cube p = 8 mixed spectral / 30% Yukawa3D(5) harmonics p = 10,
  repeated geometries): every served result equal to its solo solve and within its
  class's gate of the oracle or the direct sum.

The reference's checks of the same kind: device vs CPU driver 1e-11
(test/test_device_passes.jl:6-31), vs direct 1e-6 at p = 8 (test/test_3d.jl:155-163)."""
import numpy as np
import pytest

import harmonic as HO
import oracle as O

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def mrd(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim == 2:
        return float(np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1)))
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def mixture():
    """60,000 Gaussian-mixture sources (seed 0, the full-size runs' generator), N(0,1)
    charges, 15,000 uniform targets; the oracle's FP64 matvec on the same tree."""
    import ffmm
    from ffmm import workloads as W
    n, nt = 60_000, 15_000
    pts = W.gaussian_mixture(n, seed=0)
    tau = W.charges(n, seed=1)
    tgt = W.uniform(nt, seed=2)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    plan = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=cache)
    tree = plan.download_tree()
    ref = O.FMM(cache.host_operators(), pts, 64, *BOX, perm=tree.perm)
    print(f"[workloads] mixture n={n} levels={plan.nlevels} nodes={plan.nnodes}", flush=True)
    u_ref, v_ref = ref.matvec(tau, grad=False, targets=tgt)
    g_ref = ref.evaluate(ref.Lc, np.ascontiguousarray(tau[tree.perm]), tgt, grad=True)
    rng = np.random.default_rng(7)
    rows = np.sort(rng.choice(n, 5000, replace=False))
    trows = np.sort(rng.choice(nt, 5000, replace=False))
    ud = O.direct_self(pts, tau, rows=rows)
    vd = O.direct(pts, tau, tgt[trows])
    gd = O.direct(pts, tau, tgt[trows], grad=True)
    return dict(ffmm=ffmm, m=m, cache=cache, pts=pts, tau=tau, tgt=tgt, plan=plan, u_ref=u_ref, v_ref=v_ref,
                g_ref=g_ref, rows=rows, trows=trows, ud=ud, vd=vd, gd=gd)


@pytest.mark.parametrize("m2l", ["dense", "spectral"])
def test_mixture_plain_f64(mixture, m2l):
    """The plain FP64 form on the Gaussian mixture: an adaptive tree of depth >= 8, the
    oracle to 1e-11, the direct sum to rel L2 1e-6."""
    s = mixture
    ffmm = s["ffmm"]
    print(f"[workloads] mixture f64 {m2l}", flush=True)
    plan = s["plan"] if m2l == "dense" else ffmm.DevicePlan(s["m"], s["pts"], ppl=64, bbox=BOX, cache=s["cache"],
                                                            m2l="spectral")
    assert plan.nlevels >= 9                       # depth >= 8 below the root
    u = plan.matvec(s["tau"]).cpu().numpy()
    assert mrd(u, s["u_ref"]) < 1e-11
    assert rel(u[s["rows"]], s["ud"]) < 1e-6


def test_mixture_mixed_spectral_targets(mixture):
    """The mixed form (FP32 spectral M2L on MFMA, FP32 near field and L2T with
    compensated sums, FP64 pseudo-inverses): potential at the sources, potential and
    gradient at the targets, each within the 1e-6 rel L2 gate of the FP64 direct sum on
    10,000 sampled points (5,000 sources, 5,000 targets) and within 2e-6 (max-norm) of
    the oracle's FP64 FMM."""
    s = mixture
    ffmm = s["ffmm"]
    print("[workloads] mixture mixed spectral + targets", flush=True)
    plan = ffmm.DevicePlan(s["m"], s["pts"], ppl=64, bbox=BOX, cache=s["cache"], precision="mixed", m2l="spectral")
    plan.build_expansions(s["tau"])
    u = plan.evaluate_at_sources(ffmm.Potential).cpu().numpy()
    v = plan.evaluate(s["tgt"], ffmm.PotentialGradient).cpu().numpy()
    assert mrd(u, s["u_ref"]) < 2e-6
    assert mrd(v[:, 0], s["v_ref"]) < 2e-6
    assert mrd(v[:, 1:], s["g_ref"]) < 2e-5
    e = (rel(u[s["rows"]], s["ud"]), rel(v[s["trows"], 0], s["vd"]), rel(v[s["trows"], 1:], s["gd"]))
    print(f"[workloads] mixed rel L2 vs direct: sources {e[0]:.2e}, targets {e[1]:.2e}, gradient {e[2]:.2e}",
          flush=True)
    assert max(e) <= 1e-6


def test_mixture_400k_mixed_spectral_both_hadamard_kernels():
    """400,000 Gaussian-mixture sources, mixed precision, spectral M2L. At this size a few
    tiles of 16 sibling groups name more than 200 distinct colleague boxes, so the
    Hadamard products run in both kernels (the one staging each tile's boxes once, and
    the per-group one for those tiles); the potential at the sources within 2e-6
    (max-norm) of the oracle's FP64 FMM on the same tree, and within the 1e-6 rel L2
    gate of the FP64 direct sum on 3,000 sampled sources."""
    import ffmm
    from ffmm import workloads as W
    n = 400_000
    pts = W.gaussian_mixture(n, seed=0)
    tau = W.charges(n, seed=1)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[workloads] mixture n={n}, mixed spectral, both Hadamard kernels", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=cache, precision="mixed", m2l="spectral")
    staged, per_group = plan.spectral_tiles()
    print(f"[workloads] tiles staged {staged.tolist()}, per group {per_group.tolist()}", flush=True)
    assert staged.sum() > 0 and per_group.sum() > 0
    plan.build_expansions(tau)
    u = plan.evaluate_at_sources(ffmm.Potential).cpu().numpy()
    tree = plan.download_tree()
    ref = O.FMM(cache.host_operators(), pts, 64, *BOX, perm=tree.perm)
    u_ref = ref.matvec(tau, m2l="spectral")
    assert mrd(u, u_ref) < 2e-6
    rows = np.sort(np.random.default_rng(11).choice(n, 3000, replace=False))
    assert rel(u[rows], O.direct_self(pts, tau, rows=rows)) <= 1e-6


def test_mixture_3m_mixed_m2m_in_column_chunks():
    """3,000,000 Gaussian-mixture sources: some levels have more than 10,240 sibling
    groups, so the mixed form's FP32 M2M runs in several column chunks there.  The
    potential at every source agrees with the plain FP64 plan on the same points (which
    makes one M2M call per level) to the mixed form's 1e-6 rel L2."""
    import ffmm
    from ffmm import workloads as W
    n = 3_000_000
    pts = W.gaussian_mixture(n, seed=5)
    tau = W.charges(n, seed=6)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    print(f"[workloads] mixture n={n}, mixed vs FP64, chunked M2M", flush=True)
    mixed = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=cache, precision="mixed", m2l="spectral")
    groups = np.asarray(mixed.level_count) // 8
    print(f"[workloads] sibling groups per level {groups.tolist()}", flush=True)
    assert groups[3:].max() > 10240
    mixed.build_expansions(tau)
    u = mixed.evaluate_at_sources(ffmm.Potential).cpu().numpy()
    del mixed
    plain = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=cache, precision="f64", m2l="spectral")
    plain.build_expansions(tau)
    u64 = plain.evaluate_at_sources(ffmm.Potential).cpu().numpy()
    err = rel(u, u64)
    print(f"[workloads] mixed vs FP64 rel L2 {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("precision,tol_oracle", [("mixed", 1e-6), ("f64", 1e-11)])
def test_stokes_ellipsoids(precision, tol_oracle):
    """StokesHarmonics p = 10 on 64 ellipsoids (20,000 points), stokeslets plus
    stresslets with the surface normals, velocity at the sources: the harmonic oracle
    (FP64, same tree and operators) and the FP64 Stokes direct sum (the truncation error
    of p = 10 on surfaces, a few 1e-4)."""
    import ffmm
    from ffmm import workloads as W
    n = 20_000
    pts, nrm = W.ellipsoids(n, seed=4)
    dens = W.stokes_densities(nrm, seed=5)
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=10)
    cache = ffmm.OperatorCache(m)
    print(f"[workloads] stokes ellipsoids n={n} {precision}", flush=True)
    plan = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=cache, precision=precision)
    u = plan.matvec(dens, ffmm.Velocity).cpu().numpy()
    W0 = cache.W0
    f = HO.HarmonicFMM("stokes", lambda l, w: (cache.harmonic_operators(w), W0 / w), pts, 64, *BOX,
                       perm=plan.download_tree().perm)
    uo = f.matvec(dens)
    assert rel(u, uo) < tol_oracle
    rows = np.sort(np.random.default_rng(9).choice(n, 2000, replace=False))
    ud = HO.direct("stokes", pts, dens, pts[rows], skip=rows)
    if precision == "f64":
        # the device FP64 direct sums (velocity and pressure) against the oracle's
        assert rel(ffmm.direct(ffmm.Stokeslet3D(), pts, dens, pts[rows], skip=rows), ud) < 1e-12
        pd = HO.direct("stokes", pts, dens, pts[rows], 3, skip=rows)
        assert rel(ffmm.direct(ffmm.Stokeslet3D(), pts, dens, pts[rows], ffmm.Pressure, skip=rows), pd) < 1e-12
        pr = plan.evaluate_at_sources(ffmm.Pressure).cpu().numpy()
        assert rel(pr, f.matvec(dens, output=3)) < 1e-11
        assert rel(pr[rows], pd) < 1e-3
    err = rel(u[rows], ud)
    print(f"[workloads] stokes {precision} rel L2 vs direct {err:.2e}", flush=True)
    assert err < 1e-3


def test_served_stream_matches_solo_and_references():
    """The served stream (ffmm.workloads.served_stream) at small sizes: 14 requests of
    10,000-40,000 points, 70% Laplace cube p = 8 (mixed, spectral M2L) / 30% Yukawa3D(5)
    harmonics p = 10 (mixed), 40% repeated geometries, all submitted at once to a 4-slot
    server.  Each served result equals its solo DevicePlan solve to 1e-12; each is within
    2e-6 of the oracle's FP64 FMM (cube) or 1e-6 of the harmonic oracle (Yukawa); each
    meets its class's gate against the FP64 direct sum at 300 sampled sources (1e-6 for
    the cube, 1e-5 for Yukawa p = 10, whose truncation error at k = 5 is a few 1e-6 --
    the reference's end-to-end Yukawa tolerance, test/test_harmonics_yukawa.jl:107-109 --
    and 1e-3 for the coarse class)."""
    import ffmm
    from ffmm import workloads as W
    from ffmm.serving import Server
    reqs = W.served_stream(14, nmin=10_000, nmax=40_000, seed=11)
    cube = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    yuk = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=10)
    caches = {True: ffmm.OperatorCache(cube), False: ffmm.OperatorCache(yuk)}
    form = {True: ("mixed", "spectral"), False: ("mixed", "dense")}
    data = [(W.request_points(r), W.request_charges(r)) for r in reqs]
    assert len({r["gid"] for r in reqs}) < len(reqs)            # some geometries repeat
    print(f"[workloads] served stream: {len(reqs)} requests", flush=True)
    with Server(nslots=4, hbm_budget=16 << 30) as srv:
        tickets = [srv.submit(caches[r["lap"]], pts, tau, geometry=r["gid"], ppl=64, bbox=BOX,
                              precision=form[r["lap"]][0], m2l=form[r["lap"]][1])
                   for r, (pts, tau) in zip(reqs, data)]
        outs = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["plans_reused"] >= 1 and st["plans_built"] + st["plans_reused"] == len(reqs)
    for k, (r, (pts, tau), u) in enumerate(zip(reqs, data, outs)):
        m = cube if r["lap"] else yuk
        plan = ffmm.DevicePlan(m, pts, ppl=64, bbox=BOX, cache=caches[r["lap"]], precision=form[r["lap"]][0],
                               m2l=form[r["lap"]][1])
        us = plan.matvec(tau).cpu().numpy()
        assert np.max(np.abs(u - us)) <= 1e-12 * np.max(np.abs(us)), k
        perm = plan.download_tree().perm
        rows = np.sort(np.random.default_rng(k).choice(len(pts), 300, replace=False))
        if r["lap"]:
            ref = O.FMM(caches[True].host_operators(), pts, 64, *BOX, perm=perm).matvec(tau)
            assert mrd(u, ref) < 2e-6, k
            ud = O.direct_self(pts, tau, rows=rows)
            gate = 1e-6 if r["fine"] else 1e-3
        else:
            c = caches[False]
            f = HO.HarmonicFMM("yukawa", lambda l, w: (c.harmonic_operators(w), 1.0), pts, 64, *BOX, kappa=5.0,
                               perm=perm)
            assert rel(u, f.matvec(tau)) < 1e-6, k
            ud = HO.direct("yukawa", pts, tau, pts[rows], kappa=5.0, skip=rows)
            gate = 1e-5 if r["fine"] else 1e-3
        err = rel(u[rows], ud)
        print(f"[workloads] request {k}: {'cube' if r['lap'] else 'yukawa'} n={len(pts)} rel L2 {err:.2e}",
              flush=True)
        assert err <= gate, k
