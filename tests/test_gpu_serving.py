"""The serving loop on the GPU: a mixed stream of solves of uneven sizes and methods
(cube Laplace mixed/spectral and FP64/dense, Yukawa and Stokes harmonics), some on a
repeated geometry with new densities, submitted at once to a 4-slot server.  Every
result equals the same solve run alone through a DevicePlan; the repeated geometries
reuse their resident plans; a tight memory budget still completes every request."""
import numpy as np
import pytest

from pointsets import mixed, uniform

pytestmark = pytest.mark.gpu

BOX = (np.zeros(3), 1.0)


def solo(method, cache, pts, dens, tgt, ppl, output, precision, m2l):
    import ffmm
    plan = ffmm.DevicePlan(method, pts, ppl=ppl, bbox=BOX, cache=cache, precision=precision, m2l=m2l)
    u = plan.matvec(dens, output).cpu().numpy()
    v = None if tgt is None else plan.evaluate(tgt, output).cpu().numpy()
    return u, v


def stream():
    import ffmm
    rng = np.random.default_rng(21)
    cube = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    yuk = ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=6)
    sto = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=4)
    caches = {id(m): ffmm.OperatorCache(m) for m in (cube, yuk, sto)}
    reqs = []
    g_pts = mixed(rng, 30000)
    for k, n in enumerate([3000, 40000, 12000, 70000, 6000]):
        pts = mixed(rng, n) if k % 2 else uniform(rng, n)
        reqs.append(dict(method=cube, pts=pts, dens=rng.standard_normal(n), tgt=uniform(rng, 500), ppl=40,
                         output=ffmm.PotentialGradient, precision="mixed" if k % 2 else "f64",
                         m2l="spectral" if k % 2 else "dense", geometry=-1))
    for _ in range(3):      # one geometry, three right-hand sides
        reqs.append(dict(method=cube, pts=g_pts, dens=rng.standard_normal(len(g_pts)), tgt=None, ppl=40,
                         output=ffmm.Potential, precision="mixed", m2l="spectral", geometry=7))
    for n in (4000, 15000):
        reqs.append(dict(method=yuk, pts=mixed(rng, n), dens=rng.standard_normal(n), tgt=uniform(rng, 200), ppl=40,
                         output=ffmm.PotentialGradient, precision="f64", m2l="dense", geometry=-1))
    for n in (5000,):
        d = rng.standard_normal((n, 9))
        d[:, 6:] /= np.linalg.norm(d[:, 6:], axis=1)[:, None]
        reqs.append(dict(method=sto, pts=mixed(rng, n), dens=d, tgt=None, ppl=40, output=ffmm.Velocity,
                         precision="f64", m2l="dense", geometry=-1))
    return reqs, caches


def run(reqs, caches, nslots, budget):
    from ffmm.serving import Server
    with Server(nslots=nslots, hbm_budget=budget) as srv:
        tickets = [srv.submit(caches[id(r["method"])], r["pts"], r["dens"], r["tgt"], geometry=r["geometry"],
                              ppl=r["ppl"], bbox=BOX, output=r["output"], precision=r["precision"], m2l=r["m2l"])
                   for r in reqs]
        outs = [srv.wait(t) for t in tickets]
        return outs, srv.stats()


def test_mixed_stream_matches_solo_solves():
    reqs, caches = stream()
    print(f"[serving] {len(reqs)} requests, 4 slots", flush=True)
    outs, st = run(reqs, caches, 4, 32 << 30)
    assert st["completed"] == len(reqs)
    assert st["plans_reused"] >= 2 and st["plans_built"] == len(reqs) - st["plans_reused"]
    for r, (u, v, lat) in zip(reqs, outs):
        us, vs = solo(r["method"], caches[id(r["method"])], r["pts"], r["dens"], r["tgt"], r["ppl"], r["output"],
                      r["precision"], r["m2l"])
        assert lat > 0
        scale = np.max(np.abs(us))
        assert np.max(np.abs(u - us)) <= 1e-12 * scale, type(r["method"]).__name__
        if r["tgt"] is not None:
            assert np.max(np.abs(v - vs)) <= 1e-12 * np.max(np.abs(vs))


def test_tight_budget_serialises_and_completes():
    reqs, caches = stream()
    reqs = reqs[:6]
    print(f"[serving] {len(reqs)} requests, 1 MiB budget", flush=True)
    outs, st = run(reqs, caches, 3, 1 << 20)
    assert st["completed"] == len(reqs)
    for r, (u, v, lat) in zip(reqs, outs):
        us, _ = solo(r["method"], caches[id(r["method"])], r["pts"], r["dens"], r["tgt"], r["ppl"], r["output"],
                     r["precision"], r["m2l"])
        assert np.max(np.abs(u - us)) <= 1e-12 * np.max(np.abs(us))


def test_request_errors_are_reported_per_request():
    import ffmm
    from ffmm.serving import Server
    rng = np.random.default_rng(22)
    sto = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=4)
    cache = ffmm.OperatorCache(sto)
    pts = uniform(rng, 2000)
    with Server(nslots=2, hbm_budget=4 << 30) as srv:
        bad = srv.submit(cache, pts, rng.standard_normal(2000), bbox=BOX, output=ffmm.Velocity)   # charges on Stokes
        good = srv.submit(cache, pts, rng.standard_normal((2000, 3)), bbox=BOX, output=ffmm.Velocity)
        with pytest.raises(ffmm.ArgumentError, match="stokeslets"):
            srv.wait(bad)
        u, _, _ = srv.wait(good)
        assert u.shape == (2000, 3) and np.all(np.isfinite(u))


def test_served_stokes_pressure_matches_plan():
    """A served Stokes request for the pressure (one value per point) against the same
    solve through DevicePlan; a pressure request on a Laplace cache is refused."""
    import ffmm
    from ffmm.serving import Server
    rng = np.random.default_rng(23)
    sto = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=6)
    cache = ffmm.OperatorCache(sto)
    pts = uniform(rng, 3000)
    dens = rng.standard_normal((3000, 9))
    dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
    tgt = uniform(rng, 200)
    lap = ffmm.SphericalHarmonics(ffmm.Laplace3D(), p=4)
    lcache = ffmm.OperatorCache(lap)
    with Server(nslots=2, hbm_budget=4 << 30) as srv:
        t = srv.submit(cache, pts, dens, tgt, bbox=BOX, output=ffmm.Pressure)
        u, v, _ = srv.wait(t)
        with pytest.raises(ffmm.ArgumentError, match="Stokes"):
            srv.submit(lcache, pts, dens[:, 0].copy(), bbox=BOX, output=ffmm.Pressure)
    plan = ffmm.DevicePlan(sto, pts, ppl=64, bbox=BOX, cache=cache)
    pu = plan.matvec(dens, ffmm.Pressure).cpu().numpy()
    pv = plan.evaluate(tgt, ffmm.Pressure).cpu().numpy()
    assert u.shape == (3000,) and v.shape == (200,)
    assert np.allclose(u, pu, rtol=1e-12, atol=1e-12 * np.max(np.abs(pu)))
    assert np.allclose(v, pv, rtol=1e-12, atol=1e-12 * np.max(np.abs(pv)))
