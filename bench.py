#!/usr/bin/env python
"""Laplace3D cube kernel-independent FMM (p = 8) on one MI355X.

Workload (the configuration the headline metric is quoted on): 4,000,000 sources
from a mixture of 32 isotropic Gaussians (centres uniform in [0.1, 0.9]^3, sigma
log-uniform in [0.005, 0.05], clipped to the unit cube), charges N(0, 1), ppl = 64,
box ((0,0,0), 1); 1,000,000 uniform targets.  One step is a whole solve from
device-resident points: the device tree and its lists (the plan), the upward and
downward passes, the potential at every source, and the potential and gradient at
every target.  value = sources / second of one step.

After the timed region: the relative L2 error against the FP64 direct sum at 10,000
sampled sources (potential) and 10,000 sampled targets (potential and gradient), the roofline of the dominant kernel
from its live HIP-event duration, and the CPU baseline: the oracle (the restated
reference CPU algorithm) timed on the host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "flexiblefmm.jl_amd"))

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 78.6      # MI355X FP32 vector without packed math (packed math is not used)
PEAK_FP32_MFMA_TFLOPS = 157.3   # MI355X dense FP32 MFMA
PEAK_HBM_GBS = 8000.0
FLOPS_POT = 11               # flop conventions of the workload: Laplace potential
FLOPS_POTGRAD = 20           # potential + gradient


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def interaction_counts(tree, nleaf_targets=None):
    """P2P source counts per real leaf: sum over its colleagues of their point counts."""
    n = tree.phi - tree.plo + 1
    col = tree.colleagues
    cnt = np.where(col >= 0, n[np.maximum(col, 0)], 0).sum(axis=1)
    leaves = np.nonzero(tree.leaf)[0]
    return leaves, cnt[leaves], n[leaves]


def gemm_shapes(plan, ops, mixed: bool):
    """The GEMMs of one step of the cube method, in launch order: (name, m, n, k, type,
    flop, library: run by the BLAS library).  Per step: UC2E as two FP64 factors over the childless sourced nodes of
    levels 2 and below, M2M per level (FP32 in the mixed form, in calls of at most 10,240
    columns) from level L-2 up to 2, L2L
    per level (FP64) from level 3 down, DC2E as two FP64 factors over the leaves."""
    Ns, Nl, L = ops.nsmall, ops.nlarge, plan.nlevels
    nq_up, nq_dn = int(plan.level_s2m[2:].sum()), int(plan.level_leaves[2:].sum())
    ru, rd = ops.P_up.rank, ops.P_down.rank
    # ranks 241..304: UC2E's first factor runs inside k_s2m_uc2e (S2M fused with it), not as
    # a library GEMM; with 240 < Ns <= 304 (r_dn likewise), UC2E's second factor and DC2E's
    # first run as k_gemm_cols
    fused = 241 <= ru <= 304
    cols_up, cols_dn = 240 < Ns <= 304, 240 < rd <= 304
    out = [("UC2E Uh" + (" (in k_s2m_uc2e)" if fused else ""), ru, nq_up, Nl, "f64"),
           ("UC2E V" + (" (in k_gemm_cols)" if cols_up else ""), Ns, nq_up, ru, "f64")]
    for l in range(L - 2, 1, -1):
        ng = int(plan.level_count[l + 1]) // 8
        if mixed:   # the FP32 M2M runs in calls of at most 10,240 columns
            for c0 in range(0, ng, 10240):
                part = f" (columns {c0}-{min(ng, c0 + 10240) - 1})" if ng > 10240 else ""
                out.append((f"M2M level {l}{part}", Ns, min(10240, ng - c0), 8 * Ns, "f32"))
        else:
            out.append((f"M2M level {l}", Ns, ng, 8 * Ns, "f64"))
    for l in range(3, L):
        out.append((f"L2L level {l}", 8 * Ns, int(plan.level_count[l]) // 8, Ns, "f64"))
    out += [("DC2E Uh" + (" (in k_gemm_cols)" if cols_dn else ""), rd, nq_dn, Ns, "f64"), ("DC2E V", Nl, nq_dn, rd, "f64")]
    return [{"name": a, "m": m_, "n": n_, "k": k_, "type": t_, "flop": 2.0 * m_ * n_ * k_, "library": " (in " not in a}
            for a, m_, n_, k_, t_ in out if m_ > 0 and n_ > 0]


def traffic_from_profile(kname: str, args):
    """HBM bytes per launch of the dominant kernel from the committed two-pass
    rocprofv3 --pmc profile (FETCH_SIZE, then WRITE_SIZE, each scaled by the copy
    probe's reading), when the profile is of this kernel at this workload; else None."""
    path = os.path.join(ROOT, "profiles", "hadamard_spectral_4M_mixed_traffic.json")
    if not os.path.exists(path) or args.n != 4_000_000 or args.ppl != 64:
        return None
    prof = json.load(open(path))
    return prof["traffic_bytes_per_launch"] if kname.startswith(prof["kernel"]) else None


PUBLISHED_CPU = {"value": 0.69e6, "unit": "points/s",
                 "case": "Laplace CubeKI p=8, 1e5 uniform points, ppl=200, up 0.025 + down 0.086 + self 0.033 s, "
                         "FP64, 12 threads on an Apple M3 Max", "source": "docs/harmonics.md:79-85"}


def _cpu_solve(O, ops, pts, tau, ppl):
    """One solve of the restated reference CPU FMM (spectral M2L, FP64): tree, expansions
    (upward and downward passes), potential at the sources; the three times and the
    output in the original order."""
    t0 = time.perf_counter()
    f = O.FMM(ops, pts, ppl, np.zeros(3), 1.0)
    t1 = time.perf_counter()
    ts = np.ascontiguousarray(tau[f.tree.perm])
    Mp = f.upward(ts)
    _, Lc = f.downward(Mp, "spectral")
    t2 = time.perf_counter()
    u = f.evaluate_at_sources(Lc, ts)
    t3 = time.perf_counter()
    return (t1 - t0, t2 - t1, t3 - t2), u


def _best_of(O, ops, pts, tau, ppl, runs):
    """One warm-up solve, then the best (by total) of `runs` solves."""
    _cpu_solve(O, ops, pts, tau, ppl)
    best, u = None, None
    for _ in range(runs):
        t, u = _cpu_solve(O, ops, pts, tau, ppl)
        if best is None or sum(t) < sum(best):
            best = t
    return best, u


def cpu_baseline(n_sample: int, ppl: int, runs: int, gpu_solve):
    """The oracle's restatement of the reference CPU FMM (C/OpenMP per-point loops and
    spectral Hadamard, threaded FFTs and BLAS; FP64, spectral M2L as
    src/operators/spectral_m2l.jl:328-451) on the host cores: the best of `runs` warm
    solves on `n_sample` Gaussian-mixture sources of the same generator, split as
    tree / expansions / evaluation (bench/bench.jl:44-56), and on the reference's
    published case (1e5 uniform points, ppl = 200) beside the published rate.  The
    device result on the same sample (the measured form, `gpu_solve`) is compared with
    the restated FMM's output: the error against the reference FMM's own output."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from ffmm import operators, workloads as W
    ops = operators.level_operators(1.0, p=8, with_tables=True)
    pts = W.gaussian_mixture(n_sample, seed=0)
    tau = W.charges(n_sample, seed=1)
    (tt, te, tv), u_cpu = _best_of(O, ops, pts, tau, ppl, runs)
    total = tt + te + tv
    # the reference's published configuration, same restatement
    pu = W.uniform(100_000, seed=3)
    tu = W.charges(100_000, seed=4)
    (ut, ue, uv), _ = _best_of(O, ops, pu, tu, 200, runs)
    u_gpu = gpu_solve(pts, tau)
    rel = float(np.linalg.norm(u_gpu - u_cpu) / np.linalg.norm(u_cpu))
    rows = np.sort(np.random.default_rng(11).choice(n_sample, 2000, replace=False))
    ud = O.direct_self(pts, tau, rows=rows)
    cpu_err = float(np.linalg.norm(u_cpu[rows] - ud) / np.linalg.norm(ud))
    return {"value": n_sample / total, "unit": "points/s", "cores": O.lib().othreads(), "kind": "port",
            "sample": f"{n_sample} Gaussian-mixture sources (the same generator), ppl={ppl}, p=8, FP64 spectral "
                      f"M2L, tree + expansions + potential at the sources, best of {runs} warm runs",
            "split_s": {"tree": tt, "expansions": te, "evaluation": tv, "total": total},
            "rel_l2_vs_direct": cpu_err,
            "published_case": {"points_per_s": 1e5 / (ut + ue + uv), "split_s": {"tree": ut, "expansions": ue,
                                                                                 "evaluation": uv},
                               "published": PUBLISHED_CPU},
            "gpu_rel_l2_vs_restated_cpu_fmm": rel}


def serving_bench(nreq: int, nmin: int, nmax: int, rate: float, nslots: int, plain_every: int,
                  timeline: str | None = None):
    """The served stream (ffmm.workloads.served_stream): Poisson arrivals, point counts
    log-uniform in [nmin, nmax], 70% Laplace3D cube p = 8 (mixed, spectral M2L) / 30%
    Yukawa3D(5) harmonics p = 10 (mixed, dense class M2L), uniform or Gaussian-mixture
    layouts, 40% repeated geometries with new charges, the 1e-6 and 1e-3 classes.

    Two passes through a fresh server: arrivals replayed in real time at ``rate`` (the
    latencies at a rate the server sustains) and all requests at once (the saturated
    throughput).  Against them, the same requests solved one after another with a new
    plan each (no resident plans, no overlap): in the served forms (every request), and
    in the plain FP64 forms with the same M2L algorithms (every ``plain_every``-th
    request)."""
    import threading
    import ffmm
    from ffmm import workloads as W
    from ffmm.serving import Server
    torch = __import__("torch")
    box = (np.zeros(3), 1.0)
    methods = {True: ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8),
               False: ffmm.SphericalHarmonics(ffmm.Yukawa3D(5.0), p=10)}
    caches = {k: ffmm.OperatorCache(m) for k, m in methods.items()}
    form = {True: ("mixed", "spectral"), False: ("mixed", "dense")}
    plain_form = {True: ("f64", "spectral"), False: ("f64", "dense")}
    reqs = W.served_stream(nreq, nmin, nmax, rate, seed=5)
    t = time.perf_counter()
    geo, data = {}, []
    for r in reqs:
        if r["gid"] not in geo:
            geo[r["gid"]] = W.request_points(r)
        data.append((geo[r["gid"]], W.request_charges(r)))
    total = float(sum(r["n"] for r in reqs))
    # the server's budget: 60% of the device memory free when the leg starts; the slots'
    # idle blocks (kept for the next solve) use the rest
    torch.cuda.empty_cache()
    budget = int(0.6 * torch.cuda.mem_get_info()[0])
    log(f"[bench] serving: {nreq} requests, {total:.3e} points, {len(geo)} geometries, {nslots} slots "
        f"(generated in {time.perf_counter() - t:.1f} s)")

    def serve(time_scale: float):
        """The stream through a fresh server, arrivals at the stream's times x time_scale
        (0: all at once) replayed by a submitting thread; per request the latency and
        the submission lag, and the first served result of each (kernel, class)."""
        srv = Server(nslots=nslots, hbm_budget=budget)
        for k, c in caches.items():      # warm the operator caches (host precompute, device copies)
            c.fit_box(np.zeros(3), np.ones(3), box)
            srv._prepare(c, np.zeros(3), np.ones(3), box)
        tickets = [None] * nreq
        t_sub = [0.0] * nreq

        def submitter():
            t0 = time.perf_counter()
            for i, r in enumerate(reqs):
                dt = time_scale * r["t"] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                t_sub[i] = time.perf_counter()
                prec, m2l = form[r["lap"]]
                tickets[i] = srv.submit(caches[r["lap"]], data[i][0], data[i][1], geometry=r["gid"], ppl=64,
                                        bbox=box, precision=prec, m2l=m2l)

        t0 = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        lat, first = [], {}
        for i in range(nreq):
            while tickets[i] is None:
                time.sleep(0.0002)
            u, _, ms = srv.wait(tickets[i])
            lat.append(ms)
            key = (reqs[i]["lap"], reqs[i]["fine"])
            if key not in first:
                first[key] = (i, u)
        wall = time.perf_counter() - t0
        th.join()
        lag = np.array([t_sub[i] - t0 - time_scale * reqs[i]["t"] for i in range(nreq)])
        st = srv.stats()
        srv.close()
        lat = np.array(lat)
        sizes = np.array([r["n"] for r in reqs])
        if timeline:
            with open(f"{timeline}.{'rate' if time_scale else 'burst'}.json", "w") as fh:
                json.dump([{"i": i, "t_submit": t_sub[i] - t0, "latency_ms": float(lat[i]), "n": int(reqs[i]["n"]),
                            "lap": bool(reqs[i]["lap"]), "gid": int(reqs[i]["gid"])} for i in range(nreq)], fh)
        by_size = {}
        for lo_, hi_ in ((0, 1e5), (1e5, 5e5), (5e5, 1e7)):
            sel = (sizes >= lo_) & (sizes < hi_)
            if sel.any():
                by_size[f"{int(lo_)}-{int(hi_)}"] = {"requests": int(sel.sum()),
                                                      "p50": float(np.percentile(lat[sel], 50)),
                                                      "p99": float(np.percentile(lat[sel], 99))}
        return {"wall_s": wall, "served_points_per_s": total / wall, "p50_latency_ms": float(np.percentile(lat, 50)),
                "p99_latency_ms": float(np.percentile(lat, 99)), "max_latency_ms": float(lat.max()),
                "submit_lag_ms": {"mean": 1e3 * float(lag.mean()), "max": 1e3 * float(lag.max())},
                "plans_built": st["plans_built"], "plans_reused": st["plans_reused"],
                "evictions": st["evictions"], "peak_device_bytes": st["peak_bytes"], "budget_bytes": budget,
                "requeued_out_of_memory": st["requeued_out_of_memory"], "allocation_flushes": st["allocation_flushes"],
                "latency_by_size_ms": by_size,
                "worst_plan_over_estimate": st["worst_plan_over_estimate_permille"] / 1000}, first

    sustained, first = serve(1.0)
    log(f"[bench] serving at {rate}/s: " + json.dumps(sustained))
    saturated, _ = serve(0.0)
    log("[bench] serving, all at once: " + json.dumps(saturated))

    def sequential(fm, every):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pts_done = 0
        for i in range(0, nreq, every):
            r = reqs[i]
            prec, m2l = fm[r["lap"]]
            plan = ffmm.DevicePlan(methods[r["lap"]], data[i][0], ppl=64, bbox=box, cache=caches[r["lap"]],
                                   precision=prec, m2l=m2l)
            plan.matvec(data[i][1]).cpu()
            del plan
            pts_done += r["n"]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        return {"wall_s": dt, "points": pts_done, "points_per_s": pts_done / dt,
                "requests": len(range(0, nreq, every))}

    seq = sequential(form, 1)
    log("[bench] sequential, served forms: " + json.dumps(seq))
    seq_plain = sequential(plain_form, plain_every)
    log("[bench] sequential, plain FP64 forms: " + json.dumps(seq_plain))
    # the accuracy of the served results: the first request of each (kernel, class)
    # against the FP64 direct sum at 300 sampled sources (Laplace on the device; Yukawa in numpy)
    acc = {}
    gates = {(True, True): 1e-6, (True, False): 1e-3, (False, True): 1e-5, (False, False): 1e-3}
    for key, (i, u) in sorted(first.items()):
        pts, tau = data[i]
        rows = np.random.default_rng(5).choice(len(pts), min(300, len(pts)), replace=False)
        if key[0]:
            ref = ffmm.direct(ffmm.Laplace3D(), pts, tau, pts[rows], skip=rows)
        else:
            ref = np.empty(len(rows))
            for a, j in enumerate(rows):
                d = np.linalg.norm(pts - pts[j], axis=1)
                d[j] = np.inf
                ref[a] = np.sum(tau * np.exp(-5.0 * d) / (4 * np.pi * d))
        err = float(np.linalg.norm(u[rows] - ref) / np.linalg.norm(ref))
        name = ("laplace_cube_p8" if key[0] else "yukawa_harmonic_p10") + ("_1e-6" if key[1] else "_1e-3")
        acc[name] = {"rel_l2": err, "gate": gates[key], "ok": err <= gates[key], "points": len(pts),
                     "form": "/".join(form[key[0]])}
    return {"requests": nreq, "points": total, "geometries": len(geo), "slots": nslots, "n_range": [nmin, nmax],
            "arrival_rate_per_s": rate, "arrival_span_s": reqs[-1]["t"], "sustained": sustained,
            "saturated": saturated, "p50_latency_ms": sustained["p50_latency_ms"],
            "p99_latency_ms": sustained["p99_latency_ms"], "served_points_per_s": saturated["served_points_per_s"],
            "sequential_same_forms": seq, "sequential_plain_forms": seq_plain,
            "speedup_vs_sequential_same_forms": saturated["served_points_per_s"] / seq["points_per_s"],
            "speedup_vs_sequential_plain": saturated["served_points_per_s"] / seq_plain["points_per_s"],
            "accuracy": acc,
            "note": "served forms: cube p=8 mixed/spectral, Yukawa p=10 mixed/dense for both classes (the orders "
                    "are fixed by the workload); plain forms: FP64 with the same M2L algorithms; sequential legs "
                    "build a new plan per request; the plain leg runs every %d-th request" % plain_every}


def stokes_bench(n: int, ppl: int, nerr: int):
    """The Stokes surface workload: n points on 64 ellipsoids, stokeslets + stresslets
    with the surface normals, StokesHarmonics p=10, velocity at the sources; the mixed
    form (FP32 class M2L) and the FP64 form, one warm-up and one timed solve each; the
    error against the device FP64 direct sum at sampled sources."""
    import ffmm
    from ffmm import workloads as W
    torch = __import__("torch")
    pts, nrm = W.ellipsoids(n, seed=4)
    dens = W.stokes_densities(nrm, seed=5)
    m = ffmm.StokesHarmonics(ffmm.Stokeslet3D(), p=10)
    cache = ffmm.OperatorCache(m)
    box = (np.zeros(3), 1.0)
    dev = "cuda:0"
    xs_d = torch.from_numpy(pts).to(dev)
    d_d = torch.from_numpy(dens).to(dev)
    u_d = torch.empty((n, 3), dtype=torch.float64, device=dev)
    res = {"workload": f"StokesHarmonics p=10, {n} points on 64 ellipsoids, stokeslets + stresslets, "
                       f"velocity at the sources, ppl={ppl}"}
    rows = np.sort(np.random.default_rng(9).choice(n, nerr, replace=False))
    ud = None
    for prec in ("mixed", "f64"):
        def solve():
            plan = ffmm.DevicePlan(m, xs_d, ppl=ppl, bbox=box, cache=cache, precision=prec)
            plan.matvec(d_d, ffmm.Velocity, out=u_d)
            return plan
        del_plan = solve()
        del del_plan
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan = solve()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t)
        if ud is None:
            ud = ffmm.direct_self(ffmm.Stokeslet3D(), pts, dens, rows=rows)
            pd = ffmm.direct_self(ffmm.Stokeslet3D(), pts, dens, ffmm.Pressure, rows=rows)
        stages = plan.stage_times()   # the solve's, before the pressure evaluation below
        # the pressure at the sources from the same expansions (Pressure(), one value a point)
        t = time.perf_counter()
        pr = plan.evaluate_at_sources(ffmm.Pressure)
        torch.cuda.synchronize()
        pressure_ms = 1e3 * (time.perf_counter() - t)
        u = u_d.cpu().numpy()[rows]
        err = float(np.linalg.norm(u - ud) / np.linalg.norm(ud))
        pr = pr.cpu().numpy()[rows]
        perr = float(np.linalg.norm(pr - pd) / np.linalg.norm(pd))
        res[prec] = {"ms_per_solve": ms, "points_per_s": n / (ms * 1e-3), "rel_l2_velocity": err,
                     "stages_ms": stages, "levels": plan.nlevels, "nodes": plan.nnodes,
                     "far_pairs": int(plan.level_pairs.sum()), "pressure_eval_ms": pressure_ms,
                     "rel_l2_pressure": perr}
        log(f"[bench] stokes {prec}: {ms:.1f} ms, {n / (ms * 1e-3):.3e} points/s, rel L2 {err:.2e} "
            f"(pressure {perr:.2e})")
        del plan
    res["mixed_over_f64"] = res["f64"]["ms_per_solve"] / res["mixed"]["ms_per_solve"]
    return res


PUBLISHED_2D = {"matvec_s": 0.31, "points": 1_000_000, "error": 8e-14,
                "case": "2D Laplace, 10^6 uniform points, KernelIndependent nequiv=60 (circles 60), matvec with cached "
                        "operators (tree build excluded), FP64, 12 Apple cores", "source": "docs/kifmm.md:169"}


def circle2d_bench(n: int, nt: int, n_pub: int, nerr: int):
    """The 2D circle method (the Laplace2D configuration: KernelIndependent nequiv=60,
    FP64, ppl=50, box ((0,0),1)): n uniform sources with N(0,1) charges and nt uniform
    targets, potential at the sources and at the targets; the plan (host quadtree and
    lists, operators per level, uploads) and the solve (expansions + evaluations) timed
    apart, warm; the error against an FP64 direct sum (numpy) at nerr sampled sources and
    targets.  Then the reference's published 2D case (n_pub uniform points, matvec =
    expansions + evaluation at the sources) beside its published time."""
    import ffmm
    torch = __import__("torch")
    rng = np.random.default_rng(21)
    m = ffmm.KernelIndependent(ffmm.Laplace2D(), nequiv=60)
    cache = ffmm.CircleOperatorCache(m)
    box = ((0.0, 0.0), 1.0)
    C = -1.0 / (4.0 * np.pi)

    def direct(src, tau, x, skip=None):
        out = np.empty(len(x))
        for a in range(0, len(x), 256):
            d = x[a:a + 256, None, :] - src[None, :, :]
            r2 = d[..., 0] ** 2 + d[..., 1] ** 2
            if skip is not None:
                r2[np.arange(len(r2)), skip[a:a + 256]] = 1.0
            out[a:a + 256] = (C * np.log(r2)) @ tau
        return out

    def run(pts, tau, tgt):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan = ffmm.CirclePlan(m, pts, ppl=50, bbox=box, cache=cache)
        t1 = time.perf_counter()
        tau_d = torch.from_numpy(tau).cuda()
        tgt_d = None if tgt is None else torch.from_numpy(tgt).cuda()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        u = plan.matvec(tau_d)
        v = None if tgt is None else plan.evaluate(tgt_d)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return plan, u, v, (t1 - t0), (t3 - t2)

    pts = rng.random((n, 2))
    tau = rng.standard_normal(n)
    tgt = rng.random((nt, 2))
    run(pts, tau, tgt)                       # warm: operators of every width, library workspaces
    plan, u, v, tplan, tsolve = run(pts, tau, tgt)
    rows = np.sort(rng.choice(n, min(nerr, n), replace=False))
    trows = np.sort(rng.choice(nt, min(nerr, nt), replace=False))
    u = u.cpu().numpy()
    v = v.cpu().numpy()
    ud = direct(pts, tau, pts[rows], rows)
    vd = direct(pts, tau, tgt[trows])
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))   # noqa: E731
    res = {"workload": f"Laplace2D KernelIndependent nequiv=60, {n} uniform sources + {nt} uniform targets, ppl=50, "
                       f"FP64, potential at sources and targets",
           "levels": plan.nlevels, "nodes": plan.nnodes, "far_pairs": int(plan.level_pairs.sum()),
           "plan_ms": 1e3 * tplan, "solve_ms": 1e3 * tsolve, "points_per_s": n / tsolve,
           "rel_l2_sources": rel(u[rows], ud), "rel_l2_targets": rel(v[trows], vd)}
    log(f"[bench] circle2d: plan {1e3 * tplan:.1f} ms, solve {1e3 * tsolve:.1f} ms, errors "
        f"{res['rel_l2_sources']:.1e} / {res['rel_l2_targets']:.1e}")
    if n_pub > 0:
        pp = rng.random((n_pub, 2))
        tp = rng.standard_normal(n_pub)
        run(pp, tp, None)
        plan, up, _, tplan, tsolve = run(pp, tp, None)
        rows = np.sort(rng.choice(n_pub, 1000, replace=False))
        err = rel(up.cpu().numpy()[rows], direct(pp, tp, pp[rows], rows))
        res["published_case"] = {"points": n_pub, "plan_ms": 1e3 * tplan, "matvec_ms": 1e3 * tsolve,
                                 "rel_l2": err, "published": PUBLISHED_2D,
                                 "speedup_vs_published": PUBLISHED_2D["matvec_s"] / tsolve}
        log(f"[bench] circle2d 1e6: matvec {1e3 * tsolve:.1f} ms (published 310 ms), plan {1e3 * tplan:.1f} ms, "
            f"error {err:.1e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--nt", type=int, default=1_000_000)
    ap.add_argument("--ppl", type=int, default=64)
    ap.add_argument("--precision", default="mixed", choices=["f64", "mixed"])
    ap.add_argument("--m2l", default="spectral", choices=["dense", "spectral"])
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 steps")
    ap.add_argument("--nerr", type=int, default=10_000)
    ap.add_argument("--serve-requests", type=int, default=2000, help="served-stream requests (0: skip the leg)")
    ap.add_argument("--serve-nmin", type=int, default=10_000)
    ap.add_argument("--serve-nmax", type=int, default=2_000_000)
    ap.add_argument("--serve-rate", type=float, default=50.0,
                    help="Poisson arrivals per second of the latency pass (about half the saturated rate)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-timeline", default=None,
                    help="write each served request's submission time, latency and size to PATH.{rate,burst}.json")
    ap.add_argument("--serve-plain-every", type=int, default=8,
                    help="the plain-form sequential leg runs every k-th request")
    ap.add_argument("--stokes-n", type=int, default=2_000_000,
                    help="points of the Stokes surface leg (0: skip the leg)")
    ap.add_argument("--circle-n", type=int, default=100_000,
                    help="sources of the 2D circle leg (0: skip the leg); targets are half as many")
    ap.add_argument("--circle-pub-n", type=int, default=1_000_000,
                    help="points of the 2D leg's published comparison (0: skip)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="before the steps, copy a buffer of this size once on the device (a kernel of known "
                         "traffic, to scale the FETCH_SIZE / WRITE_SIZE counters of a --pmc profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU: --gpus N with N > 1 is not supported "
                                   "(the library runs one GPU's share; sharding over several GPUs is not built)"}))
        sys.exit(2)

    import torch
    import ffmm
    from ffmm import workloads as W

    log(f"[bench] workload: {args.n} Gaussian-mixture sources, {args.nt} uniform targets, ppl={args.ppl}, p=8, "
        f"{args.precision}, M2L {args.m2l}")
    pts = W.gaussian_mixture(args.n, seed=0)
    tau = W.charges(args.n, seed=1)
    tgt = W.uniform(args.nt, seed=2)
    dev = "cuda:0"
    xs_d = torch.from_numpy(pts).to(dev)
    tau_d = torch.from_numpy(tau).to(dev)
    tgt_d = torch.from_numpy(tgt).to(dev)
    u_d = torch.empty(args.n, dtype=torch.float64, device=dev)
    v_d = torch.empty((args.nt, 4), dtype=torch.float64, device=dev)
    m = ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8)
    cache = ffmm.OperatorCache(m)
    box = (np.zeros(3), 1.0)
    if args.copy_probe_gib > 0:
        nb = int(args.copy_probe_gib * (1 << 30)) // 8
        a_ = torch.ones(nb, dtype=torch.float64, device=dev)
        b_ = torch.empty_like(a_)
        torch.cuda.synchronize()
        torch.mul(a_, 2.0, out=b_)          # one elementwise kernel: reads a_, writes b_
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {8 * nb} bytes read and {8 * nb} written by one elementwise (mul) kernel")
        del a_, b_
    torch.cuda.synchronize()

    def step():
        plan = ffmm.DevicePlan(m, xs_d, ppl=args.ppl, bbox=box, cache=cache, precision=args.precision, m2l=args.m2l)
        plan.build_expansions(tau_d)
        plan.evaluate_at_sources(ffmm.Potential, out=u_d)
        plan.evaluate(tgt_d, ffmm.PotentialGradient, out=v_d)
        return plan

    for w in range(args.warmup):
        t = time.perf_counter()
        plan = step()
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f} s")
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        plan = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"[bench] step {k}: {times[-1]:.3f} s")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ms = 1e3 * total / args.steps
    stages = plan.stage_times()
    log("[bench] stages (ms, last step): " + ", ".join(f"{k} {v:.2f}" for k, v in stages.items()))
    log(f"[bench] tree: {plan.nlevels} levels, {plan.nnodes} nodes, {int(plan.level_pairs.sum())} far pairs, "
        f"{int(plan.level_leaves.sum())} leaves")
    # matvec alone (plan reused): expansions + potential at the sources
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        plan.build_expansions(tau_d)
        plan.evaluate_at_sources(ffmm.Potential, out=u_d)
    torch.cuda.synchronize()
    matvec_ms = 1e3 * (time.perf_counter() - t) / args.steps

    # the plain FP64 forms on the same workload, one warm-up and one timed step each: the
    # same spectral M2L in FP64 (the like-for-like ratio of the mixed form) and the dense
    # class GEMMs (the reference device plan's form)
    plain = None
    if not args.no_plain and args.precision == "mixed":
        plain = {}
        for m2l in (("spectral", "dense") if args.m2l == "spectral" else ("dense",)):
            def plain_step():
                p = ffmm.DevicePlan(m, xs_d, ppl=args.ppl, bbox=box, cache=cache, precision="f64", m2l=m2l)
                p.build_expansions(tau_d)
                p.evaluate_at_sources(ffmm.Potential, out=u_d)
                p.evaluate(tgt_d, ffmm.PotentialGradient, out=v_d)
                return p
            plan = None
            p0 = plain_step()
            del p0
            torch.cuda.synchronize()
            t = time.perf_counter()
            p0 = plain_step()
            torch.cuda.synchronize()
            plain_ms = 1e3 * (time.perf_counter() - t)
            del p0
            plain[m2l] = {"form": f"f64 {m2l}", "ms_per_step": plain_ms, "speedup_of_measured_form": plain_ms / ms}
            log(f"[bench] plain FP64 {m2l} step: {plain_ms:.1f} ms ({plain_ms / ms:.2f}x the measured form)")
        plan = step()   # the measured form's plan again, for the checks below
        torch.cuda.synchronize()

    # accuracy against the FP64 direct sum: at nerr sampled targets (potential and
    # gradient) and at nerr sampled sources (potential)
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(args.n, min(args.nerr, args.n), replace=False))
    trows = np.sort(rng.choice(args.nt, min(args.nerr, args.nt), replace=False))
    u = u_d.cpu().numpy()
    v = v_d.cpu().numpy()
    ud = ffmm.direct_self(ffmm.Laplace3D(), pts, tau, rows=rows)
    vd = ffmm.direct(ffmm.Laplace3D(), pts, tau, tgt[trows], ffmm.PotentialGradient)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))   # noqa: E731
    err = {"rel_l2_potential_sources": rel(u[rows], ud),
           "rel_l2_potential_targets": rel(v[trows, 0], vd[:, 0]),
           "rel_l2_gradient_targets": rel(v[trows, 1:], vd[:, 1:]),
           "max_rel_potential_sources": float(np.max(np.abs(u[rows] - ud)) / np.max(np.abs(ud)))}
    err["gate_1e-6"] = bool(max(err["rel_l2_potential_sources"], err["rel_l2_potential_targets"],
                                err["rel_l2_gradient_targets"]) <= 1e-6)
    log("[bench] errors: " + json.dumps(err))

    # work of the evaluation at the sources: L2T + P2P interactions x 11 flops
    tree = plan.download_tree()
    leaves, near, npts = interaction_counts(tree)
    p2p = float(np.sum(npts * near) - args.n)            # self excluded
    l2t = float(args.n) * m.nlarge
    flops_eval = (p2p + l2t) * FLOPS_POT
    pairs = int(plan.level_pairs.sum())
    if args.m2l == "spectral":
        # the dominant kernel: the Hadamard products of the spectral M2L, one launch per
        # level (mixed: the union kernel, then on some levels the per-group kernel for the
        # tiles the union staging does not take, timed together); algorithmic work = far
        # pairs x 2304 frequencies x 8 flops (a complex multiply-add), over the summed
        # HIP-event time on the plan's stream
        nfreq = 9 * 16 * 16
        flops = float(pairs) * nfreq * 8
        t_k = stages["hadamard"] * 1e-3
        peak = PEAK_FP64_TFLOPS if args.precision == "f64" else PEAK_FP32_MFMA_TFLOPS
        kname = ("k_spec_hadamard_union + k_spec_hadamard_mfma" if args.precision == "mixed"
                 else "k_spec_hadamard<double, 8>") + " (spectral M2L Hadamard products)"
    else:
        flops = float(pairs) * 2 * m.nsurface ** 2
        t_k = stages["m2l"] * 1e-3
        peak = PEAK_FP64_TFLOPS
        kname = "dense class M2L (gather + dgemm + scatter)"
    achieved = flops / t_k / 1e12
    nlaunch = int(np.count_nonzero(plan.level_pairs))
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": traffic_from_profile(kname, args), "kernel": kname, "launches": nlaunch,
                "avg_launch_ms": 1e3 * t_k / max(nlaunch, 1), "flops_per_launch": flops / max(nlaunch, 1),
                "note": ("FP32 MFMA kernels (v_mfma_f32_16x16x4_f32) against the dense FP32 MFMA peak, a "
                         "launch being one level (the union kernel, then the per-group kernel for the "
                         "tiles it does not take); achieved counts the far pairs' products only, the "
                         "MFMAs also multiply the near pairs' zero table rows of each 8 x 8 block and the "
                         "missing colleagues' zero boxes") if args.precision == "mixed" else
                        "VALU FMA kernel: MI355X's vector FP64 and FP64-MFMA peaks are both 78.6 TFLOP/s",
                "eval_sources": {"flops": flops_eval, "ms": stages["eval_sources"],
                                 "achieved_tflops": flops_eval / (stages["eval_sources"] * 1e-3) / 1e12}}

    tree_stats = {"levels": plan.nlevels, "nodes": plan.nnodes, "far_pairs": int(plan.level_pairs.sum()),
                  "leaves": int(plan.level_leaves.sum()),
                  "gemms": gemm_shapes(plan, cache.host_operators(), args.precision == "mixed")}
    serving = None
    if args.serve_requests > 0:
        del plan                 # the measured plan's device memory goes back before the stream
        torch.cuda.synchronize()
        serving = serving_bench(args.serve_requests, args.serve_nmin, args.serve_nmax, args.serve_rate,
                                args.serve_slots, args.serve_plain_every, args.serve_timeline)
        log("[bench] serving: " + json.dumps(serving))

    stokes = None
    if args.stokes_n > 0:
        stokes = stokes_bench(args.stokes_n, args.ppl, min(args.nerr, 1000))

    circle = None
    if args.circle_n > 0:
        circle = circle2d_bench(args.circle_n, args.circle_n // 2, args.circle_pub_n, min(args.nerr, 2000))

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline: restated CPU FMM on {args.cpu_sample} points, best of {args.cpu_runs}")

        def gpu_solve(p_, t_):
            pl = ffmm.DevicePlan(m, p_, ppl=args.ppl, bbox=box, cache=cache, precision=args.precision, m2l=args.m2l)
            return pl.matvec(t_).cpu().numpy()
        cpu = cpu_baseline(args.cpu_sample, args.ppl, args.cpu_runs, gpu_solve)
        log("[bench] cpu baseline: " + json.dumps(cpu))
        err["rel_l2_vs_restated_cpu_fmm"] = cpu["gpu_rel_l2_vs_restated_cpu_fmm"]
        err["restated_cpu_fmm_sample"] = args.cpu_sample

    line = {"metric": "Laplace3D 4M-pt FMM points/s at 1e-6 rel L2; served p50/p99 solve latency",
            "value": args.n / (ms * 1e-3), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if args.precision == "f64" else "f32+f64",
            "data": "synthetic (seeded Gaussian mixture / uniform)",
            "config": {"workload": f"Laplace3D CubeKI p=8, {args.n} Gaussian-mixture sources + {args.nt} uniform "
                                   f"targets, potential at sources + potential and gradient at targets, ppl={args.ppl}, "
                                   f"{args.precision}, M2L {args.m2l}",
                       "nsources": args.n, "ntargets": args.nt, "ppl": args.ppl, "p": 8},
            "errors": err, "stages_ms": stages, "matvec_ms": matvec_ms,
            "tree": tree_stats,
            "roofline": roofline, "cpu_baseline": cpu, "serving": serving, "plain_f64": plain,
            "stokes_surface": stokes, "laplace2d_circle": circle}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
