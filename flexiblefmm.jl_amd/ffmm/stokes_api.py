"""The Stokes front end of the reference (src/stokes_api.jl) over the device plan.

``StokesFMMPlan(sources; targets, method, p, ppl, bbox)`` holds the tree and operators
of a fixed geometry; ``stokes_fmm_eval_`` (``stokes_fmm_eval!``) and
``stokes_fmm_eval_src_trg_`` (``stokes_fmm_eval_src_trg!``) take separate arrays of
stokeslets, stresslets with their stressvecs (normals), optional per-source weights,
and evaluate the velocity at the sources (self, i = j skipped) or at the plan's
targets; ``StokesFMMKernel`` is the one-shot form (src/stokes_api.jl:160-186).

On the device the 3D method is ``StokesHarmonics`` (``method="harmonics"``, the
reference's analytic Stokes method; its default p here is the reference's 16).  The
kernel-independent cube method for the Stokes kernel, doublets and velocity gradients
are not built: asking for them raises ``ArgumentError``; ``output=Pressure`` gives the
pressure.  The densities
become the device plan's (N, 9) rows [f, g, nu]: stokeslet f, stresslet strength g and
normal nu, each strength multiplied by the source's weight.
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, DimensionMismatch
from .fmm import DevicePlan, OperatorCache, Pressure, StokesHarmonics, Stokeslet3D, Velocity, _output, square_bbox


def _points(x) -> np.ndarray:
    """N x 3 from N x 3 or 3 x N (the reference accepts D x N matrices)."""
    a = np.asarray(x.cpu().numpy() if hasattr(x, "data_ptr") else x, dtype=np.float64)
    if a.ndim != 2 or 3 not in a.shape:
        raise ArgumentError("points must be N x 3 or 3 x N")
    return np.ascontiguousarray(a if a.shape[1] == 3 else a.T)


def stokes_densities(n: int, weights=None, stokeslets=None, stresslets=None, stressvecs=None,
                     doublets=None) -> np.ndarray:
    """The (n, 9) rows [w f, w g, nu] (src/stokes_api.jl:72-83)."""
    if (stresslets is None) != (stressvecs is None):
        raise ArgumentError("stresslets and stressvecs must be given together")
    for name, a in (("weights", weights), ("stokeslets", stokeslets), ("stresslets", stresslets),
                    ("stressvecs", stressvecs), ("doublets", doublets)):
        if a is not None and len(a) != n:
            raise DimensionMismatch(f"{name} has length {len(a)}, expected {n}")
    if stokeslets is None and stresslets is None and doublets is None:
        raise ArgumentError("provide at least one of stokeslets, stresslets, doublets")
    if doublets is not None:
        raise ArgumentError("doublets are not built on the device")
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64).reshape(n)
    d = np.zeros((n, 9))
    if stokeslets is not None:
        d[:, 0:3] = _points(stokeslets) * w[:, None] if n else 0.0
    if stresslets is not None:
        d[:, 3:6] = _points(stresslets) * w[:, None] if n else 0.0
        d[:, 6:9] = _points(stressvecs)
    return d


class StokesFMMPlan:
    """``StokesFMMPlan(sources; targets=nothing, method, p, ppl, bbox, cache)``
    (src/stokes_api.jl:40-62): the device plan of a fixed Stokes geometry."""

    def __init__(self, sources, targets=None, method: str = "harmonics", p: int | None = None, ppl: int | None = None,
                 bbox=None, cache: OperatorCache | None = None, precision: str = "f64", device: int = 0):
        if method in ("cube", "square", "ki", "bb"):
            raise ArgumentError(f"the {method} method for the Stokes kernel is not built on the device; "
                                "use method='harmonics'")
        if method != "harmonics":
            raise ArgumentError("method must be :ki, :square (= :cube), :bb or :harmonics")
        p = 16 if p is None else int(p)
        ppl = 200 if ppl is None else int(ppl)     # 3D trees fragment on surfaces and fibres
        pts = _points(sources)
        tgts = None if targets is None else _points(targets)
        if bbox is None:
            allp = pts if tgts is None else np.vstack([pts, tgts])
            # per coordinate: numpy's axis-0 reduction of an N x 3 array is several times slower
            bbox = square_bbox(np.array([allp[:, d].min() for d in range(3)]),
                               np.array([allp[:, d].max() for d in range(3)]))
        m = StokesHarmonics(Stokeslet3D(), p=p)
        if cache is None:
            cache = OperatorCache(m)
        self.method, self.cache = m, cache
        self.sources, self.targets = pts, tgts
        self.plan = DevicePlan(m, pts, ppl=ppl, bbox=bbox, cache=cache, precision=precision, device=device)

    @property
    def n(self) -> int:
        return len(self.sources)


def _check_output(output):
    """Velocity (the default) or Pressure (src/stokes_api.jl:100-104); the output object."""
    o = _output(output)
    if type(o) not in (type(_output(Velocity)), type(_output(Pressure))):
        raise ArgumentError("the device Stokes path evaluates the velocity or the pressure")
    return o


def _check_same(given, planned, what: str):
    if given is not None and not np.array_equal(_points(given), planned):
        raise ArgumentError(f"the {what} differ from the plan's (the device library takes no dual-number positions)")


def stokes_fmm_eval_(velocities, plan: StokesFMMPlan, *, stokeslets=None, stresslets=None, stressvecs=None,
                     doublets=None, weights=None, sources=None, targets=None, output=Velocity):
    """``stokes_fmm_eval!`` (src/stokes_api.jl:107-123): the velocity (or, with
    ``output=Pressure``, the pressure) at the plan's targets, or at the sources for a
    self plan, into ``velocities`` (n x 3; n for the pressure)."""
    o = _check_output(output)
    _check_same(sources, plan.sources, "sources")
    d = stokes_densities(plan.n, weights, stokeslets, stresslets, stressvecs, doublets)
    p = plan.plan
    p.build_expansions(d)
    if plan.targets is None and targets is None:
        v = p.evaluate_at_sources(o).cpu().numpy()
    else:
        tg = plan.targets if targets is None else _points(targets)
        v = p.evaluate(tg, o).cpu().numpy()
    if velocities is None:
        return v
    if np.shape(velocities) != v.shape:
        raise DimensionMismatch(f"expected {len(v)} velocities")
    velocities[...] = v
    return velocities


def stokes_fmm_eval_src_trg_(vel_src, vel_trg, plan: StokesFMMPlan, *, stokeslets=None, stresslets=None,
                             stressvecs=None, doublets=None, weights=None, sources=None, targets=None,
                             output=Velocity):
    """``stokes_fmm_eval_src_trg!`` (src/stokes_api.jl:131-146): one expansion build,
    the velocity (or the pressure) at the sources (self) and at the plan's targets."""
    if plan.targets is None and targets is None:
        raise ArgumentError("the plan has no targets")
    o = _check_output(output)
    _check_same(sources, plan.sources, "sources")
    d = stokes_densities(plan.n, weights, stokeslets, stresslets, stressvecs, doublets)
    p = plan.plan
    p.build_expansions(d)
    vs = p.evaluate_at_sources(o).cpu().numpy()
    tg = plan.targets if targets is None else _points(targets)
    vt = p.evaluate(tg, o).cpu().numpy()
    if vel_src is not None:
        vel_src[...] = vs
    if vel_trg is not None:
        vel_trg[...] = vt
    return (vs if vel_src is None else vel_src), (vt if vel_trg is None else vel_trg)


def StokesFMMKernel(sources, *, weights=None, targets=None, stokeslets=None, stresslets=None, stressvecs=None,
                    doublets=None, force_self_eval: bool = False, output=Velocity, **kwargs) -> np.ndarray:
    """``StokesFMMKernel(sources; ...)`` (src/stokes_api.jl:172-186): a plan from the
    positions, then one evaluation; the velocities are returned."""
    self_eval = targets is None or targets is sources or force_self_eval
    plan = StokesFMMPlan(sources, targets=None if self_eval else targets, **kwargs)
    return stokes_fmm_eval_(None, plan, stokeslets=stokeslets, stresslets=stresslets, stressvecs=stressvecs,
                            doublets=doublets, weights=weights, output=output)
