"""The host API, mirroring the reference's Julia interface on top of libffmm_hip.

Reference entry points and their counterparts here (same argument meaning, same
error phrases; ``!`` becomes a trailing underscore):

=====================================================  ==========================================
``Laplace3D()``  (src/kernels.jl:179-200)               ``Laplace3D()``
``CubeKernelIndependent{3}(K; p, pc, d, pinv_rtol)``    ``CubeKernelIndependent(K, p=, pc=, d=)``
  (src/operators/cubeki.jl:18-48)
``OperatorCache(method)``, ``fit_box!``                 ``OperatorCache(method)``, ``.fit_box``
  (src/operators/cache.jl:5-83)
``FMM(method, sources; ppl, bbox, cache, max_levels)``  ``FMM(method, sources, ppl=, bbox=, ...)``
  (src/fmm.jl:31-49)
``build_expansions(f, sources, tau)`` (src/fmm.jl:306)  ``build_expansions(f, sources, tau)``
``build_expansions!(state, f, sources, tau)`` (:271)    ``build_expansions_(state, f, sources, tau)``
``evaluate_at_sources(f, st; output)``                  ``evaluate_at_sources(f, st, output=)``
  (src/evaluate.jl:129-132)
``evaluate(f, st, targets; output)`` (src/evaluate.jl:98)             ``evaluate(f, st, targets, output=)``
``DevicePlan(m, xs; ppl, bbox, cache, max_levels)``     ``DevicePlan(m, xs, ppl=, bbox=, ...)``
  (src/device_passes.jl:113-134)
``device_build_expansions!(plan, ds; sorted)`` (:511)   ``device_build_expansions_(plan, ds, sorted=)``
``device_matvec!(out, plan, ds; output, sorted)`` (:619)``device_matvec_(out, plan, ds, output=, ...)``
``device_tree`` / ``Tree(dt)`` (src/tree/device_builder.jl:398, 508)  ``DevicePlan.download_tree()``
=====================================================  ==========================================

Device memory and the hand-off of device pointers go through torch tensors (plumbing
only); every computation runs in libffmm_hip.
"""
from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, call
from . import operators as _ops


# ---------------------------------------------------------------------------
# kernels, outputs, methods
# ---------------------------------------------------------------------------
class Laplace3D:
    """Free-space Laplace Green's function 1/(4 pi |t - s|) (src/kernels.jl:175-200)."""
    homogeneity = -1

    def __repr__(self):
        return "Laplace3D()"


class Yukawa3D:
    """Modified Helmholtz (Yukawa) Green's function e^{-k r}/(4 pi r) (src/kernels.jl:220-240)."""
    homogeneity = None

    def __init__(self, k: float):
        self.k = float(k)

    def __repr__(self):
        return f"Yukawa3D({self.k})"


class Stokeslet3D:
    """The Stokeslet (1/8 pi)(I + r r^T / r^2)/r with unit viscosity (src/kernels.jl:241-272)."""
    homogeneity = -1

    def __repr__(self):
        return "Stokeslet3D()"


class Potential:
    code = _lib.POTENTIAL
    ncomp = 1


Velocity = Potential     # the reference's alias for the Stokes kernels (src/sources.jl:31-44)


class Gradient:
    code = _lib.GRADIENT
    ncomp = 3


class PotentialGradient:
    """Potential and gradient from one evaluation (4 values per point: u, du/dx, du/dy, du/dz)."""
    code = _lib.POTENTIAL_GRADIENT
    ncomp = 4


class Pressure:
    """The pressure of the Stokes kernels, one value per point (src/sources.jl:34-42)."""
    code = _lib.PRESSURE
    ncomp = 1


def _output(o):
    o = o() if isinstance(o, type) else o
    if not hasattr(o, "code"):
        raise ArgumentError(f"output {o!r} is not available")
    return o


class CubeKernelIndependent:
    """``CubeKernelIndependent{3}(kernel; p=8, pc=2p, d=0.1, pinv_rtol=50 eps)``
    (src/operators/cubeki.jl:18-48): p points per edge on the small surface, pc on the
    large one, margin d."""

    def __init__(self, kernel, p: int = 8, pc: int | None = None, d: float = 0.1,
                 pinv_rtol: float = 50 * np.finfo(np.float64).eps):
        pc = 2 * p if pc is None else pc
        if p < 2:
            raise ArgumentError("p must be at least 2")
        if pc < p:
            raise ArgumentError("pc must be at least p")
        if not (0 < d < 2 / 3):
            raise ArgumentError("d must lie in (0, 2/3): the large surface, at (3 − 2d)w/2, must lie outside the "
                                "small one, at (1 + d)w/2")
        if not isinstance(kernel, Laplace3D):
            raise ArgumentError("the device library takes Laplace3D")
        self.kernel, self.p, self.pc, self.d, self.pinv_rtol = kernel, int(p), int(pc), float(d), float(pinv_rtol)

    @property
    def nsurface(self) -> int:
        return self.p ** 3 - (self.p - 2) ** 3

    @property
    def nlarge(self) -> int:
        return self.pc ** 3 - (self.pc - 2) ** 3

    def __repr__(self):
        return f"CubeKernelIndependent{{3}}({self.kernel!r}; p={self.p}, pc={self.pc}, d={self.d})"

    @property
    def expansion_length(self) -> int:
        return self.nsurface


class SphericalHarmonics:
    """``SphericalHarmonics(kernel; p=16, tol=0)`` for Laplace3D or Yukawa3D charges
    (src/operators/harmonics.jl:313-322, src/operators/harmonics_yukawa.jl:27-33):
    solid-harmonic expansions of degree p, dense class M2L (tol = 0)."""

    def __init__(self, kernel, p: int = 16, tol: float = 0.0):
        if not isinstance(kernel, (Laplace3D, Yukawa3D)):
            raise ArgumentError("SphericalHarmonics takes Laplace3D or Yukawa3D")
        if isinstance(kernel, Yukawa3D):
            if not 1 <= p <= 60:
                raise ArgumentError("p must lie in 1:60")
        elif p < 1:
            raise ArgumentError("p must be at least 1")
        if not 0 <= tol < 1:
            raise ArgumentError("tol must lie in [0, 1)")
        if tol > 0:
            raise ArgumentError("the device library takes the dense M2L (tol = 0)")
        if p > 60:
            raise ArgumentError("the device library takes p <= 60")
        self.kernel, self.p, self.tol = kernel, int(p), float(tol)

    nblocks = 1

    @property
    def expansion_length(self) -> int:
        return self.nblocks * (self.p + 1) ** 2

    def __repr__(self):
        return f"SphericalHarmonics({self.kernel!r}; p={self.p})"


class StokesHarmonics(SphericalHarmonics):
    """``StokesHarmonics(Stokeslet3D(); p=16, tol=0)`` (src/operators/harmonics_stokes.jl:34-40):
    four Laplace expansions [Phi_1; Phi_2; Phi_3; Psi] per box; stokeslet and stresslet
    densities, velocity output."""
    nblocks = 4

    def __init__(self, kernel, p: int = 16, tol: float = 0.0):
        if not isinstance(kernel, Stokeslet3D):
            raise ArgumentError("StokesHarmonics takes Stokeslet3D")
        if p < 1:
            raise ArgumentError("p must be at least 1")
        if not 0 <= tol < 1:
            raise ArgumentError("tol must lie in [0, 1)")
        if tol > 0:
            raise ArgumentError("the device library takes the dense M2L (tol = 0)")
        if p > 60:
            raise ArgumentError("the device library takes p <= 60")
        self.kernel, self.p, self.tol = kernel, int(p), float(tol)

    def __repr__(self):
        return f"StokesHarmonics({self.kernel!r}; p={self.p})"


def _is_harmonic(m) -> bool:
    return isinstance(m, SphericalHarmonics)


# ---------------------------------------------------------------------------
# device contexts (one per device, on its own stream)
# ---------------------------------------------------------------------------
_CONTEXTS: dict = {}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU is visible: libffmm_hip needs an MI355X (gfx950)")
    return torch


def context(device: int = 0):
    h = _CONTEXTS.get(device)
    if h is None:
        _torch()
        h = ctypes.c_void_p()
        call("ffmm_context_create", device, None, ctypes.byref(h))
        _CONTEXTS[device] = h
    return h


def _dev_f64(x, device: int, cols: int | None = None):
    """A contiguous float64 tensor on the device (copied when it is not one already)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x.to(device=f"cuda:{device}", dtype=torch.float64).contiguous()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).to(f"cuda:{device}")
    if cols is not None:
        if t.ndim == 1 and cols == 1:
            pass
        elif t.ndim != 2 or t.shape[1] != cols:
            raise DimensionMismatch(f"expected an N × {cols} array, got {tuple(t.shape)}")
    torch.cuda.synchronize(device)
    return t


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# direct sums (src/kernels.jl:442-506), on the device in FP64
# ---------------------------------------------------------------------------
def direct(K, sources, tau, targets, output=Potential, skip=None, device: int = 0) -> np.ndarray:
    """``direct(K, sources, tau, targets; output)``: the O(N M) sum at the targets,
    excluding source ``skip[j]`` at target ``j`` when ``skip`` is given."""
    if isinstance(K, Stokeslet3D):
        o = _output(output)
        if o.code not in (_lib.POTENTIAL, _lib.PRESSURE):
            raise ArgumentError("the Stokes direct sum gives the velocity (Potential) or the Pressure")
        return _direct_stokes(sources, tau, targets, skip, device, pressure=o.code == _lib.PRESSURE)
    if not isinstance(K, Laplace3D):
        raise ArgumentError("the device direct sum takes Laplace3D or Stokeslet3D")
    o = _output(output)
    torch = _torch()
    s = _dev_f64(sources, device, cols=3)
    t = _dev_f64(tau, device)
    x = _dev_f64(targets, device, cols=3)
    if t.ndim != 1 or t.shape[0] != s.shape[0]:
        raise DimensionMismatch("sources and τ differ in length")
    nt = x.shape[0]
    out = torch.empty((nt,) if o.ncomp == 1 else (nt, o.ncomp), dtype=torch.float64, device=f"cuda:{device}")
    sk = None
    if skip is not None:
        sk = torch.as_tensor(np.asarray(skip, dtype=np.int64)).to(f"cuda:{device}")
        if sk.shape[0] != nt:
            raise DimensionMismatch("skip and targets differ in length")
    call("ffmm_direct", context(device), _ptr(s), _ptr(t), s.shape[0], _ptr(x), nt,
         None if sk is None else _ptr(sk), _ptr(out), o.code)
    return out.cpu().numpy()


def _direct_stokes(sources, dens, targets, skip, device: int, pressure: bool = False) -> np.ndarray:
    """The Stokes velocity (or the pressure) at the targets: dens (N, 3) stokeslets or
    (N, 9) rows [f, g, nu]."""
    torch = _torch()
    s = _dev_f64(sources, device, cols=3)
    d = _dev_f64(dens, device)
    x = _dev_f64(targets, device, cols=3)
    if d.ndim != 2 or d.shape[1] not in (3, 9) or d.shape[0] != s.shape[0]:
        raise DimensionMismatch("expected (N, 3) stokeslets or (N, 9) rows [f, g, nu]")
    nt = x.shape[0]
    out = torch.empty((nt,) if pressure else (nt, 3), dtype=torch.float64, device=f"cuda:{device}")
    sk = None
    if skip is not None:
        sk = torch.as_tensor(np.asarray(skip, dtype=np.int64)).to(f"cuda:{device}")
        if sk.shape[0] != nt:
            raise DimensionMismatch("skip and targets differ in length")
    call("ffmm_direct_stokes_pressure" if pressure else "ffmm_direct_stokes", context(device), _ptr(s), _ptr(d), int(d.shape[1]), s.shape[0], _ptr(x), nt,
         None if sk is None else _ptr(sk), _ptr(out))
    return out.cpu().numpy()


def direct_self(K, sources, tau, output=Potential, rows=None, device: int = 0) -> np.ndarray:
    """``direct_self(K, sources, tau; output)``: the sum at every source (or at the
    sources ``rows``) over all other sources."""
    src = np.asarray(sources, dtype=np.float64) if not hasattr(sources, "data_ptr") else sources.cpu().numpy()
    rows = np.arange(len(src), dtype=np.int64) if rows is None else np.asarray(rows, dtype=np.int64)
    return direct(K, src, tau, src[rows], output, skip=rows, device=device)


# ---------------------------------------------------------------------------
# operator cache (src/operators/cache.jl)
# ---------------------------------------------------------------------------
class OperatorCache:
    """Operators of one method by width exponent ``w = ldexp(W0, e)``; for the
    homogeneous Laplace kernel one reference set at ``W0`` serves every width.  The
    device copies are built once per device and shared by every plan made with this
    cache (OperatorCache.device, src/device_passes.jl:136-146)."""

    def __init__(self, method):
        self.method = method
        self.W0 = math.nan
        self._host = None
        self._device = {}
        self._widths = {}          # harmonic methods: width -> HarmonicOperators (Yukawa: one per width)
        self._put = {}             # device -> widths uploaded
        self._factors = {}         # (device, tol) -> ranks of the uploaded low-rank class factors

    def has_base_width(self) -> bool:
        return not math.isnan(self.W0)

    def fit_box(self, lo, hi, bbox):
        """``fit_box!(cache, lo, hi, bbox)`` (src/operators/cache.jl:57-83): a box of
        width ``ldexp(W0, e)`` containing ``bbox``; the first call fixes ``W0``."""
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        if not (np.all(np.isfinite(lo)) and np.all(np.isfinite(hi))):
            raise ArgumentError("the points must be finite")
        origin, W = np.asarray(bbox[0], dtype=np.float64).copy(), float(bbox[1])
        if not W > 0:
            raise ArgumentError("box width must be positive")
        if not (np.all(origin <= lo) and np.all(origin + W >= hi)):
            raise ArgumentError("points lie outside the requested box")
        if not self.has_base_width():
            self.W0 = W
            return origin, W, 0
        f, e = math.frexp(W / self.W0)
        e = e - 1 if f == 0.5 else e
        cen = origin + W / 2
        for _ in range(2):
            Wn = math.ldexp(self.W0, e)
            on = np.minimum(cen - Wn / 2, lo)
            if np.all(on + Wn >= hi):
                return on, Wn, e
            e += 1
        raise ArgumentError("could not fit a box")

    def host_operators(self):
        if not self.has_base_width():
            raise ArgumentError("fit a box before requesting operators")
        if self._host is None:
            m = self.method
            if _is_harmonic(m):
                self._host = self.harmonic_operators(self.W0)
            else:
                self._host = _ops.level_operators(self.W0, p=m.p, pc=m.pc, d=m.d, rtol=m.pinv_rtol)
        return self._host

    def harmonic_operators(self, w: float):
        """``get_operators!`` of a harmonic method for boxes of width ``w``
        (src/operators/cache.jl:33-46): Laplace and Stokes take the reference set of
        width W0 for every width (homogeneous), Yukawa one set per width."""
        from . import harmonics as _h
        m = self.method
        if not isinstance(m.kernel, Yukawa3D):
            w = self.W0
        ops = self._widths.get(w)
        if ops is None:
            if isinstance(m.kernel, Yukawa3D):
                ops = _h.yukawa_level_operators(w, m.p, m.kernel.k)
            else:
                ops = _h.laplace_level_operators(w, m.p)
            self._widths[w] = ops
        return ops

    def _put_width(self, device: int, w: float):
        ops = self.harmonic_operators(w)
        put = self._put.setdefault(device, set())
        if ops.width in put:
            return
        F = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731
        M2M, L2L = F(ops.M2M), F(ops.L2L)
        reps = np.ascontiguousarray(np.stack([F(r).T for r in ops.reps]))   # each N x N column major
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_opcache_put_harmonic_level", self._device[device], ops.width, P(M2M), P(L2L), P(reps))
        put.add(ops.width)

    def _device_harmonic(self, device: int):
        from . import harmonics as _h
        m = self.method
        kern = (_lib.KERNEL_STOKESLET if isinstance(m, StokesHarmonics) else
                _lib.KERNEL_YUKAWA if isinstance(m.kernel, Yukawa3D) else _lib.KERNEL_LAPLACE)
        kappa = m.kernel.k if isinstance(m.kernel, Yukawa3D) else 0.0
        h = ctypes.c_void_p()
        call("ffmm_opcache_create_harmonic", context(device), kern, m.p, kappa, self.W0, ctypes.byref(h))
        self._device[device] = h
        sc = _h.harmonic_symmetry_classes(m.p)
        keep = [np.ascontiguousarray(sc.klass, dtype=np.int32), np.ascontiguousarray(sc.idx, dtype=np.int32),
                np.ascontiguousarray(sc.sgn, dtype=np.int8), np.ascontiguousarray(sc.idx_t, dtype=np.int32),
                np.ascontiguousarray(sc.sgn_t, dtype=np.int8)]
        call("ffmm_opcache_put_harmonic_classes", h, len(sc.reps), *[a.ctypes.data_as(ctypes.c_void_p) for a in keep])
        if not isinstance(m.kernel, Yukawa3D):
            self._put_width(device, self.W0)
        return h

    def ensure_widths(self, device: int, widths):
        """Upload the operators of these box widths (Yukawa: per width)."""
        self.device(device)
        for w in widths:
            self._put_width(device, float(w))

    def device(self, device: int = 0):
        h = self._device.get(device)
        if h is None and _is_harmonic(self.method):
            if not self.has_base_width():
                raise ArgumentError("fit a box before requesting operators")
            return self._device_harmonic(device)
        if h is None:
            o = self.host_operators()
            m = self.method
            h = ctypes.c_void_p()
            call("ffmm_opcache_create", context(device), m.p, m.pc, m.d, ctypes.byref(h))
            F = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731 (column major)
            keep = dict(small=np.ascontiguousarray(o.small), large=np.ascontiguousarray(o.large),
                        Uh_up=F(o.P_up.Uh), invS_up=np.ascontiguousarray(o.P_up.invS), V_up=F(o.P_up.V),
                        Uh_dn=F(o.P_down.Uh), invS_dn=np.ascontiguousarray(o.P_down.invS), V_dn=F(o.P_down.V),
                        M2M=F(o.M2M), L2L=F(o.L2L),
                        class_ops=np.ascontiguousarray(np.stack([c.T for c in o.class_ops])),
                        far_class=np.ascontiguousarray(o.sc.klass, dtype=np.int32),
                        idx=np.ascontiguousarray(o.sc.idx, dtype=np.int32),
                        sgn=np.ascontiguousarray(o.sc.sgn, dtype=np.int8),
                        tables=np.ascontiguousarray(o.tables.astype(np.complex128)))
            P = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
            call("ffmm_opcache_put_cube", h, o.width, P(keep["small"]), P(keep["large"]), o.P_up.rank,
                 P(keep["Uh_up"]), P(keep["invS_up"]), P(keep["V_up"]), o.P_down.rank, P(keep["Uh_dn"]),
                 P(keep["invS_dn"]), P(keep["V_dn"]), P(keep["M2M"]), P(keep["L2L"]), len(o.class_ops),
                 P(keep["class_ops"]), P(keep["far_class"]), P(keep["idx"]), P(keep["sgn"]), P(keep["tables"]))
            self._device[device] = h
        return h

    def class_factors(self, device: int, tol: float):
        """The low-rank class factors of tolerance ``tol`` on the device (host SVDs of the
        class representatives, ``class_factors``, src/operators/symmetry.jl:169-178),
        uploaded once per tolerance; their ranks."""
        if _is_harmonic(self.method):
            raise ArgumentError("tol > 0 takes the cube method (low-rank class operators)")
        got = self._factors.get((device, tol))
        if got is not None:
            return got
        oc = self.device(device)
        ranks, A, Bt = _ops.class_factors(self.host_operators(), tol)
        r32 = np.ascontiguousarray(ranks, dtype=np.int32)
        a = np.ascontiguousarray(np.concatenate([x.ravel(order="F") for x in A]))
        b = np.ascontiguousarray(np.concatenate([x.ravel(order="F") for x in Bt]))
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_opcache_put_class_factors", oc, float(tol), P(r32), P(a), P(b))
        self._factors[(device, tol)] = ranks
        return ranks

    def __len__(self):
        return 0 if self._host is None else 1


def plan_data_check(data: bytes) -> None:
    """The host checks of plan data (``ffmm_plan_data_check``): raises ``ArgumentError``
    with the first violation of the tree's structure; needs no device."""
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    call("ffmm_plan_data_check", buf, len(data))


def square_bbox(lo, hi):
    """The square box centred on the bounding box ``[lo, hi]`` with side its largest
    extent (1 if all points coincide), widened until it contains it
    (src/tree/tree.jl:97-114)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    if not (np.all(np.isfinite(lo)) and np.all(np.isfinite(hi))):
        raise ArgumentError("the points must be finite")
    W0 = float(np.max(hi - lo))
    if not math.isfinite(W0):
        raise ArgumentError("the points' extent overflows")
    if not W0 > 0:
        W0 = 1.0
    c = lo + (hi - lo) / 2
    origin = c - W0 / 2
    for _ in range(64):
        if np.all(origin <= lo) and np.all(origin + W0 >= hi):
            return origin, W0
        W0 *= 1 + 4 * np.finfo(np.float64).eps
        origin = c - W0 / 2
    raise ArgumentError("could not fit a box around the points")


# ---------------------------------------------------------------------------
# the device plan
# ---------------------------------------------------------------------------
_PREC = {"f64": _lib.PREC_F64, "mixed": _lib.PREC_MIXED}
_M2L = {"dense": _lib.M2L_DENSE, "spectral": _lib.M2L_SPECTRAL}


@dataclass
class TreeArrays:
    """The flat tree downloaded from the device (0-based, -1 = none): the fields of the
    reference's ``Tree`` (src/tree/tree.jl:40-72) with ``points`` as ``plo``/``phi``."""
    origin: np.ndarray
    W0: float
    level_start: np.ndarray
    cell: np.ndarray
    level: np.ndarray
    parent: np.ndarray
    child_first: np.ndarray
    plo: np.ndarray
    phi: np.ndarray
    leaf: np.ndarray
    ghost: np.ndarray
    xlist: np.ndarray
    has_source: np.ndarray
    colleagues: np.ndarray
    perm: np.ndarray
    point_node: np.ndarray
    points: np.ndarray

    @property
    def nlevels(self):
        return len(self.level_start) - 1


def _check_plan_kwargs(method, m2l: str, tol: float, fast: bool, eltype):
    """The reference's DevicePlan keywords (src/device_passes.jl:90, 113-115): ``tol`` >
    0 (low-rank class operators) takes the cube method with the dense class M2L; the
    operators are FP64, and ``eltype`` must hold them ("eltype must hold the operators'
    element type Float64", src/device_passes.jl:168-169); ``fast`` (relaxed arithmetic
    in the evaluation, src/device.jl:12-20) is not offered: ``precision="mixed"`` is the
    reduced-precision form."""
    if not 0 <= tol < 1:
        raise ArgumentError("tol must lie in [0, 1)")
    if tol > 0 and (_is_harmonic(method) or m2l != "dense"):
        raise ArgumentError("tol > 0 takes the cube method with the dense class M2L")
    if eltype is not None and np.dtype(eltype) not in (np.dtype(np.float64), np.dtype(np.complex128)):
        raise ArgumentError("eltype must hold the operators' element type Float64")
    if fast:
        raise ArgumentError("fast=true (relaxed evaluation arithmetic) is not offered; precision=\"mixed\" is the "
                            "reduced-precision form")


class DevicePlan:
    """``DevicePlan(method, xs; ppl=50, bbox, cache, max_levels=30)``: the device tree,
    its lists, the operators' device copies and the work buffers
    (src/device_passes.jl:98-134).  ``precision`` is ``"f64"`` or ``"mixed"``; ``m2l`` is
    ``"dense"`` (class GEMMs) or ``"spectral"`` (batched FFT).

    ``xs`` may be a tuple of point sets, one per group (``DevicePlan(method, groups)``,
    src/device_passes.jl:113-134): one tree over all of them; densities are then given per
    group (``None``: targets only) and outputs come back per group."""

    def __init__(self, method, xs, *, ppl: int = 50, bbox=None, cache: OperatorCache = None,
                 max_levels: int = 30, precision: str = "f64", m2l: str = "dense", device: int = 0,
                 tol: float = 0.0, fast: bool = False, eltype=None):
        _check_plan_kwargs(method, m2l, tol, fast, eltype)
        if isinstance(xs, tuple):
            if len(xs) == 0:
                raise ArgumentError("no groups")
            parts = [_dev_f64(g, device, cols=3) for g in xs]
            self.group_sizes = [int(g.shape[0]) for g in parts]
            xs = parts[0] if len(parts) == 1 else _torch().cat(parts, 0)
        else:
            self.group_sizes = None
        cache = OperatorCache(method) if cache is None else cache
        if cache.method is not method:
            raise ArgumentError("the cache was built for a different method")
        if precision not in _PREC:
            raise ArgumentError(f"unknown precision {precision!r}")
        if m2l not in _M2L:
            raise ArgumentError(f"unknown M2L form {m2l!r}")
        self.method, self.cache, self.device = method, cache, device
        self.precision, self.m2l = precision, m2l
        xs = _dev_f64(xs, device, cols=3)
        n = xs.shape[0]
        if n == 0:
            raise ArgumentError("no points")
        ctx = context(device)
        lo, hi, fin = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), ctypes.c_int()
        call("ffmm_bounds", ctx, _ptr(xs), n, lo, hi, ctypes.byref(fin))
        if not fin.value:
            raise ArgumentError("the points must be finite")
        lo, hi = np.array(lo[:]), np.array(hi[:])
        if bbox is None:
            bbox = square_bbox(lo, hi)
        origin, W, e0 = cache.fit_box(lo, hi, bbox)
        self.origin, self.W0, self.e0 = np.asarray(origin, dtype=np.float64), float(W), int(e0)
        oc = cache.device(device)
        h = ctypes.c_void_p()
        org = (ctypes.c_double * 3)(*self.origin)
        call("ffmm_plan_create", oc, _ptr(xs), n, int(ppl), org, self.W0, self.e0, int(max_levels), _PREC[precision],
             _M2L[m2l], ctypes.byref(h))
        self._h = h
        self.N = n
        self.ppl = int(ppl)
        if self.group_sizes is None:
            self.group_sizes = [n]
        self.group_offsets = np.concatenate([[0], np.cumsum(self.group_sizes)]).astype(np.int64)
        self._gperm = None
        self._read_sizes()
        if _is_harmonic(method) and isinstance(method.kernel, Yukawa3D):
            # the per-width Yukawa operators of the levels with translations (2 .. L-1)
            cache.ensure_widths(device, [math.ldexp(self.W0, -l) for l in range(2, self.nlevels)])
        self.tol = 0.0
        if tol > 0:
            self.set_tol(tol)

    def set_tol(self, tol: float):
        """The class M2L through the low-rank class factors of tolerance ``tol``
        (``DevicePlan(...; tol)``, src/device_passes.jl:14-27, 570-576; 0: the dense
        representatives); the per-class ranks are ``self.class_ranks``.  A captured pass
        sequence is dropped."""
        self._captured = None
        tol = float(tol)
        self.class_ranks = self.cache.class_factors(self.device, tol) if tol > 0 else None
        call("ffmm_plan_set_tol", self._h, tol)
        self.tol = tol

    # -- plan data: warm restarts without the tree build --
    _FILE_MAGIC = b"FFMMPLAN\n"

    def save(self, path) -> None:
        """Write the plan's tree (``ffmm_plan_save``) to ``path``, after a one-line JSON
        header with the host-side fields (box, groups, forms).  ``DevicePlan.load`` makes
        the same plan from it without building the tree again."""
        size = ctypes.c_int64()
        call("ffmm_plan_save", self._h, None, 0, ctypes.byref(size))
        buf = (ctypes.c_uint8 * size.value)()
        call("ffmm_plan_save", self._h, buf, size.value, ctypes.byref(size))
        head = dict(N=self.N, ppl=self.ppl, origin=[float(v) for v in self.origin], W0=self.W0, e0=self.e0,
                    precision=self.precision, m2l=self.m2l, group_sizes=[int(g) for g in self.group_sizes])
        with open(path, "wb") as f:
            f.write(self._FILE_MAGIC)
            f.write(json.dumps(head).encode() + b"\n")
            f.write(bytes(buf))

    @classmethod
    def load(cls, method, path, *, cache: OperatorCache = None, device: int = 0) -> "DevicePlan":
        """A plan from a file written by ``save``, with a cache of the same method (the
        lists and buffers are rebuilt; the tree is uploaded)."""
        with open(path, "rb") as f:
            if f.readline() != cls._FILE_MAGIC:
                raise ArgumentError("not a plan file")
            head = json.loads(f.readline())
            data = f.read()
        cache = OperatorCache(method) if cache is None else cache
        if cache.method is not method:
            raise ArgumentError("the cache was built for a different method")
        self = cls.__new__(cls)
        self.method, self.cache, self.device = method, cache, device
        self.precision, self.m2l = head["precision"], head["m2l"]
        self.origin, self.W0, self.e0 = np.asarray(head["origin"], dtype=np.float64), float(head["W0"]), int(head["e0"])
        self.N, self.ppl = int(head["N"]), int(head["ppl"])
        w_ref = math.ldexp(self.W0, -self.e0)   # the cache's reference width when the plan was made
        if not cache.has_base_width():
            cache.W0 = w_ref
        elif cache.W0 != w_ref:
            raise ArgumentError("the plan data was made with another reference width")
        oc = cache.device(device)
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        h = ctypes.c_void_p()
        call("ffmm_plan_load", oc, buf, len(data), ctypes.byref(h))
        self._h = h
        self.group_sizes = [int(g) for g in head["group_sizes"]]
        self.group_offsets = np.concatenate([[0], np.cumsum(self.group_sizes)]).astype(np.int64)
        self._gperm = None
        self._read_sizes()
        if _is_harmonic(method) and isinstance(method.kernel, Yukawa3D):
            cache.ensure_widths(device, [math.ldexp(self.W0, -l) for l in range(2, self.nlevels)])
        self.tol = 0.0
        return self

    @classmethod
    def from_tree(cls, method, tree, *, cache: OperatorCache = None, ppl: int = 50, precision: str = "f64",
                  m2l: str = "dense", device: int = 0, tol: float = 0.0, fast: bool = False,
                  eltype=None) -> "DevicePlan":
        """``DevicePlan(f::FMM, to)`` (src/device_passes.jl:90-96): the plan from a tree
        built on the host -- any object with the fields of ``TreeArrays`` (``origin``,
        ``W0``, ``level_start``, ``cell``, ``parent``, ``child_first``, ``plo``, ``phi``,
        ``leaf``, ``ghost``, ``xlist``, ``colleagues``, ``perm`` and the tree-order
        ``points``), 0-based.  The tree is checked in full on the host before its upload
        (``ffmm_plan_from_tree``); the lists and buffers are built on the device.  The
        root box must be a width of the cache (``w_ref * 2^e``), which it is when the
        cache fitted it (``fit_box``) or holds no width yet."""
        _check_plan_kwargs(method, m2l, tol, fast, eltype)
        cache = OperatorCache(method) if cache is None else cache
        if cache.method is not method:
            raise ArgumentError("the cache was built for a different method")
        if precision not in _PREC:
            raise ArgumentError(f"unknown precision {precision!r}")
        if m2l not in _M2L:
            raise ArgumentError(f"unknown M2L form {m2l!r}")
        origin = np.ascontiguousarray(tree.origin, dtype=np.float64).reshape(3)
        W = float(tree.W0)
        if not cache.has_base_width():
            cache.W0 = W
        f, e0 = math.frexp(W / cache.W0)
        if f != 0.5:
            raise ArgumentError("the tree's box width is not a width of the cache (w_ref * 2^e)")
        e0 -= 1
        c = lambda a, t: np.ascontiguousarray(a, dtype=t)   # noqa: E731
        ls = c(tree.level_start, np.int64)
        arrs = [c(tree.cell, np.int32), c(tree.parent, np.int32), c(tree.child_first, np.int32),
                c(tree.plo, np.int32), c(tree.phi, np.int32), c(tree.leaf, np.uint8), c(tree.ghost, np.uint8),
                c(tree.xlist, np.uint8), c(tree.colleagues, np.int32), c(tree.perm, np.int32),
                c(tree.points, np.float64)]
        N = len(arrs[9])
        nn = int(ls[-1])
        if arrs[0].size != 3 * nn or arrs[8].size != 27 * nn or arrs[10].size != 3 * N or \
                any(len(a) != nn for a in arrs[1:8]):
            raise DimensionMismatch("the tree's arrays differ in length")
        oc = cache.device(device)
        h = ctypes.c_void_p()
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_plan_from_tree", oc, len(ls) - 1, P(ls), N, *[P(a) for a in arrs], int(ppl),
             (ctypes.c_double * 3)(*origin), W, e0, _PREC[precision], _M2L[m2l], ctypes.byref(h))
        self = cls.__new__(cls)
        self._h = h
        self.method, self.cache, self.device = method, cache, device
        self.precision, self.m2l = precision, m2l
        self.origin, self.W0, self.e0 = origin.copy(), W, e0
        self.N, self.ppl = N, int(ppl)
        self.group_sizes = [N]
        self.group_offsets = np.array([0, N], dtype=np.int64)
        self._gperm = None
        self._read_sizes()
        if _is_harmonic(method) and isinstance(method.kernel, Yukawa3D):
            cache.ensure_widths(device, [math.ldexp(self.W0, -l) for l in range(2, self.nlevels)])
        self.tol = 0.0
        if tol > 0:
            self.set_tol(tol)
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().ffmm_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def _read_sizes(self):
        nl, nn, np_ = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        call("ffmm_plan_sizes", self._h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(np_), None, None, None, None,
             None)
        L = nl.value
        arrs = [np.zeros(L, dtype=np.int64) for _ in range(5)]
        P = [a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
        call("ffmm_plan_sizes", self._h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(np_), *P)
        self.nlevels, self.nnodes = L, nn.value
        self.level_start, self.level_count, self.level_pairs, self.level_s2m, self.level_leaves = arrs

    # -- downloads (tests) --
    def download_tree(self) -> TreeArrays:
        n, N = self.nnodes, self.N
        a = dict(cell=np.empty((n, 3), np.int32), level=np.empty(n, np.int32), parent=np.empty(n, np.int32),
                 child_first=np.empty(n, np.int32), plo=np.empty(n, np.int32), phi=np.empty(n, np.int32),
                 leaf=np.empty(n, np.uint8), ghost=np.empty(n, np.uint8), xlist=np.empty(n, np.uint8),
                 has_source=np.empty(n, np.uint8), colleagues=np.empty((n, 27), np.int32),
                 perm=np.empty(N, np.int32), point_node=np.empty(N, np.int32), points=np.empty((N, 3), np.float64))
        order = ["cell", "level", "parent", "child_first", "plo", "phi", "leaf", "ghost", "xlist", "has_source",
                 "colleagues", "perm", "point_node", "points"]
        call("ffmm_plan_download_tree", self._h, *[a[k].ctypes.data_as(ctypes.c_void_p) for k in order])
        ls = np.concatenate([self.level_start, [self.nnodes]]).astype(np.int32)
        return TreeArrays(origin=self.origin.copy(), W0=self.W0, level_start=ls, **a)

    def download_lists(self, level: int):
        nl = int(self.level_count[level])
        npairs = int(self.level_pairs[level])
        ptr, src, far = np.empty(nl + 1, np.int32), np.empty(npairs, np.int32), np.empty(npairs, np.int32)
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_plan_download_lists", self._h, level, P(ptr), P(src), P(far))
        return ptr, src, far

    def download_expansions(self):
        if _is_harmonic(self.method):
            Ns = Nl = self.method.expansion_length
        else:
            o = self.cache.host_operators()
            Ns, Nl = o.nsmall, o.nlarge
        nleaves = int(self.level_leaves.sum())
        M = np.empty((self.nnodes, Ns))
        Pl = np.empty((self.nnodes, Ns))
        Lc = np.empty((nleaves, Nl))
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_plan_download_expansions", self._h, P(M), P(Pl), P(Lc))
        return M.T, Pl.T, Lc.T

    def spectral_tiles(self):
        """Per level, the tiles of the mixed spectral M2L's Hadamard products run by the
        kernel staging each tile's distinct colleague boxes, and by the per-group kernel."""
        a, b = np.zeros(self.nlevels, dtype=np.int64), np.zeros(self.nlevels, dtype=np.int64)
        call("ffmm_plan_spectral_tiles", self._h, a.ctypes.data_as(ctypes.c_void_p),
             b.ctypes.data_as(ctypes.c_void_p), self.nlevels)
        return a, b

    def stage_times(self) -> dict:
        ms = (ctypes.c_double * 9)()
        call("ffmm_plan_stage_times", self._h, ms, 9)
        names = ["tree", "lists", "upward", "downward", "m2l", "eval_sources", "eval_targets", "gather", "hadamard"]
        return dict(zip(names, ms[:]))

    @property
    def ngroups(self) -> int:
        return len(self.group_sizes)

    def _group_positions(self):
        """Per group, the tree-order positions of its points, in tree order (the
        reference's per-group gperm, src/tree/device_builder.jl:86-116)."""
        if self._gperm is None:
            torch = _torch()
            perm = torch.empty(self.N, dtype=torch.int32, device=f"cuda:{self.device}")
            call("ffmm_plan_perm", self._h, _ptr(perm))
            perm = perm.to(torch.int64)
            self._gperm = [torch.nonzero((perm >= a) & (perm < b)).flatten()
                           for a, b in zip(self.group_offsets[:-1], self.group_offsets[1:])]
        return self._gperm

    def _join_densities(self, tau, sorted: bool):
        """One density array over all points from the groups' (``None``: targets only,
        zero densities); a Stokes group of stokeslets alone (nd = 3) joins others with
        stresslets (nd = 9) with a zero stresslet part, and a group of charges (nd = 1)
        joins groups of charges and dipoles (nd = 4) with a zero dipole."""
        torch = _torch()
        G = self.ngroups
        if not isinstance(tau, (tuple, list)):
            if G != 1:
                raise DimensionMismatch(f"expected {G} groups of sources and densities")
            return _dev_f64(tau, self.device)
        if len(tau) != G:
            raise DimensionMismatch(f"expected {G} groups of sources and densities")
        parts = [None if t is None else _dev_f64(t, self.device) for t in tau]
        nds = {1 if t.ndim == 1 else int(t.shape[1]) for t in parts if t is not None}
        if not nds:
            nds = {3 if isinstance(self.method, StokesHarmonics) else 1}
        nd = 9 if nds == {3, 9} else max(nds)
        if len(nds) > 1 and nds not in ({3, 9}, {1, 4}):
            raise ArgumentError("the groups' densities do not combine")
        for g, t in enumerate(parts):
            if t is not None and t.shape[0] != self.group_sizes[g]:
                raise DimensionMismatch(f"expected {self.group_sizes[g]} entries in group {g + 1}")
        shape = (self.N,) if nd == 1 else (self.N, nd)
        out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        pos = self._group_positions() if sorted else None
        for g, t in enumerate(parts):
            if t is None:
                continue
            if t.ndim == 2 and t.shape[1] == 3 and nd == 9:
                t = torch.cat([t, torch.zeros((t.shape[0], 6), dtype=t.dtype, device=t.device)], 1)
            if t.ndim == 1 and nd == 4:
                t = torch.cat([t[:, None], torch.zeros((t.shape[0], 3), dtype=t.dtype, device=t.device)], 1)
            if sorted:
                out[pos[g]] = t
            else:
                a, b = self.group_offsets[g], self.group_offsets[g + 1]
                out[a:b] = t
        torch.cuda.synchronize(self.device)
        return out

    def _split(self, out, sorted: bool = False):
        """The per-group outputs (original order within each group)."""
        return tuple(out[a:b] for a, b in zip(self.group_offsets[:-1], self.group_offsets[1:]))

    # -- the passes (device tensors in and out) --
    def build_expansions(self, tau, sorted: bool = False):
        """Densities: N charges; for SphericalHarmonics(Laplace3D()) also (N, 4) rows
        (q, d) of a charge q and a dipole d (``Dipole{3}``, src/sources.jl:17-29; a
        dipole alone has q = 0); for Stokes (N, 3) stokeslets or (N, 9) rows (f, g, nu)
        of a stokeslet f plus a stresslet (g, nu); with groups, a tuple of them, one per
        group (``None`` for a group of targets only)."""
        t = self._join_densities(tau, sorted and self.ngroups > 1)
        if t.shape[0] != self.N or t.ndim > 2:
            raise DimensionMismatch(f"expected {self.N} entries in group 1")
        nd = 1 if t.ndim == 1 else t.shape[1]
        call("ffmm_build_expansions_nd", self._h, _ptr(t), nd, int(bool(sorted)))
        return self

    def _ncomp(self, o) -> int:
        if isinstance(self.method, StokesHarmonics):
            return 1 if o.code == _lib.PRESSURE else 3
        return o.ncomp

    def evaluate_at_sources(self, output=Potential, out=None):
        o = _output(output)
        torch = _torch()
        nc = self._ncomp(o)
        shape = (self.N,) if nc == 1 else (self.N, nc)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        elif tuple(out.shape) != shape:
            raise DimensionMismatch(f"expected {self.N} outputs for group 1")
        call("ffmm_evaluate_at_sources", self._h, _ptr(out), o.code)
        return out

    def evaluate(self, targets, output=Potential, out=None, nodes=None):
        """The output at ``targets``, located on the device, or in the real leaves
        ``nodes`` (0-based node indices, e.g. from the host ``locate``) when given:
        ``evaluate!(out, dev, ds, targets, nodes)`` (src/device.jl:127-145)."""
        o = _output(output)
        torch = _torch()
        t = _dev_f64(targets, self.device, cols=3)
        nt = t.shape[0]
        if nodes is not None:
            nodes = torch.as_tensor(np.asarray(nodes) if not torch.is_tensor(nodes) else nodes).to(
                device=f"cuda:{self.device}", dtype=torch.int32).contiguous()
            if nodes.shape != (nt,):
                raise DimensionMismatch("out, targets and nodes differ in length")
        nc = self._ncomp(o)
        shape = (nt,) if nc == 1 else (nt, nc)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        elif tuple(out.shape) != shape:
            raise DimensionMismatch("out and targets differ in length")
        if nodes is None:
            call("ffmm_evaluate", self._h, _ptr(t), nt, _ptr(out), o.code)
        else:
            call("ffmm_evaluate_nodes", self._h, _ptr(t), nt, _ptr(nodes), _ptr(out), o.code)
        return out

    def capture(self, output=Potential, nd: int = 1):
        """Capture the pass sequence of ``nd`` densities per point and ``output`` at the
        sources as one HIP graph (``ffmm_plan_capture``), for solves repeated on these
        points; ``replay`` then runs one."""
        if self.ngroups != 1:
            raise ArgumentError("a plan of several groups is not captured")
        o = _output(output)
        call("ffmm_plan_capture", self._h, int(nd), o.code)
        self._captured = (int(nd), o)
        return self

    def replay(self, tau, out=None):
        """One captured solve (``ffmm_plan_replay``): the densities in input order, the
        output at the sources as ``matvec`` returns it."""
        cap = getattr(self, "_captured", None)
        if cap is None:
            raise ArgumentError("capture the plan first")
        nd, o = cap
        t = _dev_f64(tau, self.device)
        if t.shape[0] != self.N or (t.ndim == 1 and nd != 1) or (t.ndim == 2 and t.shape[1] != nd):
            raise DimensionMismatch(f"expected {self.N} x {nd} densities")
        torch = _torch()
        nc = self._ncomp(o)
        shape = (self.N,) if nc == 1 else (self.N, nc)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        elif tuple(out.shape) != shape:
            raise DimensionMismatch(f"expected an output of shape {shape}")
        call("ffmm_plan_replay", self._h, _ptr(t), _ptr(out))
        return out

    def matvec(self, tau, output=Potential, sorted: bool = False, out=None):
        """``device_matvec!``; with groups, ``output`` may be a tuple of one output type
        per group (``None``: not wanted) and the result is a tuple of per-group outputs
        (src/device_passes.jl:619-640)."""
        if self.ngroups == 1:
            if isinstance(output, tuple):
                raise ArgumentError("several outputs for one group")
            self.build_expansions(tau, sorted)
            return self.evaluate_at_sources(output, out)
        if out is not None:
            raise ArgumentError("several groups: outputs are returned per group")
        outputs = output if isinstance(output, tuple) else (output,) * self.ngroups
        if len(outputs) != self.ngroups:
            raise DimensionMismatch(f"expected {self.ngroups} outputs")
        self.build_expansions(tau, sorted)
        res, cache = [], {}
        for g, o in enumerate(outputs):
            if o is None:
                res.append(None)
                continue
            if o not in cache:
                cache[o] = self._split(self.evaluate_at_sources(o))
            v = cache[o][g]
            if sorted:
                a = self.group_offsets[g]
                torch = _torch()
                perm = torch.empty(self.N, dtype=torch.int32, device=v.device)
                call("ffmm_plan_perm", self._h, _ptr(perm))
                # the group's outputs in its tree order
                v = v[(perm[self._group_positions()[g]].to(torch.int64) - a)]
            res.append(v)
        return tuple(res)


def device_build_expansions_(plan: DevicePlan, ds, sorted: bool = False) -> DevicePlan:
    """``device_build_expansions!(plan, ds; sorted)`` (src/device_passes.jl:511)."""
    return plan.build_expansions(ds["tau"] if isinstance(ds, dict) else ds[1], sorted)


def device_matvec_(out, plan: DevicePlan, ds, output=Potential, sorted: bool = False):
    """``device_matvec!(out, plan, ds; output, sorted)`` (src/device_passes.jl:619):
    ``out`` (a device tensor) receives the output at every source in the original order."""
    tau = ds["tau"] if isinstance(ds, dict) else ds[1]
    return plan.matvec(tau, output, sorted, out)


# ---------------------------------------------------------------------------
# the FMM interface of the host driver (src/fmm.jl, src/evaluate.jl)
# ---------------------------------------------------------------------------
class FMM:
    """``FMM(method, sources; ppl=50, bbox=None, cache=OperatorCache(method), max_levels=30)``
    (src/fmm.jl:24-49): here the tree, its lists and the operators live on the GPU (a
    ``DevicePlan``)."""

    def __init__(self, method, sources, *, ppl: int = 50, bbox=None, cache=None,
                 max_levels: int = 30, precision: str = "f64", m2l: str = "dense", device: int = 0):
        if isinstance(sources, tuple) and len(sources) == 1:
            sources = sources[0]
        self.method = method
        from . import circle2d
        if isinstance(method, circle2d.KernelIndependent):
            # the 2D circle method (src/operators/kifmm.jl): FP64, its own plan
            if precision != "f64" or m2l != "dense":
                raise ArgumentError("the circle method runs in FP64 with its dense M2L")
            self.plan = circle2d.CirclePlan(method, sources, ppl=ppl, bbox=bbox, cache=cache, max_levels=max_levels,
                                            device=device)
        else:
            self.plan = DevicePlan(method, sources, ppl=ppl, bbox=bbox, cache=cache, max_levels=max_levels,
                                   precision=precision, m2l=m2l, device=device)
        self.cache = self.plan.cache

    @property
    def nlevels(self):
        return self.plan.nlevels

    def __repr__(self):
        d = 2 if hasattr(self.plan, "download_tree") and type(self.plan).__name__ == "CirclePlan" else 3
        return (f"FMM({type(self.method).__name__}, Tree{{{d},Float64}}: {self.plan.N} points, {self.plan.nnodes} nodes "
                f"on {self.plan.nlevels} levels)")


class FMMState:
    """The expansions of one right-hand side (src/fmm.jl:58-94); they live in the plan."""

    def __init__(self, f: FMM):
        self.fmm = f
        self.tau = None


def build_expansions_(state: FMMState, f: FMM, sources, tau) -> FMMState:
    """``build_expansions!(state, fmm, sources, tau)`` (src/fmm.jl:264-298)."""
    if state.fmm is not f:
        raise ArgumentError("the state belongs to another FMM")
    G = getattr(f.plan, "ngroups", 1)
    if G > 1:
        # one tuple entry per group (src/fmm.jl:270-281)
        if not isinstance(tau, (tuple, list)) or len(tau) != G or (
                sources is not None and (not isinstance(sources, (tuple, list)) or len(sources) != G)):
            raise DimensionMismatch(f"expected {G} groups of sources and densities")
        for g in range(G):
            n = f.plan.group_sizes[g]
            if sources is not None and len(sources[g]) != n:
                raise DimensionMismatch(f"expected {n} sources in group {g + 1}")
            if tau[g] is not None and len(tau[g]) != n:
                raise DimensionMismatch(f"expected {n} densities in group {g + 1}")
        f.plan.build_expansions(tuple(tau))
        state.tau = tau
        return state
    n = f.plan.N
    if isinstance(sources, tuple) and len(sources) == 1:
        sources = sources[0]
    if isinstance(tau, tuple) and len(tau) == 1:
        tau = tau[0]
    if sources is not None and len(sources) != n:
        raise DimensionMismatch(f"expected {n} sources in group 1")
    tau = np.asarray(tau) if not hasattr(tau, "data_ptr") else tau
    if len(tau) != n:
        raise DimensionMismatch(f"expected {n} densities in group 1")
    f.plan.build_expansions(tau)
    state.tau = tau
    return state


def build_expansions(f: FMM, sources, tau) -> FMMState:
    """``build_expansions(fmm, sources, tau)`` (src/fmm.jl:300-314)."""
    return build_expansions_(FMMState(f), f, sources, tau)


def evaluate_at_sources(f: FMM, state: FMMState, g: int = 1, output=Potential) -> np.ndarray:
    """``evaluate_at_sources(fmm, state, g=1; output)`` (src/evaluate.jl:104-132): the
    output at every source of group g (1-based) in the original order."""
    G = getattr(f.plan, "ngroups", 1)
    if not 1 <= g <= G:
        raise ArgumentError(f"group {g} of {G}")
    if state.fmm is not f:
        raise ArgumentError("the state belongs to another FMM")
    if not hasattr(f.plan, "ngroups"):          # the 2D circle method: the potential
        if not isinstance(_output(output), Potential):
            raise ArgumentError("the circle method evaluates the potential")
        return f.plan.evaluate_at_sources().cpu().numpy()
    out = f.plan.evaluate_at_sources(output)
    a, b = f.plan.group_offsets[g - 1], f.plan.group_offsets[g]
    return out[a:b].cpu().numpy()


def evaluate(f: FMM, state: FMMState, targets, output=Potential) -> np.ndarray:
    """``evaluate(fmm, state, targets; output)`` (src/evaluate.jl:81-102)."""
    if state.fmm is not f:
        raise ArgumentError("the state belongs to another FMM")
    if not hasattr(f.plan, "ngroups"):          # the 2D circle method: the potential
        if not isinstance(_output(output), Potential):
            raise ArgumentError("the circle method evaluates the potential")
        return f.plan.evaluate(targets).cpu().numpy()
    return f.plan.evaluate(targets, output).cpu().numpy()
