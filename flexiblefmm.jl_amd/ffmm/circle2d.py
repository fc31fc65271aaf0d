"""The 2D kernel-independent method on circles (``KernelIndependent(Laplace2D();
nequiv)``, src/operators/kifmm.jl:1-87): host operators (FP64) and the host API over
the device plan of ``libffmm_hip`` (``ffmm_circle_*``).

A box of width w carries two concentric circles of ``nequiv`` equispaced points, the
small one (radius w sqrt(2)/2: upward equivalent, downward check) and the large one
(radius w (4 - sqrt(2))/2: upward check, downward equivalent).  The check-to-equivalent
maps are truncated-SVD pseudo-inverses kept as factors, M2M and L2L have them folded
in, and M2L is one dense matrix per far offset (the 40 offsets of (-3:3)^2 outside
(-1:1)^2).  The kernel c log r^2 is not homogeneous, so every level width has its own
operators (src/device_passes.jl:242-255: the dense per-offset branch, per level).

The quadtree and its interaction lists are built by the library's host code (the
reference's host builder, src/tree/builder.jl, for D = 2); upward pass, M2L, downward
pass and the evaluation run on the device in FP64.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, call

KI_SMALL_RADIUS = math.sqrt(2) / 2          # src/operators/kifmm.jl:35
KI_LARGE_RADIUS = (4 - math.sqrt(2)) / 2    # src/operators/kifmm.jl:36


class Laplace2D:
    """``Laplace2D()`` (src/kernels.jl:54-63): ``c log |t - s|^2`` with ``c = -1/(4 pi)``."""
    c = -1.0 / (4.0 * math.pi)

    def __repr__(self):
        return "Laplace2D()"


class KernelIndependent:
    """``KernelIndependent(kernel; nequiv=50, pinv_rtol=50 eps)`` (src/operators/kifmm.jl:17-33)."""

    def __init__(self, kernel, nequiv: int = 50, pinv_rtol: float = 50 * np.finfo(np.float64).eps):
        if not isinstance(kernel, Laplace2D):
            raise ArgumentError("the device library takes Laplace2D for the circle method")
        if nequiv < 4:
            raise ArgumentError("nequiv must be at least 4")
        if nequiv > 64:
            raise ArgumentError("the device library takes nequiv <= 64")
        self.kernel, self.nequiv, self.pinv_rtol = kernel, int(nequiv), float(pinv_rtol)

    def __repr__(self):
        return f"KernelIndependent({self.kernel!r}; nequiv={self.nequiv})"


# ---------------------------------------------------------------------------
# enumerations (src/util.jl), D = 2
# ---------------------------------------------------------------------------
def child_offsets() -> np.ndarray:
    """Slot k has bit pattern k with coordinate 0 least significant (src/util.jl:19-29)."""
    return np.array([[k & 1, (k >> 1) & 1] for k in range(4)], dtype=np.int64)


def far_offsets() -> np.ndarray:
    """The 7^2 - 3^2 = 40 offsets of (-3:3)^2 with Chebyshev norm >= 2, coordinate 0
    fastest (src/util.jl:59-84)."""
    out = [(i % 7 - 3, i // 7 - 3) for i in range(49)]
    return np.array([o for o in out if max(abs(o[0]), abs(o[1])) >= 2], dtype=np.int64)


def circle_points(n: int, r: float) -> np.ndarray:
    """``circle_points(n, r)`` (src/operators/kifmm.jl:68-70): r (cos t, sin t), t = 2 pi k / n."""
    t = (2 * math.pi / n) * np.arange(n)
    return np.stack([r * np.cos(t), r * np.sin(t)], axis=1)


def laplace2d_matrix(sources: np.ndarray, targets: np.ndarray) -> np.ndarray:
    """``kernel_matrix(K, sources, targets)``: entry (i, j) = c log |t_i - s_j|^2."""
    d = targets[:, None, :] - sources[None, :, :]
    return Laplace2D.c * np.log(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])


class CircleOperators:
    """The operators of one box width: surfaces relative to the box centre, the factored
    pseudo-inverses (Uh: r x n, Vs = V diag(invS): n x r), M2M (n x 4n), L2L (4n x n) and
    the 40 M2L matrices (n x n: rows the target's small circle, columns the source's)."""

    def __init__(self, nequiv: int, w: float, rtol: float):
        from .operators import pinv_rtol
        n = nequiv
        self.width, self.n = float(w), n
        self.small = circle_points(n, KI_SMALL_RADIUS * w)
        self.large = circle_points(n, KI_LARGE_RADIUS * w)
        child_small = circle_points(n, KI_SMALL_RADIUS * w / 2)
        # _ki_dense_operators (src/operators/ki_common.jl:176-199)
        self.P_up = pinv_rtol(laplace2d_matrix(self.small, self.large), rtol)
        self.P_down = pinv_rtol(laplace2d_matrix(self.large, self.small), rtol)
        centers = [(o - 0.5) * (w / 2) for o in child_offsets().astype(np.float64)]
        M2Mc = np.hstack([laplace2d_matrix(child_small + c, self.large) for c in centers])
        L2Lc = np.vstack([laplace2d_matrix(self.large, child_small + c) for c in centers])
        P, Q = self.P_up, self.P_down
        self.M2M = P.V @ (P.invS[:, None] * (P.Uh @ M2Mc))
        self.L2L = ((L2Lc @ Q.V) * Q.invS[None, :]) @ Q.Uh
        # _multipole_to_check_matrix(K, small + o w, o w, w, small, 0) per far offset
        # (src/operators/kifmm.jl:83-84)
        self.M2L = np.array([laplace2d_matrix(self.small + o * w, self.small)
                             for o in far_offsets().astype(np.float64)])


class CircleOperatorCache:
    """Operators by box width (``OperatorCache``, src/operators/cache.jl:5-46: a
    non-homogeneous kernel keeps one set per width)."""

    def __init__(self, method: KernelIndependent):
        self.method = method
        self._ops = {}

    def operators(self, w: float) -> CircleOperators:
        ops = self._ops.get(w)
        if ops is None:
            ops = CircleOperators(self.method.nequiv, w, self.method.pinv_rtol)
            self._ops[w] = ops
        return ops


# ---------------------------------------------------------------------------
# the plan over the C ABI
# ---------------------------------------------------------------------------
def _torch():
    from .fmm import _torch as t
    return t()


def _f64(a, cols):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if a.ndim != 2 or a.shape[1] != cols:
        raise DimensionMismatch(f"expected an N × {cols} array, got {a.shape}")
    return a


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class CircleTree:
    """The plan's quadtree (``Tree`` fields, 0-based, levels concatenated)."""


class CirclePlan:
    """The device plan of a 2D circle FMM: the host-built quadtree and its lists on the
    device, the operators of every level width, the work buffers
    (``DevicePlan(f::FMM)``, src/device_passes.jl:90-96, for D = 2)."""

    def __init__(self, method: KernelIndependent, sources, *, ppl: int = 50, bbox=None,
                 cache: CircleOperatorCache = None, max_levels: int = 30, device: int = 0):
        from .fmm import context
        if not isinstance(method, KernelIndependent):
            raise ArgumentError("CirclePlan takes KernelIndependent(Laplace2D())")
        cache = CircleOperatorCache(method) if cache is None else cache
        if cache.method is not method:
            raise ArgumentError("the cache was built for a different method")
        xy = _f64(sources, 2)
        n = len(xy)
        if n == 0:
            raise ArgumentError("no points")
        if not np.all(np.isfinite(xy)):
            raise ArgumentError("the points must be finite")
        if bbox is None:
            lo, hi = xy.min(axis=0), xy.max(axis=0)
            W = float(np.max(hi - lo)) or 1.0
            c = lo + (hi - lo) / 2
            origin = c - W / 2
            while not (np.all(origin <= lo) and np.all(origin + W >= hi)):
                W *= 1 + 4 * np.finfo(np.float64).eps
                origin = c - W / 2
        else:
            origin, W = np.asarray(bbox[0], dtype=np.float64).reshape(2), float(bbox[1])
        if not W > 0:
            raise ArgumentError("box width must be positive")
        self.method, self.cache, self.device = method, cache, device
        self.origin, self.W0, self.N, self.ppl = origin.copy(), W, n, int(ppl)
        h = ctypes.c_void_p()
        call("ffmm_circle_create", context(device), method.nequiv, xy.ctypes.data_as(ctypes.c_void_p), n,
             int(ppl), (ctypes.c_double * 2)(*origin), W, int(max_levels), ctypes.byref(h))
        self._h = h
        L = ctypes.c_int()
        nn = ctypes.c_int64()
        call("ffmm_circle_sizes", h, ctypes.byref(L), ctypes.byref(nn), None, None)
        ls = np.zeros(L.value + 1, dtype=np.int64)
        lp = np.zeros(L.value, dtype=np.int64)
        call("ffmm_circle_sizes", h, ctypes.byref(L), ctypes.byref(nn), ls.ctypes.data_as(ctypes.c_void_p),
             lp.ctypes.data_as(ctypes.c_void_p))
        self.nlevels, self.nnodes, self.level_start, self.level_pairs = L.value, nn.value, ls, lp
        F = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731 (column major)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)        # noqa: E731
        for l in range(2, self.nlevels):
            o = cache.operators(math.ldexp(W, -l))
            keep = [np.ascontiguousarray(o.large.T), F(o.P_up.Uh), F(o.P_up.V * o.P_up.invS[None, :]),
                    F(o.P_down.Uh), F(o.P_down.V * o.P_down.invS[None, :]), F(o.M2M), F(o.L2L),
                    np.ascontiguousarray(np.stack([F(m).ravel(order="F") for m in o.M2L]))]
            call("ffmm_circle_put_level", h, l, P(keep[0]), o.P_up.rank, P(keep[1]), P(keep[2]), o.P_down.rank,
                 P(keep[3]), P(keep[4]), P(keep[5]), P(keep[6]), P(keep[7]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                call("ffmm_circle_destroy", h)
            except Exception:
                pass
            self._h = None

    def build_expansions(self, tau):
        """``device_build_expansions!`` (src/device_passes.jl:511-597): charges in the
        input order."""
        torch = _torch()
        t = torch.as_tensor(np.ascontiguousarray(tau, dtype=np.float64) if not isinstance(tau, torch.Tensor) else tau)
        t = t.to(device=f"cuda:{self.device}", dtype=torch.float64).contiguous()
        if t.ndim != 1 or t.shape[0] != self.N:
            raise DimensionMismatch(f"expected {self.N} charges")
        torch.cuda.synchronize(self.device)
        call("ffmm_circle_build_expansions", self._h, _ptr(t))
        return self

    def evaluate_at_sources(self):
        """``evaluate_at_sources!``: the potential at every source, the source itself
        excluded, in the input order."""
        torch = _torch()
        out = torch.empty(self.N, dtype=torch.float64, device=f"cuda:{self.device}")
        call("ffmm_circle_evaluate_at_sources", self._h, _ptr(out))
        return out

    def evaluate(self, targets):
        """``evaluate!`` at targets inside the root box (located on the device)."""
        torch = _torch()
        if isinstance(targets, torch.Tensor):
            x = targets.to(device=f"cuda:{self.device}", dtype=torch.float64).contiguous()
            if x.ndim != 2 or x.shape[1] != 2:
                raise DimensionMismatch(f"expected an N × 2 array, got {tuple(x.shape)}")
        else:
            x = torch.as_tensor(_f64(targets, 2)).to(f"cuda:{self.device}")
        out = torch.empty(x.shape[0], dtype=torch.float64, device=f"cuda:{self.device}")
        torch.cuda.synchronize(self.device)
        call("ffmm_circle_evaluate", self._h, _ptr(x), x.shape[0], _ptr(out))
        return out

    def matvec(self, tau):
        self.build_expansions(tau)
        return self.evaluate_at_sources()

    def download_tree(self) -> CircleTree:
        nn, N = self.nnodes, self.N
        t = CircleTree()
        t.level_start = self.level_start.copy()
        t.cell = np.zeros((nn, 2), np.int32)
        t.parent = np.zeros(nn, np.int32)
        t.child_first = np.zeros(nn, np.int32)
        t.plo = np.zeros(nn, np.int32)
        t.phi = np.zeros(nn, np.int32)
        t.leaf = np.zeros(nn, np.uint8)
        t.ghost = np.zeros(nn, np.uint8)
        t.xlist = np.zeros(nn, np.uint8)
        t.colleagues = np.zeros((nn, 9), np.int32)
        t.perm = np.zeros(N, np.int32)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        call("ffmm_circle_download_tree", self._h, P(t.cell), P(t.parent), P(t.child_first), P(t.plo), P(t.phi),
             P(t.leaf), P(t.ghost), P(t.xlist), P(t.colleagues), P(t.perm))
        t.lists = []
        for l in range(self.nlevels):
            nl = int(self.level_start[l + 1] - self.level_start[l])
            ptr = np.zeros(nl + 1, np.int32)
            src = np.zeros(int(self.level_pairs[l]), np.int32)
            far = np.zeros(int(self.level_pairs[l]), np.int32)
            call("ffmm_circle_download_lists", self._h, l, P(ptr), P(src), P(far))
            t.lists.append((ptr, src, far))
        return t


def _circle_lib_check():
    return _lib.lib()
