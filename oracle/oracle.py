"""ORACLE -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
import this module, and only as the checker or as the timed CPU baseline; the
product library never loads it.

A CPU restatement of the reference's host FMM for Laplace3D on cube
kernel-independent surfaces:

* the ghost-cell tree, its interaction lists and ``locate``: oracle/tree_oracle.c
  (src/tree/builder.jl, src/tree/lists.jl, src/tree/locate.jl);
* the FP64 direct sums and the per-point S2M / L2T / near-field loops:
  oracle/fmm_oracle.c (src/kernels.jl:442-487, src/operators/ki_common.jl:256-315,
  src/evaluate.jl:12-58);
* the upward and downward passes (``upward_pass!``/``downward_pass!``,
  src/fmm.jl:170-262) with numpy/BLAS for the GEMM stages, M2L either per far
  offset through the symmetry-class representatives (``m2l_level!`` default,
  src/fmm.jl:100-127, the form the reference device plan uses) or spectral
  (``_spectral_m2l_level!``, src/operators/spectral_m2l.jl:328-451, the form the
  reference CPU driver uses for the cube method);
* ``evaluate_at_sources!`` / ``evaluate!`` (src/evaluate.jl:81-127).

The operators (surfaces, pseudo-inverse factors, folded M2M/L2L, class
representatives, spectral tables) are supplied by the caller from the host
precompute ``ffmm.operators``; they are pinned independently by
tests/test_operators.py (field error on all 316 offsets, M2M/L2L against the exact
fields, class reproduction, spectral against dense).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    """The built oracle library (``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.otree_build.restype = ctypes.c_int
        L.otree_build.argtypes = [f64p, ctypes.c_int64, ctypes.c_int, f64p, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_void_p)]
        L.otree_last_error.restype = ctypes.c_char_p
        L.otree_nlevels.argtypes = [ctypes.c_void_p]
        L.otree_nnodes.argtypes = [ctypes.c_void_p]
        L.otree_level_npairs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.otree_level_npairs.restype = ctypes.c_int64
        L.otree_get.argtypes = [ctypes.c_void_p, i32p, i32p, i32p, i32p, i32p, i32p, i32p, u8p, u8p, u8p, u8p,
                                i32p, i32p, i32p, f64p]
        L.otree_get_lists.argtypes = [ctypes.c_void_p, ctypes.c_int, i32p, i32p, i32p]
        L.otree_locate.argtypes = [ctypes.c_void_p, f64p, ctypes.c_int64, i32p]
        L.otree_free.argtypes = [ctypes.c_void_p]
        L.odirect.argtypes = [f64p, f64p, ctypes.c_int64, f64p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, f64p]
        L.os2m.argtypes = [i32p, ctypes.c_int64, f64p, i32p, f64p, f64p, ctypes.c_int, i32p, i32p, f64p, f64p, f64p]
        L.oeval.argtypes = [f64p, i32p, i32p, i64p, ctypes.c_int64, f64p, i64p, f64p, f64p, ctypes.c_int,
                            i32p, f64p, i32p, i32p, i32p, f64p, f64p, ctypes.c_int, f64p]
        i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
        L.om2l_gather.argtypes = [f64p, ctypes.c_int, i64p, i64p, ctypes.c_int64, i32p, i8p, f64p]
        L.om2l_scatter.argtypes = [f64p, ctypes.c_int, i64p, i64p, i64p, ctypes.c_int64, i32p, i8p, f64p,
                                   ctypes.c_double]
        L.ohadamard.argtypes = [i32p, i32p, i32p, i64p, ctypes.c_int64, f64p, f64p, ctypes.c_int64, f64p]
        L.othreads.restype = ctypes.c_int
        tabs = [f64p] * 6
        L.oh_s2m.argtypes = [ctypes.c_int, ctypes.c_int, *tabs, i32p, ctypes.c_int64, f64p, i32p, f64p, f64p, i32p,
                             i32p, f64p, f64p, ctypes.c_int, f64p]
        L.oh_l2t.argtypes = [ctypes.c_int, ctypes.c_int, *tabs, f64p, i32p, ctypes.c_int64, f64p, f64p, i32p, f64p,
                             f64p, ctypes.c_int, f64p]
        L.oh_near.argtypes = [ctypes.c_int, ctypes.c_double, f64p, i32p, i32p, ctypes.c_int64, i32p, i32p, i32p,
                              f64p, f64p, ctypes.c_int, ctypes.c_int, f64p]
        L.oh_direct.argtypes = [ctypes.c_int, ctypes.c_double, f64p, f64p, ctypes.c_int, ctypes.c_int64, f64p,
                                ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, f64p]
        _LIB = L
    return _LIB


class OracleError(ValueError):
    pass


@dataclass
class Tree:
    """The flat tree (0-based; -1 = none), the fields of the reference's ``Tree``
    (src/tree/tree.jl:40-72) with ``points`` as ``plo``/``phi``."""
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
    colleagues: np.ndarray     # (n, 27)
    perm: np.ndarray
    point_node: np.ndarray
    points: np.ndarray         # sorted points (N, 3)
    lists: list                # per level: (ptr, src, far), targets in node order

    @property
    def nlevels(self) -> int:
        return len(self.level_start) - 1

    @property
    def nnodes(self) -> int:
        return len(self.level)

    def level_range(self, l: int) -> range:
        return range(int(self.level_start[l]), int(self.level_start[l + 1]))

    def group_parent(self, l: int) -> np.ndarray:
        if l == 0:
            return np.zeros(0, dtype=np.int32)
        r = self.level_range(l)
        return self.parent[r.start:r.stop:8].copy()

    def width(self, l: int) -> float:
        return float(np.ldexp(self.W0, -l))

    def centers(self) -> np.ndarray:
        w = np.ldexp(self.W0, -self.level.astype(np.int64))
        return self.origin[None, :] + (self.cell.astype(np.float64) + 0.5) * w[:, None]

    def locate(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        # the tree is rebuilt on demand from its arrays: keep a handle alive instead
        out = np.empty(len(x), dtype=np.int32)
        lib().otree_locate(self._handle, x, len(x), out)
        if np.any(out < 0):
            raise OracleError("point lies outside the tree's box")
        return out


def build_tree(points: np.ndarray, ppl: int, origin, W0: float, max_levels: int = 30) -> Tree:
    """``Tree(points; ppl, bbox=(origin, W0), max_levels)`` (src/tree/builder.jl:482-495)."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    o = np.ascontiguousarray(origin, dtype=np.float64)
    h = ctypes.c_void_p()
    L = lib()
    rc = L.otree_build(pts, len(pts), int(ppl), o, float(W0), int(max_levels), ctypes.byref(h))
    if rc != 0:
        raise OracleError(L.otree_last_error().decode())
    nl, n, N = L.otree_nlevels(h), L.otree_nnodes(h), len(pts)
    a = dict(level_start=np.empty(nl + 1, np.int32), cell=np.empty((n, 3), np.int32), level=np.empty(n, np.int32),
             parent=np.empty(n, np.int32), child_first=np.empty(n, np.int32), plo=np.empty(n, np.int32),
             phi=np.empty(n, np.int32), leaf=np.empty(n, np.uint8), ghost=np.empty(n, np.uint8),
             xlist=np.empty(n, np.uint8), has_source=np.empty(n, np.uint8), colleagues=np.empty((n, 27), np.int32),
             perm=np.empty(N, np.int32), point_node=np.empty(N, np.int32), points=np.empty((N, 3), np.float64))
    L.otree_get(h, a["level_start"], a["cell"], a["level"], a["parent"], a["child_first"], a["plo"], a["phi"],
                a["leaf"], a["ghost"], a["xlist"], a["has_source"], a["colleagues"], a["perm"], a["point_node"],
                a["points"])
    lists = []
    for l in range(nl):
        nlv = int(a["level_start"][l + 1] - a["level_start"][l])
        npr = L.otree_level_npairs(h, l)
        ptr, src, far = np.empty(nlv + 1, np.int32), np.empty(npr, np.int32), np.empty(npr, np.int32)
        L.otree_get_lists(h, l, ptr, src, far)
        lists.append((ptr, src, far))
    t = Tree(origin=o.copy(), W0=float(W0), lists=lists, **a)
    t._handle = h
    return t


def direct(sources, tau, targets, grad: bool = False) -> np.ndarray:
    """FP64 direct sum at separate targets (``direct``, src/kernels.jl:463-472)."""
    s = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.empty((len(t), 3) if grad else len(t))
    lib().odirect(s, tau, len(s), t, len(t), None, int(grad), out)
    return out


def direct_self(sources, tau, grad: bool = False, rows=None) -> np.ndarray:
    """FP64 direct sum at the sources, self excluded (``direct_self``,
    src/kernels.jl:479-487); ``rows`` restricts it to some of the sources."""
    s = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    rows = np.arange(len(s), dtype=np.int64) if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    t = np.ascontiguousarray(s[rows])
    out = np.empty((len(rows), 3) if grad else len(rows))
    lib().odirect(s, tau, len(s), t, len(t), rows.ctypes.data_as(ctypes.c_void_p), int(grad), out)
    return out


# ---------------------------------------------------------------------------
# the passes
# ---------------------------------------------------------------------------
class FMM:
    """``FMM(method, sources; ppl, bbox)`` with the operators ``ops`` (reference width
    ``ops.width``) shared by every level through the homogeneous rescale
    (src/fmm.jl:31-49, src/operators/cache.jl:33-46)."""

    def __init__(self, ops, points, ppl: int, origin, W0: float, max_levels: int = 30, perm=None):
        self.ops = ops
        self.tree = build_tree(points, ppl, origin, W0, max_levels)
        t = self.tree
        if perm is not None:
            # the same tree with another order of the points inside each leaf (the device
            # sorts them by their full key): sums then run in that order
            perm = np.asarray(perm, dtype=np.int32)
            for i in np.nonzero(t.leaf)[0]:
                a, b = t.plo[i], t.phi[i] + 1
                if not np.array_equal(np.sort(perm[a:b]), np.sort(t.perm[a:b])):
                    raise OracleError("perm reorders points across leaves")
            t.perm = perm.copy()
            t.points = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[perm])
        L = t.nlevels
        self.lam = np.array([t.width(l) / ops.width for l in range(L)])
        self.centers = np.ascontiguousarray(t.centers())
        s2m = (t.child_first < 0) & (t.has_source != 0)
        leaf = t.leaf != 0
        self.s2m_nodes, self.leaf_nodes = [], []
        for l in range(L):
            r = t.level_range(l)
            self.s2m_nodes.append(np.ascontiguousarray(r.start + np.nonzero(s2m[r.start:r.stop])[0], dtype=np.int32))
            self.leaf_nodes.append(np.ascontiguousarray(r.start + np.nonzero(leaf[r.start:r.stop])[0], dtype=np.int32))
        self.surf = np.ascontiguousarray(ops.large)

    # upward_pass! (src/fmm.jl:170-205); expansions are stored node major (row per node)
    def upward(self, tau_sorted: np.ndarray) -> np.ndarray:
        t, ops = self.tree, self.ops
        L = t.nlevels
        Ns, Nl = ops.nsmall, ops.nlarge
        Mp = np.zeros((t.nnodes, Ns))
        P = ops.P_up
        for l in range(L - 1, 1, -1):
            nodes = self.s2m_nodes[l]
            if len(nodes):
                pM = np.empty((len(nodes), Nl))          # row q: the check potentials of node q
                lib().os2m(nodes, len(nodes), self.centers, t.level, self.surf, self.lam, Nl, t.plo, t.phi,
                           t.points, tau_sorted, pM)
                # partial_to_multipole!: pinv_scale V (invS .* (Uh pM))
                Mp[nodes] = self.lam[l] * (((pM @ P.Uh.T) * P.invS[None, :]) @ P.V.T)
            if l < L - 1:
                r = t.level_range(l + 1)
                # the children block, one stacked row of 8 children per sibling group
                B = Mp[r.start:r.stop].reshape(-1, 8 * Ns)
                Mp[t.group_parent(l + 1)] = B @ ops.M2M.T
        return Mp

    # downward_pass! (src/fmm.jl:207-262)
    def downward(self, Mp: np.ndarray, m2l: str = "dense"):
        t, ops = self.tree, self.ops
        L = t.nlevels
        Ns, Nl = ops.nsmall, ops.nlarge
        Pl = np.zeros((t.nnodes, Ns))
        Lc = np.zeros((t.nnodes, Nl))
        P = ops.P_down
        for l in range(2, L):
            r = t.level_range(l)
            if l > 2:
                G = Pl[t.group_parent(l)]
                pL = (G @ ops.L2L.T).reshape(-1, Ns)
            else:
                pL = np.zeros((len(r), Ns))
            alpha = 1.0 / self.lam[l]
            if m2l == "dense":
                self._m2l_dense(pL, Mp, l, alpha)
            else:
                self._m2l_spectral(pL, Mp, l, alpha)
            Pl[r.start:r.stop] = pL
            leaves = self.leaf_nodes[l]
            if len(leaves):
                # partial_to_local!: pinv_scale V (invS .* (Uh pL)) of the real leaves
                Lc[leaves] = self.lam[l] * (((Pl[leaves] @ P.Uh.T) * P.invS[None, :]) @ P.V.T)
        return Pl, Lc

    def _pairs(self, l: int):
        ptr, src, far = self.tree.lists[l]
        tgt = np.repeat(np.arange(len(ptr) - 1, dtype=np.int64), np.diff(ptr))   # level-local targets
        return tgt, src.astype(np.int64), far.astype(np.int64)

    def _m2l_dense(self, pL, Mp, l, alpha):
        """Per symmetry class: gather P_f^T m_s of the class's pairs, one GEMM with the
        class representative, scatter P_f y into the targets."""
        ops = self.ops
        sc = ops.sc
        Ns = ops.nsmall
        tgt, src, far = self._pairs(l)
        if len(tgt) == 0:
            return
        klass = sc.klass[far]
        order = np.lexsort((tgt, klass))
        tgt, src, far, klass = tgt[order], src[order], far[order], klass[order]
        bounds = np.searchsorted(klass, np.arange(len(ops.class_ops) + 1))
        idx = np.ascontiguousarray(sc.idx, dtype=np.int32)
        sgn = np.ascontiguousarray(sc.sgn, dtype=np.int8)
        for c in range(len(ops.class_ops)):
            a, b = bounds[c], bounds[c + 1]
            if a == b:
                continue
            G = np.empty((b - a, Ns))
            lib().om2l_gather(Mp, Ns, src[a:b].copy(), far[a:b].copy(), b - a, idx, sgn, G)
            Y = G @ ops.class_ops[c].T
            tc = tgt[a:b]
            runs = np.concatenate([[0], np.nonzero(np.diff(tc))[0] + 1, [b - a]]).astype(np.int64)
            lib().om2l_scatter(pL, Ns, tc.copy(), far[a:b].copy(), runs, len(runs) - 1, idx, sgn, Y, alpha)

    def _m2l_spectral(self, pL, Mp, l, alpha):
        """The level-wise spectral translation (_spectral_m2l_level!,
        src/operators/spectral_m2l.jl:328-451): transform every source box of the level's
        pairs (threaded FFTs), accumulate the Hadamard products per target in C, transform
        the targets back and extract their small-surface entries."""
        import scipy.fft as sfft
        ops = self.ops
        n = 2 * ops.p
        ptr, src, far = self.tree.lists[l]
        if len(src) == 0:
            return
        workers = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        srcs = np.unique(src)
        slot = np.full(self.tree.nnodes, -1, np.int32)
        slot[srcs] = np.arange(len(srcs), dtype=np.int32)
        # the multipoles sit in the [0, p)^3 corner of the (2p)^3 lattice: the transform
        # pads the p^3 block with zeros
        p = ops.p
        e = np.asarray(ops.embed)
        corner = (e // (n * n)) * p * p + ((e // n) % n) * p + e % n
        lat = np.zeros((len(srcs), p ** 3))
        lat[:, corner] = Mp[srcs]
        S = sfft.rfftn(lat.reshape(-1, p, p, p), s=(n, n, n), axes=(1, 2, 3), workers=workers)
        del lat
        nf = S[0].size
        S = np.ascontiguousarray(S.reshape(len(srcs), nf)).view(np.float64)
        tables = getattr(ops, "_tables_flat", None)
        if tables is None:
            tables = np.ascontiguousarray(ops.tables.reshape(len(ops.tables), nf)).view(np.float64)
            ops._tables_flat = tables
        ts = np.nonzero(np.diff(ptr))[0].astype(np.int64)
        U = np.empty((len(ts), 2 * nf))
        lib().ohadamard(np.ascontiguousarray(ptr, np.int32), np.ascontiguousarray(slot[src]),
                        np.ascontiguousarray(far, np.int32), ts, len(ts), tables, S, nf, U)
        del S
        U = U.view(np.complex128).reshape(len(ts), n, n, n // 2 + 1)
        u = sfft.irfftn(U, s=(n, n, n), axes=(1, 2, 3), workers=workers).reshape(len(ts), -1)[:, ops.embed]
        pL[ts] += alpha * u

    # evaluate_at_sources! (src/evaluate.jl:111-127)
    def evaluate_at_sources(self, Lc, tau_sorted, grad=False):
        t = self.tree
        N = len(t.perm)
        out = np.empty((N, 3) if grad else N)
        lib().oeval(t.points, t.point_node, np.arange(N, dtype=np.int32), t.perm.astype(np.int64), N,
                    np.ascontiguousarray(Lc), np.arange(t.nnodes, dtype=np.int64), self.surf, self.lam,
                    self.ops.nlarge, t.level, self.centers, t.colleagues, t.plo, t.phi, t.points, tau_sorted,
                    int(grad), out)
        return out

    # evaluate! (src/evaluate.jl:81-96)
    def evaluate(self, Lc, tau_sorted, targets, grad=False):
        t = self.tree
        x = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
        nodes = t.locate(x)
        nt = len(x)
        out = np.empty((nt, 3) if grad else nt)
        lib().oeval(x, nodes, np.full(nt, -1, np.int32), np.arange(nt, dtype=np.int64), nt,
                    np.ascontiguousarray(Lc), np.arange(t.nnodes, dtype=np.int64), self.surf, self.lam,
                    self.ops.nlarge, t.level, self.centers, t.colleagues, t.plo, t.phi, t.points, tau_sorted,
                    int(grad), out)
        return out

    def matvec(self, tau, grad=False, m2l="dense", targets=None):
        """``build_expansions`` then ``evaluate_at_sources`` (and ``evaluate`` at
        ``targets`` when given); outputs in the original order."""
        tau_sorted = np.ascontiguousarray(np.asarray(tau, dtype=np.float64)[self.tree.perm])
        Mp = self.upward(tau_sorted)
        Pl, Lc = self.downward(Mp, m2l)
        self.Lc = Lc
        # column per node (Ns x nnodes, Nl x nnodes) views, as the reference stores them
        self.multipoles, self.plocals, self.locals = Mp.T, Pl.T, Lc.T
        u = self.evaluate_at_sources(Lc, tau_sorted, grad)
        if targets is None:
            return u
        return u, self.evaluate(Lc, tau_sorted, targets, grad)
