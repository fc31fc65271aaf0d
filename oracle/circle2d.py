"""ORACLE -- test infrastructure only.  Only tests/ (and nothing in the product library)
imports this module, as the checker of the 2D circle method.

A plain-Python / numpy restatement of the reference's 2D kernel-independent FMM
(``KernelIndependent(Laplace2D(); nequiv)``) for small inputs:

  tree: split_nodes! (stable counting sort by child slot), tag_colleagues! (the parent
        rule), level_restrict!, add_ghosts!, flatten      src/tree/builder.jl:40-470
  lists: _count_pairs! / _fill_pairs!                       src/tree/lists.jl:71-124
  passes: upward_pass! / downward_pass! / m2l_level!        src/fmm.jl:100-127, 170-262
  point loops: s2m_accumulate!, l2t, evaluate_target        src/operators/ki_common.jl:256-315,
                                                            src/evaluate.jl:12-58
  direct sums: direct! / direct_self!                        src/kernels.jl:442-487

Indices are 0-based; a missing node is -1.  The operators are the product's host
operators (ffmm.circle2d.CircleOperators), pinned on their own by tests/test_circle2d.py.
"""
from __future__ import annotations

import math

import numpy as np

C = -1.0 / (4.0 * math.pi)


def kernel(sources: np.ndarray, targets: np.ndarray) -> np.ndarray:
    d = targets[:, None, :] - sources[None, :, :]
    return C * np.log(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])


def direct(sources, tau, targets) -> np.ndarray:
    """direct! (src/kernels.jl:442-462): sum over all sources, in chunks."""
    out = np.empty(len(targets))
    for a in range(0, len(targets), 512):
        out[a:a + 512] = kernel(sources, targets[a:a + 512]) @ tau
    return out


def direct_self(sources, tau, rows=None) -> np.ndarray:
    """direct_self! (src/kernels.jl:464-487): the source itself excluded."""
    rows = np.arange(len(sources)) if rows is None else np.asarray(rows)
    out = np.empty(len(rows))
    for a in range(0, len(rows), 256):
        r = rows[a:a + 256]
        d = sources[r][:, None, :] - sources[None, :, :]
        r2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
        r2[np.arange(len(r)), r] = 1.0           # log 1 = 0: the excluded source
        out[a:a + 256] = (C * np.log(r2)) @ tau
    return out


# ---------------------------------------------------------------------------
# the tree (src/tree/builder.jl) for D = 2
# ---------------------------------------------------------------------------
COLL = [(k % 3 - 1, k // 3 - 1) for k in range(9)]      # colleague_offsets, coordinate 0 fastest
SELF = 4


class _Level:
    def __init__(self):
        self.cell, self.parent, self.child_first, self.lo, self.hi = [], [], [], [], []
        self.leaf, self.ghost, self.xlist, self.depth = [], [], [], []
        self.col = []

    def n(self):
        return len(self.cell)


def _center(origin, W0, l, cell):
    w = math.ldexp(W0, -l)
    return (origin[0] + (cell[0] + 0.5) * w, origin[1] + (cell[1] + 0.5) * w)


def _split(par, which, l, chi, pts, perm, origin, W0, ghost):
    for i in which:
        lo, hi = par.lo[i], par.hi[i]
        c = _center(origin, W0, l, par.cell[i])
        slots = [int(pts[j][0] > c[0]) | (int(pts[j][1] > c[1]) << 1) for j in range(lo, hi + 1)]
        order = sorted(range(len(slots)), key=lambda q: slots[q])     # stable
        sub_p = [pts[lo + q] for q in order]
        sub_m = [perm[lo + q] for q in order]
        pts[lo:hi + 1] = sub_p
        perm[lo:hi + 1] = sub_m
        counts = [slots.count(k) for k in range(4)]
        par.child_first[i] = chi.n()
        if not ghost:
            par.leaf[i] = False
        s = lo
        for k in range(4):
            chi.cell.append((2 * par.cell[i][0] + (k & 1), 2 * par.cell[i][1] + ((k >> 1) & 1)))
            chi.parent.append(i)
            chi.child_first.append(-1)
            chi.lo.append(s)
            chi.hi.append(s + counts[k] - 1)
            chi.leaf.append(not ghost)
            chi.ghost.append(ghost)
            chi.xlist.append(False)
            chi.depth.append(0)
            s += counts[k]


def _tag(lev, par):
    lev.col = []
    for i in range(lev.n()):
        c, p = lev.cell[i], lev.parent[i]
        row = []
        for o in COLL:
            cc = (c[0] + o[0], c[1] + o[1])
            d = ((cc[0] >> 1) - (c[0] >> 1), (cc[1] >> 1) - (c[1] >> 1))
            q = par.col[p][(d[0] + 1) + 3 * (d[1] + 1)]
            if q < 0 or par.child_first[q] < 0:
                row.append(-1)
            else:
                row.append(par.child_first[q] + ((cc[0] & 1) | ((cc[1] & 1) << 1)))
        lev.col.append(row)


def _depths(levels):
    L = len(levels)
    levels[-1].depth = [0] * levels[-1].n()
    for l in range(L - 2, -1, -1):
        lev, nxt = levels[l], levels[l + 1]
        for i in range(lev.n()):
            cf = lev.child_first[i]
            lev.depth[i] = 0 if cf < 0 else 1 + max(nxt.depth[cf + k] for k in range(4))


class Tree2D:
    pass


def build_tree(points, ppl: int, origin, W0: float, max_levels: int = 30) -> Tree2D:
    pts = [tuple(p) for p in np.asarray(points, dtype=np.float64)]
    perm = list(range(len(pts)))
    origin = (float(origin[0]), float(origin[1]))
    root = _Level()
    root.cell, root.parent, root.child_first = [(0, 0)], [-1], [-1]
    root.lo, root.hi, root.leaf, root.ghost, root.xlist, root.depth = [0], [len(pts) - 1], [True], [False], [False], [0]
    levels = [root]
    while True:
        lev = levels[-1]
        which = [i for i in range(lev.n()) if lev.hi[i] - lev.lo[i] + 1 > ppl]
        if not which:
            break
        if len(levels) >= max_levels:
            raise ValueError("tree reached max_levels")
        nxt = _Level()
        _split(lev, which, len(levels) - 1, nxt, pts, perm, origin, W0, False)
        levels.append(nxt)
    root.col = [[-1] * 9]
    root.col[0][SELF] = 0
    for l in range(1, len(levels)):
        _tag(levels[l], levels[l - 1])
    # level_restrict!: levels 2 .. L-3 (0-based 1 .. L-3)
    L = len(levels)
    _depths(levels)
    while True:
        nsplit = 0
        for l in range(1, L - 2):
            lev = levels[l]
            which = [i for i in range(lev.n()) if lev.leaf[i] and
                     any(j >= 0 and abs(lev.depth[i] - lev.depth[j]) > 1 for j in lev.col[i])]
            if not which:
                continue
            _split(lev, which, l, levels[l + 1], pts, perm, origin, W0, False)
            _tag(levels[l + 1], lev)
            _depths(levels)
            nsplit += len(which)
        if nsplit == 0:
            break
    # add_ghosts!: mark all levels first, then split (levels 1 .. L-2, 0-based)
    for l in range(1, L - 1):
        lev = levels[l]
        for i in range(lev.n()):
            if lev.leaf[i] and any(j >= 0 and lev.child_first[j] >= 0 for j in lev.col[i]):
                lev.xlist[i] = True
    for l in range(1, L - 1):
        lev = levels[l]
        which = [i for i in range(lev.n()) if lev.xlist[i]]
        if which:
            _split(lev, which, l, levels[l + 1], pts, perm, origin, W0, True)
            _tag(levels[l + 1], lev)
    # flatten
    counts = [lev.n() for lev in levels]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    t = Tree2D()
    t.origin, t.W0, t.nlevels, t.level_start = origin, W0, L, offs
    nn = int(offs[-1])
    t.nnodes = nn
    t.cell = np.zeros((nn, 2), np.int32)
    t.level = np.zeros(nn, np.int32)
    t.parent = np.full(nn, -1, np.int32)
    t.child_first = np.full(nn, -1, np.int32)
    t.plo = np.zeros(nn, np.int32)
    t.phi = np.zeros(nn, np.int32)
    t.leaf = np.zeros(nn, np.uint8)
    t.ghost = np.zeros(nn, np.uint8)
    t.xlist = np.zeros(nn, np.uint8)
    t.colleagues = np.full((nn, 9), -1, np.int32)
    t.point_node = np.full(len(pts), -1, np.int64)
    for l, lev in enumerate(levels):
        for i in range(lev.n()):
            g = offs[l] + i
            t.cell[g] = lev.cell[i]
            t.level[g] = l
            t.parent[g] = -1 if l == 0 else offs[l - 1] + lev.parent[i]
            t.child_first[g] = -1 if lev.child_first[i] < 0 else offs[l + 1] + lev.child_first[i]
            t.plo[g], t.phi[g] = lev.lo[i], lev.hi[i]
            t.leaf[g], t.ghost[g], t.xlist[g] = lev.leaf[i], lev.ghost[i], lev.xlist[i]
            t.colleagues[g] = [-1 if j < 0 else offs[l] + j for j in lev.col[i]]
            if lev.leaf[i]:
                t.point_node[lev.lo[i]:lev.hi[i] + 1] = g
    t.has_source = (t.phi >= t.plo).astype(np.uint8)
    t.perm = np.array(perm, dtype=np.int32)
    t.points = np.array(pts, dtype=np.float64).reshape(-1, 2)
    t.lists = [_lists(t, l) for l in range(L)]
    return t


FAR = [(i % 7 - 3, i // 7 - 3) for i in range(49)]
FAR = [o for o in FAR if max(abs(o[0]), abs(o[1])) >= 2]
FAR_INDEX = {o: k for k, o in enumerate(FAR)}


def _lists(t, l):
    """_count_pairs! / _fill_pairs! (src/tree/lists.jl:71-124): per non-ghost target of
    level l >= 2 (0-based), the sourced children of its parent's colleagues at a far
    offset, in colleague-then-child order."""
    lo, hi = int(t.level_start[l]), int(t.level_start[l + 1])
    ptr, src, far = [0], [], []
    for i in range(lo, hi):
        if l >= 2 and not t.ghost[i]:
            p = t.parent[i]
            for k in range(9):
                cc = t.colleagues[p, k]
                if cc < 0 or t.child_first[cc] < 0:
                    continue
                for s in range(t.child_first[cc], t.child_first[cc] + 4):
                    if not t.has_source[s]:
                        continue
                    o = (int(t.cell[s, 0] - t.cell[i, 0]), int(t.cell[s, 1] - t.cell[i, 1]))
                    if max(abs(o[0]), abs(o[1])) >= 2:
                        src.append(s)
                        far.append(FAR_INDEX[o])
        ptr.append(len(src))
    return np.array(ptr, np.int32), np.array(src, np.int32), np.array(far, np.int32)


# ---------------------------------------------------------------------------
# the passes (src/fmm.jl:170-262)
# ---------------------------------------------------------------------------
class FMM2D:
    def __init__(self, ops_of_width, points, ppl, origin, W0, max_levels=30):
        self.t = build_tree(points, ppl, origin, W0, max_levels)
        self.ops = ops_of_width
        t = self.t
        w = [math.ldexp(W0, -l) for l in range(t.nlevels)]
        self.width = w
        self.centers = np.stack([t.origin[0] + (t.cell[:, 0] + 0.5) * np.array(w)[t.level],
                                 t.origin[1] + (t.cell[:, 1] + 0.5) * np.array(w)[t.level]], axis=1)

    def matvec(self, tau, targets=None):
        t = self.t
        L = t.nlevels
        n = self.ops(self.width[min(2, L - 1)]).n
        ts = np.asarray(tau, dtype=np.float64)[t.perm]
        Mp = np.zeros((t.nnodes, n))
        for l in range(L - 1, 1, -1):
            o = self.ops(self.width[l])
            r = range(int(t.level_start[l]), int(t.level_start[l + 1]))
            nodes = [i for i in r if t.child_first[i] < 0 and t.has_source[i]]
            for i in nodes:
                a, b = t.plo[i], t.phi[i] + 1
                # s2m_accumulate!: check potentials on the large circle
                pM = kernel(t.points[a:b] - self.centers[i], o.large) @ ts[a:b]
                P = o.P_up
                Mp[i] = P.V @ (P.invS * (P.Uh @ pM))
            if l < L - 1:
                cr = range(int(t.level_start[l + 1]), int(t.level_start[l + 2]))
                for g0 in range(cr.start, cr.stop, 4):
                    Mp[t.parent[g0]] = o.M2M @ Mp[g0:g0 + 4].reshape(-1)
        Pl = np.zeros((t.nnodes, n))
        Lc = np.zeros((t.nnodes, n))
        for l in range(2, L):
            o = self.ops(self.width[l])
            lo, hi = int(t.level_start[l]), int(t.level_start[l + 1])
            if l > 2:
                po = self.ops(self.width[l - 1])
                for g0 in range(lo, hi, 4):
                    Pl[g0:g0 + 4] = (po.L2L @ Pl[t.parent[g0]]).reshape(4, n)
            ptr, src, far = t.lists[l]
            for q in range(hi - lo):
                for e in range(ptr[q], ptr[q + 1]):
                    Pl[lo + q] += o.M2L[far[e]] @ Mp[src[e]]
            for i in range(lo, hi):
                if t.leaf[i]:
                    P = o.P_down
                    Lc[i] = P.V @ (P.invS * (P.Uh @ Pl[i]))
        self.Mp, self.Pl, self.Lc = Mp, Pl, Lc
        u = np.empty(len(ts))
        for i in range(t.nnodes):
            if not t.leaf[i] or t.phi[i] < t.plo[i]:
                continue
            a, b = t.plo[i], t.phi[i] + 1
            u[t.perm[a:b]] = self._eval(i, t.points[a:b], np.arange(a, b), ts)
        if targets is None:
            return u
        x = np.asarray(targets, dtype=np.float64)
        v = np.empty(len(x))
        for q, y in enumerate(x):
            v[q] = self._eval(self.locate(y), y[None, :], np.array([-1]), ts)[0]
        return u, v

    def locate(self, y):
        """locate (src/tree/locate.jl:10-22): descend by child_slot to a leaf."""
        t = self.t
        i = 0
        while not t.leaf[i]:
            c = self.centers[i]
            i = t.child_first[i] + (int(y[0] > c[0]) | (int(y[1] > c[1]) << 1))
        return i

    @np.errstate(divide="ignore", invalid="ignore")
    def _eval(self, i, x, skip, ts):
        """evaluate_target (src/evaluate.jl:37-58): l2t of the leaf plus the near field of
        its colleagues, the source `skip` excluded."""
        t = self.t
        o = self.ops(self.width[t.level[i]])
        out = kernel(o.large, x - self.centers[i]) @ self.Lc[i]
        for j in t.colleagues[i]:
            if j < 0 or t.phi[j] < t.plo[j]:
                continue
            a, b = t.plo[j], t.phi[j] + 1
            K = kernel(t.points[a:b], x)
            K[skip[:, None] == np.arange(a, b)[None, :]] = 0.0
            out = out + K @ ts[a:b]
        return out
