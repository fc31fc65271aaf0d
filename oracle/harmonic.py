"""ORACLE -- test infrastructure only (imported by tests/ as the checker).

A CPU restatement of the reference's host FMM for the harmonic methods:
``SphericalHarmonics`` with Laplace3D or Yukawa3D charges, and ``StokesHarmonics``
with stokeslet (and stresslet) densities:

* the tree and lists: oracle/tree_oracle.c, as for the cube method;
* the per-point S2M, L2T and near field: oracle/harmonic_oracle.c;
* ``upward_pass!`` / ``downward_pass!`` (src/fmm.jl:170-262) with numpy for the
  translations: M2M and L2L as the folded matrices of each level (the Stokes ones in
  their dense 4N x 32N form, src/operators/harmonics_stokes.jl:112-131), the M2L per
  symmetry class by gather / representative / scatter with the source- and
  target-side signed permutations (``m2l_level!``, src/operators/harmonics.jl:531-640),
  the Stokes coupling Psi += o.Phi in the scatter; partial_to_multipole! and
  partial_to_local! are copies (src/operators/harmonics.jl:441-442).

Operators come from ``ffmm.harmonics`` (the host precompute), pinned by
tests/test_harmonic_operators.py against closed forms and direct sums.
"""
from __future__ import annotations

import numpy as np

import oracle as O

KIND = {"laplace": 0, "yukawa": 1, "stokes": 2}


def _tabs(tb):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)   # noqa: E731
    e = f(tb.e) if len(tb.e) else np.zeros(1)
    return f(tb.a), f(tb.b), e, f(tb.cz), f(tb.cm), f(tb.cp)


def direct(kind: str, sources, dens, targets, output: int = 0, kappa: float = 0.0, skip=None) -> np.ndarray:
    """FP64 direct sums of the three kinds (``direct``, src/kernels.jl:463-487)."""
    s = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dens, dtype=np.float64)
    nd = 1 if d.ndim == 1 else d.shape[1]
    k = KIND[kind]
    nout = (1 if output == 3 else 3) if k == 2 else (1 if output == 0 else 3 if output == 1 else 4)
    out = np.empty((len(t), nout))
    sk = None
    if skip is not None:
        sk = np.ascontiguousarray(skip, dtype=np.int64)
    O.lib().oh_direct(k, float(kappa), s, d.reshape(-1), nd, len(s), t, len(t),
                      None if sk is None else sk.ctypes.data_as(O.ctypes.c_void_p), int(output), out)
    return out[:, 0] if nout == 1 else out


class HarmonicFMM:
    """``FMM(method, sources; ppl, bbox)`` for the harmonic methods.

    ``level_ops(l, w)`` returns the ``HarmonicOperators`` of level ``l`` (width ``w``) and
    the level's M2L factor: for the homogeneous kernels one reference set (width W0)
    with factor ``W0 / w``; for Yukawa the operators of the width itself and 1.
    """

    def __init__(self, kind: str, level_ops, points, ppl: int, origin, W0: float, kappa: float = 0.0,
                 max_levels: int = 30, perm=None):
        self.kind, self.k, self.kappa = kind, KIND[kind], float(kappa)
        self.tree = O.build_tree(points, ppl, origin, W0, max_levels)
        t = self.tree
        if perm is not None:
            perm = np.asarray(perm, dtype=np.int32)
            for i in np.nonzero(t.leaf)[0]:
                a, b = t.plo[i], t.phi[i] + 1
                if not np.array_equal(np.sort(perm[a:b]), np.sort(t.perm[a:b])):
                    raise O.OracleError("perm reorders points across leaves")
            t.perm = perm.copy()
            t.points = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[perm])
        L = t.nlevels
        self.ops = [level_ops(l, t.width(l)) for l in range(L)]
        op0 = self.ops[0][0]
        self.p, self.N = op0.p, op0.N
        self.nb = 4 if kind == "stokes" else 1
        self.tabs = _tabs(op0.tb)
        self.scales = np.array([1.0 / t.width(l) for l in range(L)])
        self.khat = np.array([self.kappa * t.width(l) for l in range(L)])
        self.centers = np.ascontiguousarray(t.centers())
        childless = t.child_first < 0
        self.s2m_nodes = [np.ascontiguousarray([i for i in t.level_range(l) if childless[i] and t.has_source[i]],
                                               dtype=np.int32) for l in range(L)]
        self.leaf_nodes = [np.ascontiguousarray([i for i in t.level_range(l) if t.leaf[i]], dtype=np.int32)
                           for l in range(L)]
        if kind == "stokes":
            from ffmm.harmonics import stokes_dense_operators
            self.dense = {}
            for l in range(L):
                o = self.ops[l][0]
                key = id(o)
                if key not in self.dense:
                    self.dense[key] = stokes_dense_operators(o.M2M, o.L2L, self.N)

    def _m2m_l2l(self, l):
        o = self.ops[l][0]
        if self.kind == "stokes":
            return self.dense[id(o)]
        return o.M2M, o.L2L

    def upward(self, dens_sorted: np.ndarray) -> np.ndarray:
        t = self.tree
        L = t.nlevels
        NB = self.nb * self.N
        d = np.ascontiguousarray(dens_sorted, dtype=np.float64)
        nd = 1 if d.ndim == 1 else d.shape[1]
        Mp = np.zeros((t.nnodes, NB))
        for l in range(L - 1, 1, -1):
            nodes = self.s2m_nodes[l]
            if len(nodes):
                M = np.empty((len(nodes), NB))
                O.lib().oh_s2m(self.k, self.p, *self.tabs, nodes, len(nodes), self.centers, t.level, self.scales,
                               self.khat, t.plo, t.phi, t.points, d.reshape(-1), nd, M)
                Mp[nodes] = M
            if l < L - 1:
                r = t.level_range(l + 1)
                B = Mp[r.start:r.stop].reshape(-1, 8 * NB)
                Mp[t.group_parent(l + 1)] = B @ self._m2m_l2l(l)[0].T
        return Mp

    def downward(self, Mp: np.ndarray):
        t = self.tree
        L = t.nlevels
        NB = self.nb * self.N
        Pl = np.zeros((t.nnodes, NB))
        for l in range(2, L):
            r = t.level_range(l)
            if l > 2:
                G = Pl[t.group_parent(l)]
                pL = (G @ self._m2m_l2l(l - 1)[1].T).reshape(-1, NB)
            else:
                pL = np.zeros((len(r), NB))
            self._m2l(pL, Mp, l)
            Pl[r.start:r.stop] = pL
        return Pl

    def _m2l(self, pL, Mp, l):
        from ffmm import enumerations as en
        ops, alpha = self.ops[l]
        sc = ops.sc
        N = self.N
        ptr, src, far = self.tree.lists[l]
        if len(src) == 0:
            return
        tgt = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        src = src.astype(np.int64)
        far = far.astype(np.int64)
        offs = np.asarray(en.far_offsets(), dtype=np.float64)
        klass = sc.klass[far]
        for c in range(len(ops.reps)):
            sel = np.nonzero(klass == c)[0]
            if len(sel) == 0:
                continue
            f, s_, t_ = far[sel], src[sel], tgt[sel]
            Y = []
            for b in range(self.nb):
                G = Mp[s_[:, None], b * N + sc.idx[f]] * sc.sgn[f]
                Y.append(alpha * (G @ ops.reps[c].T))
            if self.nb == 4:
                o = offs[f]
                Y[3] = Y[3] + o[:, 0:1] * Y[0] + o[:, 1:2] * Y[1] + o[:, 2:3] * Y[2]
            for b in range(self.nb):
                np.add.at(pL, (t_[:, None], b * N + sc.idx_t[f]), sc.sgn_t[f] * Y[b])

    def _eval(self, Pl, x, tnode, tskip, dens_sorted, output):
        t = self.tree
        d = np.ascontiguousarray(dens_sorted, dtype=np.float64)
        nd = 1 if d.ndim == 1 else d.shape[1]
        nout = (1 if output == 3 else 3) if self.k == 2 else (1 if output == 0 else 3 if output == 1 else 4)
        out = np.zeros((len(x), nout))
        O.lib().oh_l2t(self.k, self.p, *self.tabs, x, tnode, len(x), np.ascontiguousarray(Pl), self.centers, t.level,
                       self.scales, self.khat, int(output), out)
        O.lib().oh_near(self.k, self.kappa, x, tnode, tskip, len(x), np.ascontiguousarray(t.colleagues), t.plo, t.phi,
                        t.points, d.reshape(-1), nd, int(output), out)
        return out[:, 0] if nout == 1 else out

    def matvec(self, dens, output: int = 0, targets=None):
        """``build_expansions`` then ``evaluate_at_sources`` (and ``evaluate`` at
        ``targets``); outputs in the original order.  ``output``: 0 potential, 1
        gradient, 2 both (scalar kernels); the Stokes output is the velocity."""
        t = self.tree
        dens = np.asarray(dens, dtype=np.float64)
        ds = np.ascontiguousarray(dens[t.perm])
        Mp = self.upward(ds)
        Pl = self.downward(Mp)
        self.multipoles, self.locals = Mp, Pl
        N = len(t.perm)
        u_sorted = self._eval(Pl, t.points, t.point_node, np.arange(N, dtype=np.int32), ds, output)
        u = np.empty_like(u_sorted)
        u[t.perm] = u_sorted
        if targets is None:
            return u
        x = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
        v = self._eval(Pl, x, t.locate(x), np.full(len(x), -1, np.int32), ds, output)
        return u, v
