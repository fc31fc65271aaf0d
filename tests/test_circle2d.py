"""The 2D circle method on the CPU: the host operators (ffmm/circle2d.py) against what
the mathematics fixes, and the 2D oracle (oracle/circle2d.py) against brute-force tree
invariants and FP64 direct sums.

Reference checks of the same kind: the circle operators' accuracy statement
(src/operators/kifmm.jl:7-11: worst-case translation error ~0.547^N, floor ~1e-14 near
N = 64), the tree invariants of test/test_tree.jl:7-131 and the matvec against direct
sums (README.md:38-51; docs/kifmm.md:169: circles 60 at 8e-14)."""
import math

import numpy as np
import pytest

import circle2d as O2


@pytest.fixture(scope="module")
def C2():
    from ffmm import circle2d
    return circle2d


def test_enumerations(C2):
    offs = C2.far_offsets()
    assert len(offs) == 40 and np.all(np.max(np.abs(offs), axis=1) >= 2)
    # table order, coordinate 0 fastest (src/util.jl:59-84)
    key = (offs[:, 0] + 3) + 7 * (offs[:, 1] + 3)
    assert np.all(np.diff(key) > 0)
    assert [tuple(o) for o in offs] == O2.FAR
    assert C2.child_offsets().tolist() == [[0, 0], [1, 0], [0, 1], [1, 1]]
    p = C2.circle_points(60, 0.3)
    assert np.allclose(np.hypot(p[:, 0], p[:, 1]), 0.3, rtol=0, atol=1e-15)
    assert np.allclose(p[15], [0.0, 0.3], atol=1e-15)          # a quarter turn


def _field(sources, q, x):
    return O2.kernel(sources, x) @ q


@pytest.mark.parametrize("n,tol", [(32, 1e-6), (60, 1e-12)])
def test_translations_reproduce_fields(C2, n, tol):
    """Charges in a box of width w: the multipole (S2M + UC2E) reproduces their field
    beyond the box's neighbours; M2M reproduces the children's fields; M2L + DC2E + L2T
    reproduces a far box's field inside the target box, at every far offset; L2L keeps
    the parent's local field inside the children."""
    rng = np.random.default_rng(n)
    w = 0.25
    m = C2.KernelIndependent(C2.Laplace2D(), nequiv=n)
    o = C2.CircleOperators(n, w, m.pinv_rtol)
    oc = C2.CircleOperators(n, w / 2, m.pinv_rtol)
    src = (rng.random((40, 2)) - 0.5) * w
    q = rng.standard_normal(40)
    P = o.P_up
    M = P.V @ (P.invS * (P.Uh @ _field(src, q, o.large)))
    far_pts = np.array([[3 * w * math.cos(t), 3 * w * math.sin(t)] for t in np.linspace(0, 6, 25)])
    ex = _field(src, q, far_pts)
    assert np.max(np.abs(_field(o.small, M, far_pts) - ex)) < tol * np.max(np.abs(ex))
    # M2M: four children's multipoles to the parent
    Mc = []
    for c in C2.child_offsets():
        cc = (c - 0.5) * (w / 2)
        s = cc + (rng.random((10, 2)) - 0.5) * (w / 2)
        qq = rng.standard_normal(10)
        Mc.append((s, qq, oc.P_up.V @ (oc.P_up.invS * (oc.P_up.Uh @ _field(s - cc, qq, oc.large)))))
    Mpar = o.M2M @ np.concatenate([x[2] for x in Mc])
    exc = sum(_field(s, qq, far_pts) for s, qq, _ in Mc)
    assert np.max(np.abs(_field(o.small, Mpar, far_pts) - exc)) < tol * np.max(np.abs(exc))
    # M2L at every far offset, then the local density and its field inside the target box
    tgt = (rng.random((30, 2)) - 0.5) * w
    Q = o.P_down
    worst = 0.0
    for f, off in enumerate(C2.far_offsets()):
        shift = off * w
        pL = o.M2L[f] @ M
        Lden = Q.V @ (Q.invS * (Q.Uh @ pL))
        ex = _field(src + shift, q, tgt)
        worst = max(worst, np.max(np.abs(_field(o.large, Lden, tgt) - ex)) / np.max(np.abs(ex)))
    assert worst < tol
    # L2L: the parent's local field, evaluated from the children's local densities
    Lchild = (o.L2L @ pL).reshape(4, n)
    for k, c in enumerate(C2.child_offsets()):
        cc = (c - 0.5) * (w / 2)
        pts = cc + (rng.random((10, 2)) - 0.5) * (w / 2)
        Ld = oc.P_down.V @ (oc.P_down.invS * (oc.P_down.Uh @ Lchild[k]))
        assert np.max(np.abs(_field(oc.large, Ld, pts - cc) - _field(o.large, Lden, pts))) < 10 * tol * np.max(
            np.abs(_field(o.large, Lden, pts)))


def _layout(kind, n, rng):
    if kind == "uniform":
        return rng.random((n, 2))
    # clusters on a line and a dense blob: a deep, unbalanced tree (restriction, ghosts)
    a = np.stack([rng.random(n // 2), 0.5 + 1e-3 * rng.standard_normal(n // 2)], 1)
    b = 0.25 + 0.01 * rng.standard_normal((n - n // 2, 2))
    return np.clip(np.concatenate([a, b]), 0, 1)


@pytest.mark.parametrize("kind,n,ppl", [("uniform", 3000, 50), ("clustered", 2500, 20)])
def test_oracle_tree_invariants(kind, n, ppl):
    """test/test_tree.jl:7-131 in 2D: every point in exactly one real leaf, leaves within
    ppl, colleagues by brute force over the level's cells, 2:1 balance between
    neighbouring leaves, ghosts only below xlist leaves; lists equal V ∩ sourced."""
    rng = np.random.default_rng(2)
    t = O2.build_tree(_layout(kind, n, rng), ppl, (0.0, 0.0), 1.0)
    real = np.nonzero(t.leaf)[0]
    cnt = np.zeros(n, int)
    for i in real:
        cnt[t.perm[t.plo[i]:t.phi[i] + 1]] += 1
        assert t.phi[i] - t.plo[i] + 1 <= ppl
    assert np.all(cnt == 1)
    assert np.array_equal(np.sort(t.perm), np.arange(n))
    for l in range(t.nlevels):
        r = range(int(t.level_start[l]), int(t.level_start[l + 1]))
        where = {tuple(t.cell[i]): i for i in r}
        for i in r:
            for k, o in enumerate(O2.COLL):
                j = where.get((t.cell[i][0] + o[0], t.cell[i][1] + o[1]), -1)
                assert t.colleagues[i, k] == j
    # 2:1 balance: the children of a real leaf's colleagues have no real children
    for i in real:
        for j in t.colleagues[i]:
            if j < 0 or t.child_first[j] < 0:
                continue
            for c in range(t.child_first[j], t.child_first[j] + 4):
                assert t.child_first[c] < 0 or t.ghost[t.child_first[c]]
    assert np.all(t.xlist[t.parent[np.nonzero(t.ghost)[0]]] == 1)
    if kind == "clustered":
        assert t.ghost.sum() > 0 and t.nlevels >= 8


def test_oracle_fmm_against_direct():
    from ffmm import circle2d as C2
    rng = np.random.default_rng(3)
    for kind, n, ppl, nequiv, gate in (("uniform", 3000, 50, 60, 1e-12), ("clustered", 2500, 20, 60, 1e-12),
                                       ("uniform", 3000, 50, 32, 1e-7)):
        pts = _layout(kind, n, rng)
        tau = rng.standard_normal(n)
        tgt = rng.random((200, 2))
        cache = C2.CircleOperatorCache(C2.KernelIndependent(C2.Laplace2D(), nequiv=nequiv))
        f = O2.FMM2D(cache.operators, pts, ppl, (0.0, 0.0), 1.0)
        u, v = f.matvec(tau, tgt)
        ud = O2.direct_self(pts, tau)
        vd = O2.direct(pts, tau, tgt)
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < gate, (kind, nequiv)
        assert np.linalg.norm(v - vd) / np.linalg.norm(vd) < gate, (kind, nequiv)
