"""Brute-force structural checks of a ghost-cell tree and its interaction lists,
restating the reference's test/test_tree.jl:7-131 (every invariant of the tree) and
test/test_coverage.jl:7-84 (every (source terminal, target leaf) pair covered exactly
once, by the near field or by one V-list pair).  Applied to the oracle's trees here
and to the device's trees in the GPU tests."""
import numpy as np

from ffmm import enumerations as en


def check_tree_invariants(t, original, locate=None):
    N = len(original)
    n = len(t.level)
    L = len(t.level_start) - 1
    sp = original[t.perm]
    assert np.array_equal(np.sort(t.perm), np.arange(N))
    assert t.level_start[0] == 0 and t.level_start[-1] == n and np.all(np.diff(t.level_start) > 0)
    for l in range(L):
        assert np.all(t.level[t.level_start[l]:t.level_start[l + 1]] == l)
    assert t.plo[0] == 0 and t.phi[0] == N - 1 and t.parent[0] == -1 and np.all(t.cell[0] == 0)
    offs = en.child_offsets()
    w = np.ldexp(t.W0, -t.level.astype(np.int64))
    for i in range(n):
        l = t.level[i]
        assert np.all((0 <= t.cell[i]) & (t.cell[i] < 2 ** l))
        blo = t.origin + t.cell[i] * w[i]
        bhi = t.origin + (t.cell[i] + 1) * w[i]
        pts = sp[t.plo[i]:t.phi[i] + 1]
        assert np.all((blo <= pts) & (pts <= bhi))
        assert bool(t.has_source[i]) == (t.phi[i] >= t.plo[i])
        assert not (t.ghost[i] and t.leaf[i])
        cf = t.child_first[i]
        if cf >= 0:
            ch = np.arange(cf, cf + 8)
            assert np.all(t.parent[ch] == i) and np.all(t.level[ch] == l + 1)
            assert np.array_equal(t.cell[ch], 2 * t.cell[i] + offs)
            assert t.plo[ch[0]] == t.plo[i] and t.phi[ch[-1]] == t.phi[i]
            assert np.all(t.phi[ch[:-1]] + 1 == t.plo[ch[1:]])
            c = t.origin + (t.cell[i] + 0.5) * w[i]
            for k in range(8):
                p = sp[t.plo[ch[k]]:t.phi[ch[k]] + 1]
                slot = (p[:, 0] > c[0]).astype(int) | ((p[:, 1] > c[1]) << 1) | ((p[:, 2] > c[2]) << 2)
                assert np.all(slot == k)
            assert not t.ghost[i]
            if t.leaf[i]:
                assert t.xlist[i] and np.all(t.ghost[ch])
            else:
                assert not np.any(t.ghost[ch])
        else:
            assert not t.xlist[i]
            assert t.leaf[i] or t.ghost[i]
    # sibling groups
    for l in range(1, L):
        r = np.arange(t.level_start[l], t.level_start[l + 1])
        assert len(r) % 8 == 0
        for g in range(len(r) // 8):
            f = r[0] + 8 * g
            p = t.parent[f]
            assert t.child_first[p] == f and np.all(t.parent[f:f + 8] == p)
    # unique cells and brute-force colleagues
    co = en.colleague_offsets()
    for l in range(L):
        r = range(t.level_start[l], t.level_start[l + 1])
        lookup = {tuple(t.cell[i]): i for i in r}
        assert len(lookup) == len(r)
        for i in r:
            assert t.colleagues[i, en.SELF_SLOT] == i
            for k, o in enumerate(co):
                assert t.colleagues[i, k] == lookup.get(tuple(t.cell[i] + o), -1)
    # real depths and 2:1 balance on colleagues
    rdepth = np.zeros(n, int)
    for l in range(L - 1, -1, -1):
        for i in range(t.level_start[l], t.level_start[l + 1]):
            if t.child_first[i] >= 0 and not t.leaf[i]:
                rdepth[i] = 1 + rdepth[t.child_first[i]:t.child_first[i] + 8].max()
    for i in np.nonzero(t.leaf)[0]:
        for j in t.colleagues[i]:
            if j >= 0 and not t.ghost[j]:
                assert abs(rdepth[i] - rdepth[j]) <= 1
    # ghosts
    for i in range(n):
        if not t.leaf[i] and not t.ghost[i]:
            for j in t.colleagues[i]:
                if j >= 0:
                    assert not t.ghost[j] and t.child_first[j] >= 0
        if t.leaf[i]:
            nonleaf = any(j >= 0 and not t.leaf[j] and not t.ghost[j] for j in t.colleagues[i])
            assert bool(t.xlist[i]) == nonleaf
    # point_node and the locator
    for j in range(N):
        p = t.point_node[j]
        assert t.leaf[p] and t.plo[p] <= j <= t.phi[p]
    if locate is not None:
        assert np.array_equal(locate(sp), t.point_node)


def vlist(t, y):
    p = t.parent[y]
    if p < 0:
        return set()
    cols = {j for j in t.colleagues[y] if j >= 0}
    out = set()
    for c in t.colleagues[p]:
        if c < 0 or t.child_first[c] < 0:
            continue
        for s in range(t.child_first[c], t.child_first[c] + 8):
            if s not in cols:
                out.add(s)
    return out


def check_coverage(t, lists):
    """lists[l] = (ptr, src, far) per level, targets in node order."""
    n = len(t.level)
    terminals = [i for i in range(n) if t.child_first[i] < 0 and t.has_source[i]]
    leaves = np.nonzero(t.leaf)[0]
    vl = [vlist(t, y) for y in range(n)]
    anc = []
    for i in range(n):
        a, j = [], i
        while j >= 0:
            a.append(j)
            j = t.parent[j]
        anc.append(set(a))
    bad = 0
    for T in terminals:
        for B in leaves:
            near = any(C >= 0 and C in anc[T] for C in t.colleagues[B])
            cnt = sum(1 for X in anc[T] for Y in anc[B] if X in vl[Y])
            bad += not ((near and cnt == 0) or (not near and cnt == 1))
    assert bad == 0
    far = en.far_offsets()
    pairs = set()
    total = 0
    for l, (ptr, src, fi) in enumerate(lists):
        lo = t.level_start[l]
        for q in range(len(ptr) - 1):
            for k in range(ptr[q], ptr[q + 1]):
                tg, s = lo + q, src[k]
                pairs.add((tg, s))
                total += 1
                assert t.level[s] == l and np.array_equal(t.cell[s] - t.cell[tg], far[fi[k]])
        if l < 2:
            assert ptr[-1] == 0
    assert len(pairs) == total
    expected = {(y, x) for y in range(n) if not t.ghost[y] for x in vl[y] if t.has_source[x]}
    assert pairs == expected
