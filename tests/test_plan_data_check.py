"""The host checks of plan data (ffmm_plan_data_check, the checks ffmm_plan_load and
ffmm_server_import run before anything reaches a device), without a GPU: plan data of
an oracle-built tree passes, and every corruption below is refused with the violated
rule named -- truncation, trailing bytes, sizes out of range, indices out of range
(children, parents, colleagues, point ranges, node lists), a non-permutation, a
non-finite point, and structure that the kernels' indexing relies on (child cells,
colleague offsets, partitioned point ranges, leaves covering every point once)."""
import struct

import numpy as np
import pytest

import oracle as O
import plandata
from pointsets import mixed


@pytest.fixture(scope="module")
def tree():
    rng = np.random.default_rng(71)
    return O.build_tree(mixed(rng, 6000), 30, np.zeros(3), 1.0)


def check(data):
    import ffmm
    ffmm.plan_data_check(data)


def refused(data, match):
    import ffmm
    with pytest.raises(ffmm.ArgumentError, match=match):
        ffmm.plan_data_check(data)


def test_valid_plan_data_passes(tree):
    assert tree.nlevels >= 5 and tree.ghost.any()
    check(plandata.encode(tree))


def test_framing_is_checked(tree):
    data = plandata.encode(tree)
    refused(data[:-3], "truncated")
    refused(data[: len(data) // 2], "truncated|wrong length")
    refused(data + b"\0", "trailing bytes")
    refused(b"NOTATREE" + data[8:], "not plan data")
    refused(data[:8] + struct.pack("<I", 9) + data[12:], "another version")
    # an array count far beyond the data
    f = plandata.fields(tree)
    bad = bytearray(data)
    off = data.index(struct.pack("<q", len(f["perm"])) + f["perm"].tobytes())
    bad[off:off + 8] = struct.pack("<q", 1 << 60)
    refused(bytes(bad), "wrong length")


def test_sizes_are_checked(tree):
    refused(plandata.encode(tree, L=40), "bad tree sizes")
    refused(plandata.encode(tree, L=0), "bad tree sizes")
    refused(plandata.encode(tree, nnodes=1 << 40), "bad tree sizes")
    refused(plandata.encode(tree, max_levels=3), "more levels than max_levels")


def corrupt(tree, key, fn):
    f = plandata.fields(tree)
    f[key] = np.array(f[key], copy=True)
    fn(f[key], f)
    return plandata.encode(tree, f)


def test_indices_out_of_range_are_refused(tree):
    n, N = tree.nnodes, len(tree.perm)
    inner = int(np.nonzero(tree.child_first >= 0)[0][3])
    leaf = int(np.nonzero(tree.leaf)[0][5])
    cases = [
        ("child_first", lambda a, f: a.__setitem__(inner, n + 100), "children outside"),
        ("child_first", lambda a, f: a.__setitem__(inner, a[inner] + 3), "children outside|parent"),
        ("parent", lambda a, f: a.__setitem__(n - 1, n + 7), "parent"),
        ("col", lambda a, f: a.__setitem__(5 * n + leaf, n + 12345), "colleague outside"),
        ("col", lambda a, f: a.__setitem__(13 * n + leaf, -1), "colleague outside"),
        ("phi", lambda a, f: a.__setitem__(leaf, N + 10), "point range"),
        ("plo", lambda a, f: a.__setitem__(leaf, -4), "point range"),
        ("perm", lambda a, f: a.__setitem__(7, N + 1), "permutation"),
        ("perm", lambda a, f: a.__setitem__(7, a[8]), "permutation"),
        ("s2m_nodes", lambda a, f: a.__setitem__(0, n + 3), "node lists"),
        ("leaf_nodes", lambda a, f: a.__setitem__(len(a) - 1, a[0]), "node lists"),
        ("s2m_off", lambda a, f: a.__setitem__(2, a[3] + 1), "node lists"),
        ("point_node", lambda a, f: a.__setitem__(11, n + 9), "point nodes"),
        ("cell", lambda a, f: a.__setitem__(0, 1), "cell outside"),
        ("cell", lambda a, f: a.__setitem__(n - 1, 1 << 25), "cell outside|slot"),
        ("level", lambda a, f: a.__setitem__(n - 1, 0), "node levels"),
        ("leaf", lambda a, f: a.__setitem__(leaf, 7), "flag other than"),
        ("has_source", lambda a, f: a.__setitem__(leaf, 1 - a[leaf]), "source flags"),
    ]
    for key, fn, match in cases:
        refused(corrupt(tree, key, fn), match)


def test_structure_is_checked(tree):
    n = tree.nnodes
    inner = int(np.nonzero((tree.child_first >= 0) & (tree.level >= 2))[0][0])
    cf = int(tree.child_first[inner])
    leaf = int(np.nonzero(tree.leaf & (tree.level >= 3) & (tree.child_first < 0))[0][0])
    has_col = int(np.nonzero((tree.colleagues[:, 0] >= 0) & (tree.level >= 3))[0][0])

    def swap_children(a, f):                 # two sibling cells exchanged
        for d in range(3):
            a[d * n + cf], a[d * n + cf + 1] = a[d * n + cf + 1], a[d * n + cf]

    def shift_range(a, f):                   # a child's points no longer follow its sibling's
        f["plo"][cf + 1] += 1

    def wrong_offset(a, f):                  # a colleague moved to another slot
        a[0 * n + has_col], a[26 * n + has_col] = a[26 * n + has_col], a[0 * n + has_col]

    def unleaf(a, f):                        # a leaf that is not flagged: childless, neither leaf nor ghost
        a[leaf] = 0

    cases = [("cell", swap_children, "slot"), ("plo", shift_range, "partition"),
             ("col", wrong_offset, "wrong offset"), ("leaf", unleaf, "neither leaf nor ghost|point")]
    for key, fn, match in cases:
        refused(corrupt(tree, key, fn), match)
    # a point outside the box, a non-finite point
    refused(corrupt(tree, "sx", lambda a, f: a.__setitem__(3, 1.5)), "outside the requested box")
    refused(corrupt(tree, "sy", lambda a, f: a.__setitem__(3, np.nan)), "finite")
    # a point covered by two leaves: one leaf's range widened into its neighbour's
    lv = np.nonzero(tree.leaf & (tree.phi >= tree.plo))[0]
    a_, b_ = int(lv[10]), int(lv[11])
    refused(corrupt(tree, "phi", lambda a, f: a.__setitem__(a_, a[b_])), "partition|two leaves")
