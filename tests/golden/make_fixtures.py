"""Writes the committed fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_fixtures.py``).

The reference ships no golden vectors (its checks are computed cross-checks), so the
fixtures are built here from the oracle and from FP64 direct sums, and each is checked
before it is written:

* ``tree_octants.npz``: eight clusters of six points, one per octant, ppl = 6: the
  closed-form tree of test/test_3d.jl:101-106 (two levels, nine nodes, no ghosts);
* ``tree_mixed300.npz``: 300 points of the sphere + cluster + background layout at
  ppl = 4 (restriction and ghosts): the oracle's structural arrays and lists, which pass
  every brute-force invariant and the exact-once coverage when generated;
* ``fmm_mixed2000.npz``: 2000 points, charges N(0, 1), 200 uniform targets, and the FP64
  direct sums at all the points (self excluded) and at the targets (potential and
  gradient): the values any p = 8 FMM must reproduce to 1e-6 (gradients 1e-5).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "flexiblefmm.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import oracle as O  # noqa: E402
from pointsets import mixed, uniform  # noqa: E402
from treecheck import check_coverage, check_tree_invariants  # noqa: E402

BOX = (np.zeros(3), 1.0)
FIELDS = ["level_start", "cell", "level", "parent", "child_first", "plo", "phi", "leaf", "ghost", "xlist",
          "has_source", "colleagues", "point_node"]


def tree_fixture(name, pts, ppl, closed_form=None):
    t = O.build_tree(pts, ppl, *BOX)
    check_tree_invariants(t, pts, locate=t.locate)
    check_coverage(t, t.lists)
    if closed_form:
        closed_form(t)
    arrays = {f: getattr(t, f) for f in FIELDS}
    for l, (ptr, src, far) in enumerate(t.lists):
        arrays[f"ptr{l}"], arrays[f"src{l}"], arrays[f"far{l}"] = ptr, src, far
    np.savez_compressed(os.path.join(HERE, name), points=pts, ppl=ppl, nlevels=t.nlevels, **arrays)


def octants_closed_form(t):
    assert t.nlevels == 2 and t.nnodes == 9 and t.ghost.sum() == 0
    assert np.array_equal(t.plo[1:], np.arange(0, 48, 6)) and np.array_equal(t.phi[1:], np.arange(5, 48, 6))
    assert np.all(t.leaf[1:]) and not t.leaf[0]


def main():
    rng = np.random.default_rng(20261202)
    oc = np.array([[0.25 + 0.5 * a, 0.25 + 0.5 * b, 0.25 + 0.5 * c] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
    tree_fixture("tree_octants.npz", np.repeat(oc, 6, axis=0) + 0.01 * rng.random((48, 3)), 6, octants_closed_form)
    tree_fixture("tree_mixed300.npz", mixed(rng, 300), 4)
    pts = mixed(rng, 2000)
    tau = rng.standard_normal(2000)
    tgt = uniform(rng, 200)
    u_src = O.direct_self(pts, tau)
    u_tgt = O.direct(pts, tau, tgt)
    g_tgt = O.direct(pts, tau, tgt, grad=True)
    # the O(N^2) sums once more in numpy: the fixture holds what both agree on
    d = tgt[:, None, :] - pts[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    assert np.allclose(u_tgt, (1 / (4 * np.pi * r)) @ tau, rtol=1e-12)
    np.savez_compressed(os.path.join(HERE, "fmm_mixed2000.npz"), points=pts, tau=tau, targets=tgt, u_sources=u_src,
                        u_targets=u_tgt, g_targets=g_tgt)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
