"""Plan data (the payload of ffmm_plan_save) written on the host from a tree of the
oracle's builder, field by field in the library's layout, so that the host checks of
ffmm_plan_data_check can be exercised without a device: the header (magic, version,
the cache's identity, the plan's parameters and sizes), the per-level node offsets and
node-list offsets, then each array as an int64 count and its elements (cells and the
27 colleagues stored structure-of-arrays, the sorted points as three coordinate
arrays)."""
import struct

import numpy as np

MAGIC = b"FFMMTREE"
VERSION = 1


def fields(tree):
    """The plan's arrays from an oracle ``Tree`` (0-based), in the library's layout."""
    n = tree.nnodes
    L = tree.nlevels
    ls = np.asarray(tree.level_start, dtype=np.int64)
    childless = tree.child_first < 0
    s2m = np.nonzero(childless & (tree.has_source > 0))[0].astype(np.int32)
    leaves = np.nonzero(tree.leaf > 0)[0].astype(np.int32)
    s2m_off = np.searchsorted(s2m, ls).astype(np.int64)
    leaf_off = np.searchsorted(leaves, ls).astype(np.int64)
    w = np.ldexp(tree.W0, -tree.level.astype(np.int64))
    centers = tree.origin[None, :] + (tree.cell + 0.5) * w[:, None]
    return dict(
        lo=ls[:L].copy(), nl=np.diff(ls), s2m_off=s2m_off, leaf_off=leaf_off,
        cell=np.ascontiguousarray(tree.cell.T, dtype=np.int32).ravel(),
        level=tree.level.astype(np.int32), parent=tree.parent.astype(np.int32),
        child_first=tree.child_first.astype(np.int32), plo=tree.plo.astype(np.int32), phi=tree.phi.astype(np.int32),
        col=np.ascontiguousarray(tree.colleagues.T, dtype=np.int32).ravel(),
        perm=tree.perm.astype(np.int32), point_node=tree.point_node.astype(np.int32),
        leaf=tree.leaf.astype(np.uint8), ghost=tree.ghost.astype(np.uint8), xlist=tree.xlist.astype(np.uint8),
        has_source=tree.has_source.astype(np.uint8),
        sx=np.ascontiguousarray(tree.points[:, 0]), sy=np.ascontiguousarray(tree.points[:, 1]),
        sz=np.ascontiguousarray(tree.points[:, 2]),
        centers=np.ascontiguousarray(centers.T).ravel(), s2m_nodes=s2m, leaf_nodes=leaves, n=n, L=L)


ORDER = ["lo", "nl", "s2m_off", "leaf_off", "cell", "level", "parent", "child_first", "plo", "phi", "col", "perm",
         "point_node", "leaf", "ghost", "xlist", "has_source", "sx", "sy", "sz", "centers", "s2m_nodes", "leaf_nodes"]


def encode(tree, f=None, *, ppl=64, p=8, w_ref=1.0, max_levels=30, precision=0, m2l=0, L=None, nnodes=None,
           N=None) -> bytes:
    """Plan data of a Laplace cube p = 8 plan (Ns = 296, Nl = 1352) over ``tree``; ``f``
    (from ``fields``, possibly altered) and the size fields may be overridden."""
    f = fields(tree) if f is None else f
    N = len(tree.perm) if N is None else N
    head = MAGIC + struct.pack("<I6i2dq4i3ddiiq", VERSION, 0, 0, p, 1, 296, 1352, w_ref, 0.0, N, ppl, max_levels,
                               precision, m2l, *[float(v) for v in tree.origin], float(tree.W0), 0,
                               f["L"] if L is None else L, f["n"] if nnodes is None else nnodes)
    parts = [head]
    for k in ORDER:
        a = np.ascontiguousarray(f[k])
        parts.append(struct.pack("<q", len(a)))
        parts.append(a.tobytes())
    return b"".join(parts)
