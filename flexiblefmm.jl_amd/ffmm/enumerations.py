"""Slot, offset and far-index orders shared by the tree, the lists and the operators.

Every function here fixes an ordering that the device tree, the interaction lists
and the translation operators must agree on, and each follows the reference's
enumeration exactly (with 0-based indices instead of Julia's 1-based ones):

* child slots: bit ``d`` of the slot is set when coordinate ``d`` lies strictly above
  the box centre (coordinate 0 least significant) -- src/util.jl:25-45;
* colleague offsets: ``(-1:1)^3`` with coordinate 0 fastest, self at slot 13 --
  src/util.jl:53-54;
* far offsets: the offsets of ``(-3:3)^3`` (coordinate 0 fastest) with Chebyshev norm
  at least 2, in table order: 316 of them -- src/util.jl:64-97;
* Morton keys: bit ``b`` of cell coordinate ``k`` at position ``3b + k`` --
  src/util.jl:132-138.
"""
from __future__ import annotations

import numpy as np

D = 3
NCHILD = 8
NCOLL = 27
SELF_SLOT = 13
NFAR = 316


def child_offsets() -> np.ndarray:
    """(8, 3) lattice offsets of the children in slot order (src/util.jl:25-28)."""
    k = np.arange(NCHILD)
    return np.stack([(k >> d) & 1 for d in range(D)], axis=1).astype(np.int64)


def _offset_table(r: int) -> np.ndarray:
    b = 2 * r + 1
    i = np.arange(b ** D)
    return np.stack([(i // b ** d) % b - r for d in range(D)], axis=1).astype(np.int64)


def colleague_offsets() -> np.ndarray:
    """(27, 3) offsets in (-1:1)^3, coordinate 0 fastest (src/util.jl:53)."""
    return _offset_table(1)


def m2l_table_offsets() -> np.ndarray:
    """(343, 3) offsets in (-3:3)^3, coordinate 0 fastest (src/util.jl:64)."""
    return _offset_table(3)


def m2l_table_index(off) -> int:
    """0-based position of an offset in the (-3:3)^3 table (src/util.jl:65-71)."""
    off = np.asarray(off)
    return int(sum((off[..., d] + 3) * 7 ** d for d in range(D)))


def is_far_offset(off) -> bool:
    return int(np.max(np.abs(off))) >= 2


def far_offsets() -> np.ndarray:
    """(316, 3) far offsets in table order (src/util.jl:79)."""
    t = m2l_table_offsets()
    return t[np.max(np.abs(t), axis=1) >= 2]


def far_index_table() -> np.ndarray:
    """Length 343: table index -> 0-based far index, or -1 for a near offset
    (src/util.jl:87-97)."""
    t = m2l_table_offsets()
    far = np.max(np.abs(t), axis=1) >= 2
    out = np.full(len(t), -1, dtype=np.int32)
    out[far] = np.arange(int(far.sum()), dtype=np.int32)
    return out


def morton_key(cell, nbits: int) -> int:
    """Interleaved bits of a cell: bit b of coordinate k at 3b + k (src/util.jl:132-138)."""
    key = 0
    for b in range(nbits):
        for k in range(D):
            key |= ((int(cell[k]) >> b) & 1) << (D * b + k)
    return key
