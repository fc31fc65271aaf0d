"""Host-side FP64 precompute of the cube kernel-independent operators.

This is set-up code, run once per operator cache on the host in FP64 with LAPACK
(numpy), exactly where the reference builds its operators: the device library only
receives the finished matrices (``ffmm_opcache_put_cube``).  It restates:

* surfaces: ``cube_boundary_points`` (src/operators/spectral_m2l.jl:103-115);
* the factored truncated-SVD pseudo-inverses ``pinv_rtol`` (src/operators/ki_common.jl:38-44);
* the folded M2M / L2L of ``_ki_dense_operators`` (src/operators/ki_common.jl:176-199)
  at the half-widths ``(1+d)w/2`` and ``(3-2d)w/2`` of ``level_operators``
  (src/operators/cubeki.jl:55-81);
* the symmetry classes of the 316 far offsets under the 48 signed permutations and
  the class representatives (src/operators/symmetry.jl:14-131);
* the spectral M2L kernel tables (src/operators/spectral_m2l.jl:119-135).

All matrices are produced at one reference width; for a homogeneous kernel
(Laplace3D, degree -1) every other width shares them through per-level scalar
factors, as ``rescale`` does (src/operators/ki_common.jl:213-222).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import enumerations as en

FOUR_PI_INV = 1.0 / (4.0 * np.pi)


# ---------------------------------------------------------------------------
# kernels (src/kernels.jl:179-200)
# ---------------------------------------------------------------------------
def laplace3d_matrix(sources: np.ndarray, targets: np.ndarray) -> np.ndarray:
    """``A[j, i] = 1 / (4 pi |t_j - s_i|)``: rows are targets (src/kernels.jl:404-417)."""
    d = targets[:, None, :] - sources[None, :, :]
    r2 = np.einsum("jik,jik->ji", d, d)
    return FOUR_PI_INV / np.sqrt(r2)


# ---------------------------------------------------------------------------
# surfaces
# ---------------------------------------------------------------------------
def cube_boundary_points(p: int, a: float):
    """Boundary points of the p^3 lattice spanning [-a, a]^3 (coordinate 0 fastest)
    and the 0-based linear index of each in the zero-padded (2p)^3 lattice
    (src/operators/spectral_m2l.jl:103-115)."""
    if p < 2:
        raise ValueError("p must be at least 2")
    coords = np.array([-a + 2 * a * float(i) / (p - 1) for i in range(p)])
    pts, emb, lat = [], [], []
    n = 2 * p
    for i3 in range(p):
        for i2 in range(p):
            for i1 in range(p):
                I = (i1, i2, i3)
                if not any(i == 0 or i == p - 1 for i in I):
                    continue
                pts.append((coords[i1], coords[i2], coords[i3]))
                emb.append(i1 + n * i2 + n * n * i3)
                lat.append(I)
    return np.array(pts, dtype=np.float64), np.array(emb, dtype=np.int64), np.array(lat, dtype=np.int64)


# ---------------------------------------------------------------------------
# factored pseudo-inverse (src/operators/ki_common.jl:25-44)
# ---------------------------------------------------------------------------
@dataclass
class FactoredPinv:
    Uh: np.ndarray      # r x m
    invS: np.ndarray    # r
    V: np.ndarray       # n x r
    sigma: np.ndarray   # all singular values

    @property
    def rank(self) -> int:
        return len(self.invS)

    def apply(self, x: np.ndarray, alpha: float = 1.0) -> np.ndarray:
        """``alpha * V (invS .* (Uh x))``, the factors in sequence
        (src/operators/ki_common.jl:52-68)."""
        return alpha * (self.V @ (self.invS[:, None] * (self.Uh @ x))) if x.ndim == 2 else \
            alpha * (self.V @ (self.invS * (self.Uh @ x)))


def pinv_rtol(A: np.ndarray, rtol: float) -> FactoredPinv:
    """Truncated-SVD pseudo-inverse keeping singular values >= rtol * sigma_1."""
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    r = int(np.count_nonzero(s >= rtol * s[0]))
    return FactoredPinv(Uh=np.ascontiguousarray(U[:, :r].T), invS=1.0 / s[:r],
                        V=np.ascontiguousarray(Vt[:r, :].T), sigma=s)


# ---------------------------------------------------------------------------
# symmetry classes (src/operators/symmetry.jl:14-131)
# ---------------------------------------------------------------------------
def hyperoctahedral_group():
    """The 48 signed permutations ``(R x)[k] = sign[k] x[perm[k]]``, in the reference's
    enumeration: permutations outer (lexicographic), sign patterns inner with the
    first axis fastest (src/operators/symmetry.jl:20-34)."""
    out = []
    for perm in itertools.permutations(range(3)):
        for s3 in (1, -1):
            for s2 in (1, -1):
                for s1 in (1, -1):
                    out.append((np.array(perm), np.array([s1, s2, s3])))
    return out


def canonical_offset(o) -> np.ndarray:
    return np.sort(np.abs(np.asarray(o)))[::-1].copy()


@dataclass
class SymmetryClasses:
    reps: np.ndarray      # (nclass, 3)
    klass: np.ndarray     # (316,) 0-based class of each far offset
    idx: np.ndarray       # (316, N) 0-based: (P x)[idx[i]] = sgn[i] x[i]
    sgn: np.ndarray       # (316, N) int8


def symmetry_classes(p: int, lattice: np.ndarray) -> SymmetryClasses:
    """Classes of the far offsets and, per offset, the signed permutation of the
    surface points that maps the class representative's operator onto the offset's:
    ``M2L_o m = P (A_rep (P^T m))`` (src/operators/symmetry.jl:68-101)."""
    offs = en.far_offsets()
    G = hyperoctahedral_group()
    slot = {tuple(m): i for i, m in enumerate(lattice)}
    reps, klass, idxs, sgns = [], [], [], []
    for o in offs:
        r = canonical_offset(o)
        c = next((k for k, q in enumerate(reps) if np.array_equal(q, r)), None)
        if c is None:
            reps.append(r)
            c = len(reps) - 1
        klass.append(c)
        perm, sign = next((pm, sg) for pm, sg in G if np.array_equal(sg * r[pm], o))
        rho = np.empty(len(lattice), dtype=np.int64)
        for i, m in enumerate(lattice):
            mp = tuple(int(m[perm[k]]) if sign[k] == 1 else p - 1 - int(m[perm[k]]) for k in range(3))
            rho[i] = slot[mp]
        idxs.append(rho)
        sgns.append(np.ones(len(lattice), dtype=np.int8))
    return SymmetryClasses(np.array(reps), np.array(klass, dtype=np.int32),
                           np.array(idxs, dtype=np.int32), np.array(sgns, dtype=np.int8))


# ---------------------------------------------------------------------------
# the operators of one width
# ---------------------------------------------------------------------------
@dataclass
class CubeOperators:
    """The reference-width operators of ``CubeKernelIndependent{3}(Laplace3D(); p, pc, d)``."""
    width: float
    p: int
    pc: int
    d: float
    small: np.ndarray          # (Ns, 3) upward equivalent / downward check surface
    large: np.ndarray          # (Nl, 3) upward check / downward equivalent surface
    embed: np.ndarray          # (Ns,) padded-lattice index of each small-surface point
    lattice: np.ndarray        # (Ns, 3) lattice index of each small-surface point
    h: float                   # small-surface lattice spacing
    P_up: FactoredPinv         # pinv of K(small -> large), Nl x Ns
    P_down: FactoredPinv       # pinv of K(large -> small), Ns x Nl
    M2M: np.ndarray            # (Ns, 8 Ns)
    L2L: np.ndarray            # (8 Ns, Ns)
    sc: SymmetryClasses = None
    class_ops: np.ndarray = None   # (nclass, Ns, Ns) class representatives
    tables: np.ndarray = None      # (316, 2p, 2p, p+1) complex half spectra, [i3, i2, i1]

    @property
    def nsmall(self) -> int:
        return len(self.small)

    @property
    def nlarge(self) -> int:
        return len(self.large)


def level_operators(w: float, p: int = 8, pc: int | None = None, d: float = 0.1,
                    rtol: float = 50 * np.finfo(np.float64).eps, with_classes: bool = True,
                    with_tables: bool = True) -> CubeOperators:
    """The cube KI operators of width ``w`` (src/operators/cubeki.jl:55-81 with
    src/operators/ki_common.jl:176-199)."""
    pc = 2 * p if pc is None else pc
    if pc < p:
        raise ValueError("pc must be at least p")
    if not (0 < d < 2 / 3):
        raise ValueError("d must lie in (0, 2/3)")
    a_small = (1 + d) * w / 2
    a_large = (3 - 2 * d) * w / 2
    small, embed, lattice = cube_boundary_points(p, a_small)
    large, _, _ = cube_boundary_points(pc, a_large)
    child_small, _, _ = cube_boundary_points(p, a_small / 2)
    P_up = pinv_rtol(laplace3d_matrix(small, large), rtol)
    P_down = pinv_rtol(laplace3d_matrix(large, small), rtol)
    centers = (en.child_offsets().astype(np.float64) - 0.5) * (w / 2)
    M2Mc = np.hstack([laplace3d_matrix(child_small + c, large) for c in centers])
    L2Lc = np.vstack([laplace3d_matrix(large, child_small + c) for c in centers])
    M2M = P_up.V @ (P_up.invS[:, None] * (P_up.Uh @ M2Mc))
    L2L = ((L2Lc @ P_down.V) * P_down.invS[None, :]) @ P_down.Uh
    h = 2 * a_small / (p - 1)
    ops = CubeOperators(width=w, p=p, pc=pc, d=d, small=small, large=large, embed=embed,
                        lattice=lattice, h=h, P_up=P_up, P_down=P_down, M2M=M2M, L2L=L2L)
    if with_classes:
        ops.sc = symmetry_classes(p, lattice)
        ops.class_ops = np.array([laplace3d_matrix(small + r * w, small) for r in ops.sc.reps])
    if with_tables:
        ops.tables = np.array([kernel_table_half(p, h, o, w) for o in en.far_offsets()])
    return ops


def class_factors(ops: CubeOperators, tol: float):
    """``class_factors`` (src/operators/symmetry.jl:169-178): each class representative
    as ``A·Bt`` from its SVD, keeping the singular values above ``tol`` times the
    largest (at least one); returns (ranks, [A_c], [Bt_c])."""
    ranks, A, Bt = [], [], []
    for M in ops.class_ops:
        U, S, Vt = np.linalg.svd(M)
        k = max(1, int(np.count_nonzero(S > tol * S[0])))
        ranks.append(k)
        A.append(U[:, :k] * S[:k])
        Bt.append(Vt[:k])
    return np.array(ranks), A, Bt


def m2l_dense(ops: CubeOperators, f: int) -> np.ndarray:
    """The dense M2L matrix of far offset ``f`` (multipole -> check data) at the
    operators' width (src/operators/symmetry.jl:321-325)."""
    o = en.far_offsets()[f].astype(np.float64) * ops.width
    return laplace3d_matrix(ops.small + o, ops.small)


def kernel_table_half(p: int, h: float, off, w: float) -> np.ndarray:
    """Half spectrum of the circulant kernel table of one far offset: entry
    ``m in [-(p-1), p-1]^3`` of ``K(off w, h m)`` at the wrapped index, real FFT over
    the three axes; layout ``[i3, i2, i1]`` with ``i1`` halved to ``p+1``
    (src/operators/spectral_m2l.jl:119-135 and :42-50)."""
    n = 2 * p
    tab = np.zeros((n, n, n))
    m = np.arange(-(p - 1), p)
    m3, m2, m1 = np.meshgrid(m, m, m, indexing="ij")
    src = np.asarray(off, dtype=np.float64) * w
    dx = h * m1 - src[0]
    dy = h * m2 - src[1]
    dz = h * m3 - src[2]
    tab[m3 % n, m2 % n, m1 % n] = FOUR_PI_INV / np.sqrt(dx * dx + dy * dy + dz * dz)
    return np.fft.rfftn(tab)


def spectral_m2l_apply(ops: CubeOperators, f: int, x: np.ndarray, alpha: float = 1.0) -> np.ndarray:
    """Single-vector spectral M2L (embed, transform, multiply, transform back,
    extract) of far offset ``f`` (src/operators/spectral_m2l.jl:144-185)."""
    n = 2 * ops.p
    lat = np.zeros(n ** 3)
    lat[ops.embed] = x
    u = np.fft.irfftn(np.fft.rfftn(lat.reshape(n, n, n)) * ops.tables[f], s=(n, n, n), axes=(0, 1, 2))
    return alpha * u.reshape(-1)[ops.embed]


# ---------------------------------------------------------------------------
# scale factors of a homogeneous kernel (src/operators/ki_common.jl:213-222)
# ---------------------------------------------------------------------------
HOMOGENEITY = -1


def level_scales(lam: float):
    """For width ``lam * w_ref``: (surface scale, pinv_scale, m2l_scale).  M2M and L2L
    carry their pseudo-inverses and are width invariant."""
    f = lam ** HOMOGENEITY
    return lam, 1.0 / f, f
