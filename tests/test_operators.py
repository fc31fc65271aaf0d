"""The host FP64 operator precompute, pinned by what the mathematics fixes: the
translated fields against the exact (direct) fields on all 316 far offsets, M2M
against the parent's own S2M and the exact field, L2L against the exact far field,
the symmetry classes against the dense per-offset operators, the spectral form
against the dense one.  Restates the reference's operator checks
(test/operator_checks.jl:51-126; test/test_3d.jl:119-128; test/test_symmetry.jl:9-48;
test/test_cubeki.jl:43-68)."""
import numpy as np
import pytest

from ffmm import enumerations as en
from ffmm import operators as op

K = op.laplace3d_matrix


def mrd(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def box_points(rng, c, w, n):
    return np.asarray(c) + w * (rng.random((n, 3)) - 0.5)


def multipole_of(ops, xs, tau, c):
    pM = K(xs - c, ops.large) @ tau              # check potentials (s2m!)
    return ops.P_up.apply(pM)                     # partial_to_multipole!


def l2t(ops, L, x):
    return K(ops.large, x) @ L


@pytest.fixture(scope="module")
def ops37():
    return op.level_operators(0.37, p=8)


def test_sizes_and_enumerations():
    assert len(en.far_offsets()) == 316 == 7 ** 3 - 3 ** 3
    fi = en.far_index_table()
    assert np.count_nonzero(fi >= 0) == 316 and sorted(fi[fi >= 0]) == list(range(316))
    s, emb, lat = op.cube_boundary_points(8, 0.55)
    assert len(s) == 6 * 64 - 12 * 8 + 8 == 296
    assert len(np.unique(emb)) == 296 and np.all((emb >= 0) & (emb < 16 ** 3))
    assert np.allclose(np.max(np.abs(s), axis=1), 0.55)
    assert s[1, 0] > s[0, 0]                     # coordinate 0 fastest
    L, _, _ = op.cube_boundary_points(16, 1.0)
    assert len(L) == 16 ** 3 - 14 ** 3 == 1352
    assert en.morton_key([1, 0, 0], 1) == 1 and en.morton_key([0, 0, 1], 1) == 4 and en.morton_key([2, 0, 0], 2) == 8


def test_m2l_all_offsets(ops37):
    """Field error of S2M -> UC2E -> M2L -> DC2E -> L2T on every far offset < 2e-6, the
    nearest offsets the worst (test/operator_checks.jl:51-74, test/test_3d.jl:119-128)."""
    rng = np.random.default_rng(17)
    w = ops37.width
    errs = []
    for f, off in enumerate(en.far_offsets()):
        cs = off * w
        xs = box_points(rng, cs, w, 40)
        tau = rng.standard_normal(40)
        ts = box_points(rng, np.zeros(3), w, 40)
        M = multipole_of(ops37, xs, tau, cs)
        pL = op.m2l_dense(ops37, f) @ M
        L = ops37.P_down.apply(pL)
        errs.append(mrd(l2t(ops37, L, ts), K(xs, ts) @ tau))
    errs = np.array(errs)
    nearest = np.max(np.abs(en.far_offsets()), axis=1) == 2
    assert errs.max() < 2e-6
    assert errs[nearest].max() >= errs[~nearest].max() / 10


def test_m2m_against_parent_s2m_and_exact():
    """test/operator_checks.jl:76-102: children multipoles through M2M give the
    parent's field; so does the parent's own S2M."""
    rng = np.random.default_rng(3)
    w = 0.61
    ops, cops = op.level_operators(w, p=8, with_tables=False), op.level_operators(w / 2, p=8, with_tables=False)
    centers = (en.child_offsets() - 0.5) * (w / 2)
    stacked, allx, allt = [], [], []
    for c in centers:
        xs = box_points(rng, c, w / 2, 25)
        tau = rng.standard_normal(25)
        stacked.append(multipole_of(cops, xs, tau, c))
        allx.append(xs)
        allt.append(tau)
    allx, allt = np.vstack(allx), np.concatenate(allt)
    Mp = ops.M2M @ np.concatenate(stacked)
    Md = multipole_of(ops, allx, allt, np.zeros(3))
    v = rng.standard_normal((50, 3))
    ts = 2.5 * w * v / np.linalg.norm(v, axis=1, keepdims=True)
    uex = K(allx, ts) @ allt
    assert mrd(K(ops.small, ts) @ Mp, uex) < 2e-6
    assert mrd(K(ops.small, ts) @ Md, uex) < 2e-6


def test_l2l_against_exact():
    """test/operator_checks.jl:104-126: a parent's partial local through L2L gives the
    children's fields."""
    rng = np.random.default_rng(5)
    w = 0.61
    ops, cops = op.level_operators(w, p=8, with_tables=False), op.level_operators(w / 2, p=8, with_tables=False)
    off = np.array([3, -2, 1])
    f = next(i for i, o in enumerate(en.far_offsets()) if np.array_equal(o, off))
    cs = off * w
    xs = box_points(rng, cs, w, 40)
    tau = rng.standard_normal(40)
    pLp = op.m2l_dense(ops, f) @ multipole_of(ops, xs, tau, cs)
    Lp = ops.P_down.apply(pLp)
    pLc = ops.L2L @ pLp
    Ns = ops.nsmall
    for k, c in enumerate((en.child_offsets() - 0.5) * (w / 2)):
        ts = box_points(rng, c, w / 2, 20)
        uex = K(xs, ts) @ tau
        Lk = cops.P_down.apply(pLc[k * Ns:(k + 1) * Ns])
        assert mrd(l2t(cops, Lk, ts - c), uex) < 2e-6
        assert mrd(l2t(ops, Lp, ts), uex) < 2e-6


def test_symmetry_classes_reproduce_dense_operators():
    """16 classes, orbit sizes [48, 24 x 9, 12, 12, 8, 8, 6, 6], every far offset's
    dense operator = P A_rep P^T to 1e-13 (test/test_symmetry.jl:9-48,
    src/operators/symmetry.jl:138-158)."""
    ops = op.level_operators(0.7, p=8, with_tables=False)
    sc = ops.sc
    assert len(sc.reps) == 16
    assert sorted(np.bincount(sc.klass)) == sorted([48] + [24] * 9 + [12, 12, 8, 8, 6, 6])
    worst = 0.0
    for f in range(316):
        A = op.m2l_dense(ops, f)
        B = np.empty_like(A)
        B[np.ix_(sc.idx[f], sc.idx[f])] = ops.class_ops[sc.klass[f]]
        worst = max(worst, np.max(np.abs(A - B)) / np.max(np.abs(A)))
        assert np.array_equal(np.sort(sc.idx[f]), np.arange(ops.nsmall))
    assert worst < 1e-13
    rng = np.random.default_rng(1)
    v = rng.standard_normal(ops.nsmall)
    for f in (0, 16, 99, 315):
        Pv = np.empty_like(v)
        Pv[sc.idx[f]] = sc.sgn[f] * v
        assert np.allclose(sc.sgn[f] * Pv[sc.idx[f]], v)


def test_spectral_equals_dense(ops37):
    """test/test_cubeki.jl:43-68: the spectral translation equals the dense kernel
    matrix on every far offset to rtol 1e-12."""
    rng = np.random.default_rng(2)
    for f in range(316):
        x = rng.standard_normal(ops37.nsmall)
        a = op.m2l_dense(ops37, f) @ x
        b = op.spectral_m2l_apply(ops37, f, x, 2.0)
        assert np.max(np.abs(b - 2.0 * a)) <= 1e-12 * np.max(np.abs(a))


def test_homogeneous_rescale():
    """Laplace3D is homogeneous of degree -1: M2M and L2L (which carry their
    pseudo-inverses) are width invariant, the pseudo-inverses scale by lambda and M2L
    by 1/lambda (src/operators/ki_common.jl:213-222)."""
    a, b = op.level_operators(1.0, p=8, with_tables=False), op.level_operators(0.25, p=8, with_tables=False)
    lam = 0.25
    assert mrd(b.M2M, a.M2M) < 1e-10 and mrd(b.L2L, a.L2L) < 1e-10
    rng = np.random.default_rng(8)
    x = rng.standard_normal(a.nlarge)
    assert mrd(b.P_up.apply(x), lam * a.P_up.apply(x)) < 1e-8
    assert mrd(op.m2l_dense(b, 7), op.m2l_dense(a, 7) / lam) < 1e-14
    assert b.P_up.rank == a.P_up.rank and b.P_down.rank == a.P_down.rank


def test_class_factors(ops8):
    """Low-rank class factors (class_factors, src/operators/symmetry.jl:169-178): each
    representative is reproduced to tol in the spectral norm, the ranks lie well below
    n = 296 (test/test_device_passes.jl:44: below n/2 at 1e-8), and the smallest ranks
    are those the reference's docstring gives for Laplace at p = 8 ("13 and 26" at 1e-6
    and 1e-8; here 13 and 24)."""
    from ffmm import operators as op
    for tol, smallest in ((1e-6, 13), (1e-8, 24)):
        ranks, A, Bt = op.class_factors(ops8, tol)
        assert len(ranks) == len(ops8.class_ops) == 16
        assert ranks.min() == smallest and ranks.max() < ops8.nsmall // 2
        for a, b, M in zip(A, Bt, ops8.class_ops):
            assert np.linalg.norm(a @ b - M, 2) <= tol * np.linalg.norm(M, 2)
    r6, _, _ = op.class_factors(ops8, 1e-6)
    r8, _, _ = op.class_factors(ops8, 1e-8)
    assert np.all(r6 <= r8)
