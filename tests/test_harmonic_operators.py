"""The harmonic methods' host operators (ffmm.harmonics) and the harmonic oracle
(oracle/harmonic.py, oracle/harmonic_oracle.c), pinned against what the mathematics
fixes: the Laplace addition theorem, the Hobson-Schmidt relation, exact M2M/L2L
translations, M2L against direct sums on every far offset, the class permutations
against the per-offset operators, the Yukawa radial functions against scipy's
modified spherical Bessel functions, and the whole oracle FMM against FP64 direct
sums (the reference ships no golden vectors for these methods)."""
import math

import numpy as np
import pytest

from ffmm import enumerations as en
from ffmm import harmonics as H

import harmonic as HO
from pointsets import mixed, uniform

P = 10


@pytest.fixture(scope="module")
def lap():
    return H.laplace_level_operators(1.0, P)


def s2m_np(ops, pts, tau, c, w, kh=0.0):
    """S2M from the vectorised fold (independent of the C oracle)."""
    p, N, tb = ops.p, ops.N, ops.tb
    X = (np.asarray(pts) - c) / w
    R, I = H.fold_regular(X, p, tb)
    U = H.yukawa_u_many((kh * np.linalg.norm(X, axis=1)) ** 2, p + 1) if kh else np.ones((p + 2, len(X)))
    col = np.zeros(N)
    for n in range(p + 1):
        q = tau * U[n]
        col[n * n] += R[H.tidx(n, 0)] @ q
        for m in range(1, n + 1):
            col[n * n + 2 * m - 1] += R[H.tidx(n, m)] @ q
            col[n * n + 2 * m] -= I[H.tidx(n, m)] @ q
    return col


def l2t_np(ops, L, x, c, w, kh=0.0):
    p, tb = ops.p, ops.tb
    X = (np.asarray(x) - c) / w
    R, I = H.fold_regular(X, p, tb)
    U = H.yukawa_u_many((kh * np.linalg.norm(X, axis=1)) ** 2, p) if kh else np.ones((p + 1, len(X)))
    acc = np.zeros(len(X))
    for n in range(p + 1):
        a = L[n * n] * R[H.tidx(n, 0)]
        for m in range(1, n + 1):
            a = a + 2 * (L[n * n + 2 * m - 1] * R[H.tidx(n, m)] + L[n * n + 2 * m] * I[H.tidx(n, m)])
        acc += U[n] * a
    return acc


def test_addition_theorem_pins_hobson_harmonics():
    # 1/|x - y| = sum_{n,m} conj(Y_n^m(y)) Theta_n^m(x) for |y| < |x| (the identity of
    # src/operators/harmonics.jl:7); truncated at degree 40 with |y| / |x| = 0.3
    rng = np.random.default_rng(0)
    p = 40
    for _ in range(5):
        x = rng.standard_normal(3)
        x *= 1.3 / np.linalg.norm(x)
        y = rng.standard_normal(3)
        y *= 0.39 / np.linalg.norm(y)
        s = np.sum(np.conj(H.regular_hobson(y, p)) * H.irregular_hobson(x, p))
        assert abs(s - 1 / np.linalg.norm(x - y)) < 1e-14 * (1 / np.linalg.norm(x - y))


def test_schmidt_fold_equals_scaled_hobson():
    # the recurrence tables generate s_n^m Y_n^m (src/operators/harmonics.jl:142-187)
    p = 12
    tb = H.harmonic_tables(p)
    s = H.scaling_factors(p)
    X = np.random.default_rng(1).uniform(-1, 1, (7, 3))
    R, I = H.fold_regular(X, p, tb)
    for q, x in enumerate(X):
        Y = H.regular_hobson(x, p)
        for n in range(p + 1):
            for m in range(n + 1):
                v = s[H.cidx(n, m)] * Y[H.cidx(n, m)]
                assert abs(complex(R[H.tidx(n, m), q], I[H.tidx(n, m), q]) - v) < 1e-13 * max(1, abs(v))


def test_real_operator_round_trip():
    rng = np.random.default_rng(2)
    p = 5
    N = (p + 1) ** 2
    v = rng.standard_normal(N)
    c = H.complex_coefficients(v, p)
    assert np.allclose(H.real_coefficients(c, p), v)
    # an operator commuting with conjugate symmetry: the M2M of a child shift
    A = np.zeros((N, N), dtype=complex)
    Y = H.regular_hobson(np.array([0.25, -0.25, 0.25]), p)
    for n1 in range(p + 1):
        for m1 in range(-n1, n1 + 1):
            for n in range(n1, p + 1):
                d = n - n1
                for mm in range(max(-n, m1 - d), min(n, m1 + d) + 1):
                    A[H.cidx(n, mm), H.cidx(n1, m1)] = np.conj(Y[H.cidx(d, mm - m1)])
    R = H.real_operator(A, p)
    assert np.allclose(R @ v, H.real_coefficients(A @ c, p), atol=1e-13)


def test_m2m_l2l_exact(lap):
    rng = np.random.default_rng(3)
    N = lap.N
    Mp = np.zeros(N)
    S, T = [], []
    for j, d in enumerate(en.child_offsets()):
        cc = (np.asarray(d) - 0.5) / 2
        s_ = cc + rng.uniform(-0.25, 0.25, (15, 3))
        t_ = rng.standard_normal(15)
        Mp += lap.M2M[:, j * N:(j + 1) * N] @ s2m_np(lap, s_, t_, cc, 0.5)
        S.append(s_)
        T.append(t_)
    Md = s2m_np(lap, np.vstack(S), np.concatenate(T), np.zeros(3), 1.0)
    assert np.max(np.abs(Mp - Md)) < 1e-13 * np.max(np.abs(Md))
    L = rng.standard_normal(N)
    for j, d in enumerate(en.child_offsets()):
        cc = (np.asarray(d) - 0.5) / 2
        xs = cc + rng.uniform(-0.25, 0.25, (6, 3))
        e1 = l2t_np(lap, L, xs, np.zeros(3), 1.0)
        e2 = l2t_np(lap, lap.L2L[j * N:(j + 1) * N] @ L, xs, cc, 0.5)
        assert np.max(np.abs(e1 - e2)) < 1e-12 * np.max(np.abs(e1))


def test_m2l_classes_reproduce_every_offset(lap):
    # check_harmonic_classes (src/operators/harmonics.jl:645-669): P_t A_rep P_s^T is the
    # operator built directly for each of the 316 offsets
    s = H.scaling_factors(2 * P)
    assert len(lap.reps) == 34
    for f, o in enumerate(en.far_offsets()):
        M = H.harmonic_m2l(o, P, H.C_LAPLACE, s)
        assert np.max(np.abs(M - lap.m2l(f))) <= 1e-13 * np.max(np.abs(M)), f


def test_m2l_against_direct_sums(lap):
    rng = np.random.default_rng(4)
    src = rng.uniform(-0.5, 0.5, (40, 3))
    tau = rng.standard_normal(40)
    M = s2m_np(lap, src, tau, np.zeros(3), 1.0)
    errs = []
    for f, o in enumerate(en.far_offsets()):
        ct = -np.asarray(o, float)
        xs = ct + rng.uniform(-0.5, 0.5, (6, 3))
        ex = (H.C_LAPLACE / np.linalg.norm(xs[:, None] - src[None], axis=2)) @ tau
        errs.append(np.max(np.abs(l2t_np(lap, lap.m2l(f) @ M, xs, ct, 1.0) - ex)) / np.max(np.abs(ex)))
    assert np.median(errs) < 1e-6 and max(errs) < 1e-3


def test_yukawa_radial_functions_against_scipy():
    sp = pytest.importorskip("scipy.special")

    def dfact(n):
        return math.prod(range(n, 0, -2)) if n > 0 else 1

    for z in [1e-4, 0.3, 2.0, 7.5, 25.0]:
        u, v = H.yukawa_u(z * z, 12), H.yukawa_v(z, 12)
        for n in range(13):
            ue = dfact(2 * n + 1) * sp.spherical_in(n, z) / z ** n
            ve = (2 / math.pi) * z ** (n + 1) * sp.spherical_kn(n, z) / dfact(2 * n - 1)
            assert abs(u[n] / ue - 1) < 1e-13 and abs(v[n] / ve - 1) < 1e-13
    z = np.array([0.0, 1e-5, 0.3, 2.0, 40.0])
    U, V = H.yukawa_u_many(z ** 2, 8), H.yukawa_v_many(z, 8)
    for i, zz in enumerate(z):
        assert np.allclose(U[:, i], H.yukawa_u(zz * zz, 8), rtol=1e-15, atol=0)
        assert np.allclose(V[:, i], H.yukawa_v(zz, 8), rtol=1e-15, atol=0)


@pytest.mark.parametrize("w", [0.25, 1 / 16])
def test_yukawa_operators_against_direct_sums(w):
    kappa = 5.0
    ops = H.yukawa_level_operators(w, P, kappa)
    kh = ops.kappa_hat
    rng = np.random.default_rng(5)
    src = rng.uniform(-0.5, 0.5, (30, 3)) * w
    tau = rng.standard_normal(30)
    M = s2m_np(ops, src, tau, np.zeros(3), w, kh)
    K = lambda d: H.C_LAPLACE * np.exp(-kappa * d) / d   # noqa: E731
    errs = []
    for f, o in enumerate(en.far_offsets()):
        ct = -np.asarray(o, float) * w
        xs = ct + rng.uniform(-0.5, 0.5, (5, 3)) * w
        ex = K(np.linalg.norm(xs[:, None] - src[None], axis=2)) @ tau
        errs.append(np.max(np.abs(l2t_np(ops, ops.m2l(f) @ M, xs, ct, w, kh) - ex)) / np.max(np.abs(ex)))
    assert np.median(errs) < 1e-5 and max(errs) < 1e-2
    # M2M by projection: the parent's multipole of the children's sources
    N = ops.N
    Mp = np.zeros(N)
    S, T = [], []
    for j, d in enumerate(en.child_offsets()):
        cc = (np.asarray(d) - 0.5) / 2 * w
        s_ = cc + rng.uniform(-0.25, 0.25, (10, 3)) * w
        t_ = rng.standard_normal(10)
        Mp += ops.M2M[:, j * N:(j + 1) * N] @ s2m_np(ops, s_, t_, cc, w / 2, kh / 2)
        S.append(s_)
        T.append(t_)
    Md = s2m_np(ops, np.vstack(S), np.concatenate(T), np.zeros(3), w, kh)
    assert np.max(np.abs(Mp - Md)) < 1e-6 * np.max(np.abs(Md))


def test_stokes_dense_operators_match_blockwise_form(lap):
    # _stokes_dense_operators against m2m_apply! / l2l_apply!
    # (src/operators/harmonics_stokes.jl:112-196)
    N = lap.N
    Md, Ld = H.stokes_dense_operators(lap.M2M, lap.L2L, N)
    rng = np.random.default_rng(6)
    B = rng.standard_normal(32 * N)       # 8 children x 4 blocks
    out = np.zeros(4 * N)
    for a in range(4):
        X = np.zeros(8 * N)
        for j, d in enumerate(en.child_offsets()):
            dh = (np.asarray(d, float) - 0.5) / 2
            r0 = j * 4 * N
            if a < 3:
                X[j * N:(j + 1) * N] = B[r0 + a * N:r0 + (a + 1) * N]
            else:
                X[j * N:(j + 1) * N] = (B[r0 + 3 * N:r0 + 4 * N] / 2 + dh[0] * B[r0:r0 + N] + dh[1] * B[r0 + N:r0 + 2 * N]
                                        + dh[2] * B[r0 + 2 * N:r0 + 3 * N])
        out[a * N:(a + 1) * N] = lap.M2M @ X
    assert np.allclose(Md @ B, out, rtol=1e-13, atol=1e-13)
    G = rng.standard_normal(4 * N)
    R = [lap.L2L @ G[a * N:(a + 1) * N] for a in range(4)]
    exp = np.zeros(32 * N)
    for j, d in enumerate(en.child_offsets()):
        dh = (np.asarray(d, float) - 0.5) / 2
        r0 = j * 4 * N
        r1, r2, r3, r4 = (R[a][j * N:(j + 1) * N] for a in range(4))
        exp[r0:r0 + N], exp[r0 + N:r0 + 2 * N], exp[r0 + 2 * N:r0 + 3 * N] = r1, r2, r3
        exp[r0 + 3 * N:r0 + 4 * N] = 2 * r4 - 2 * (dh[0] * r1 + dh[1] * r2 + dh[2] * r3)
    assert np.allclose(Ld @ G, exp, rtol=1e-13, atol=1e-13)


def test_oracle_point_loops_match_numpy_restatement(lap):
    """The C oracle's S2M and L2T (charges, Laplace and Yukawa) equal the numpy fold route."""
    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 1, (300, 3))
    tau = rng.standard_normal(300)
    for kind, kappa in (("laplace", 0.0), ("yukawa", 3.0)):
        ops = lap if kind == "laplace" else H.yukawa_level_operators(1.0, P, kappa)
        f = HO.HarmonicFMM(kind, lambda l, w: (ops, 1.0), pts, 400, np.zeros(3), 1.0, kappa=kappa)
        t = f.tree
        assert t.nnodes == 1
        ds = np.ascontiguousarray(tau[t.perm])
        M = np.empty((1, ops.N))
        HO.O.lib().oh_s2m(HO.KIND[kind], P, *f.tabs, np.zeros(1, np.int32), 1, f.centers, t.level, f.scales, f.khat,
                          t.plo, t.phi, t.points, ds, 1, M)
        ref = s2m_np(ops, t.points, ds, f.centers[0], 1.0, kappa)
        assert np.allclose(M[0], ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))
        L = rng.standard_normal(ops.N)
        x = rng.uniform(0.2, 0.8, (20, 3))
        out = np.zeros((20, 1))
        HO.O.lib().oh_l2t(HO.KIND[kind], P, *f.tabs, x, np.zeros(20, np.int32), 20, L[None, :].copy(), f.centers,
                          t.level, f.scales, f.khat, 0, out)
        assert np.allclose(out[:, 0], l2t_np(ops, L, x, f.centers[0], 1.0, kappa), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("kind", ["laplace", "yukawa", "stokes"])
def test_oracle_fmm_against_direct_sums(lap, kind):
    rng = np.random.default_rng(8)
    n = 2500
    pts = mixed(rng, n)
    tgt = uniform(rng, 150)
    kappa = 5.0 if kind == "yukawa" else 0.0
    cache = {}

    def level_ops(l, w):
        if kind == "yukawa":
            if w not in cache:
                cache[w] = H.yukawa_level_operators(w, P, kappa)
            return cache[w], 1.0
        return lap, 1.0 / w

    f = HO.HarmonicFMM(kind, level_ops, pts, 40, np.zeros(3), 1.0, kappa=kappa)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)   # noqa: E731
    if kind == "stokes":
        dens = np.hstack([rng.standard_normal((n, 6)), rng.standard_normal((n, 3))])
        dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
        u, v = f.matvec(dens, targets=tgt)
        assert rel(u, HO.direct("stokes", pts, dens, pts, skip=np.arange(n))) < 3e-4
        assert rel(v, HO.direct("stokes", pts, dens, tgt)) < 3e-4
    else:
        tau = rng.standard_normal(n)
        u, v = f.matvec(tau, output=2, targets=tgt)
        ud = HO.direct(kind, pts, tau, pts, 2, kappa=kappa, skip=np.arange(n))
        vd = HO.direct(kind, pts, tau, tgt, 2, kappa=kappa)
        assert rel(u[:, 0], ud[:, 0]) < 2e-5 and rel(v[:, 0], vd[:, 0]) < 2e-5
        assert rel(u[:, 1:], ud[:, 1:]) < 1e-4 and rel(v[:, 1:], vd[:, 1:]) < 1e-4


def test_oracle_dipole_kernel_is_the_charge_kernel_differenced():
    """The Laplace field of a dipole d at s (potential d . grad_s K, src/sources.jl:138;
    gradient -(hess_t K) d, src/sources.jl:151) against the central difference of two
    charges +-1/(2h) at s +- h d, at sources and separate targets, potential and gradient."""
    rng = np.random.default_rng(21)
    src = rng.random((40, 3))
    tgt = rng.random((30, 3)) + 1.5
    d = rng.standard_normal((40, 3))
    h = 1e-4
    dens = np.hstack([np.zeros((40, 1)), d])
    got = HO.direct("laplace", src, dens, tgt, 2)
    pair = np.vstack([src + h * d, src - h * d])
    q = np.concatenate([np.full(40, 1 / (2 * h)), np.full(40, -1 / (2 * h))])
    want = HO.direct("laplace", pair, q, tgt, 2)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-9 * np.max(np.abs(want)))
    # a charge and a dipole per point add: the (q, d) rows equal the two parts summed
    qd = rng.standard_normal(40)
    both = HO.direct("laplace", src, np.hstack([qd[:, None], d]), tgt, 2)
    assert np.allclose(both, got + HO.direct("laplace", src, qd, tgt, 2), rtol=1e-13, atol=1e-14)


def test_oracle_fmm_with_dipoles_against_direct_sums(lap):
    """Laplace charge + dipole densities through the harmonic oracle (S2M by the dipole
    accumulation of src/operators/harmonics.jl:711-752, the near field by the dipole
    kernel) against FP64 direct sums: the truncation error of p = 10."""
    rng = np.random.default_rng(22)
    n = 2500
    pts = mixed(rng, n)
    tgt = uniform(rng, 150)
    f = HO.HarmonicFMM("laplace", lambda l, w: (lap, 1.0 / w), pts, 40, np.zeros(3), 1.0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)   # noqa: E731
    dens = rng.standard_normal((n, 4))
    u, v = f.matvec(dens, output=2, targets=tgt)
    ud = HO.direct("laplace", pts, dens, pts, 2, skip=np.arange(n))
    vd = HO.direct("laplace", pts, dens, tgt, 2)
    assert rel(u[:, 0], ud[:, 0]) < 3e-5 and rel(v[:, 0], vd[:, 0]) < 3e-5
    assert rel(u[:, 1:], ud[:, 1:]) < 3e-5 and rel(v[:, 1:], vd[:, 1:]) < 3e-5


def test_oracle_stokes_pressure_is_the_laplace_dipole_field():
    """The Stokes pressure kernels against the (pinned) Laplace dipole field: a stokeslet
    f's pressure (f.r) / (4 pi r^3) (src/kernels.jl:267-271) is the potential of the
    dipole f, and a stresslet (g, nu)'s, ((g.nu) / r^3 - 3 (g.r)(nu.r) / r^5) / (2 pi)
    (src/sources.jl:237-242), is twice nu . the gradient of the dipole g."""
    rng = np.random.default_rng(23)
    src = rng.random((30, 3))
    tgt = rng.random((25, 3)) + 1.5
    f = rng.standard_normal((30, 3))
    got = HO.direct("stokes", src, f, tgt, 3)
    want = HO.direct("laplace", src, np.hstack([np.zeros((30, 1)), f]), tgt, 0)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    for j in range(5):
        g, nu = rng.standard_normal(3), rng.standard_normal(3)
        dens = np.concatenate([np.zeros(3), g, nu])[None, :]
        got = HO.direct("stokes", src[j:j + 1], dens, tgt, 3)
        grad = HO.direct("laplace", src[j:j + 1], np.concatenate([[0.0], g])[None, :], tgt, 1)
        assert np.allclose(got, 2 * grad @ nu, rtol=1e-12, atol=1e-14)


def test_oracle_stokes_pressure_fmm_against_direct_sums(lap):
    """The pressure of stokeslets + stresslets through the harmonic oracle (far field
    -s tr(grad Phi), _stokes_output(::Pressure), src/operators/harmonics_stokes.jl:348-351)
    against the FP64 direct sum, at sources and targets."""
    rng = np.random.default_rng(24)
    n = 2500
    pts = mixed(rng, n)
    tgt = uniform(rng, 150)
    f = HO.HarmonicFMM("stokes", lambda l, w: (lap, 1.0 / w), pts, 40, np.zeros(3), 1.0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)   # noqa: E731
    dens = rng.standard_normal((n, 9))
    dens[:, 6:] /= np.linalg.norm(dens[:, 6:], axis=1)[:, None]
    u, v = f.matvec(dens, output=3, targets=tgt)
    assert rel(u, HO.direct("stokes", pts, dens, pts, 3, skip=np.arange(n))) < 5e-6
    assert rel(v, HO.direct("stokes", pts, dens, tgt, 3)) < 5e-6
