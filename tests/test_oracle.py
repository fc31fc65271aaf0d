"""The oracle pinned on the CPU: its tree against every brute-force invariant and the
exact-once coverage of the reference's tests, the closed-form tree cases, its direct
sums against a plain numpy sum, and its FMM against the direct sum, the analytic
sphere field and linearity (test/test_tree.jl, test/test_coverage.jl,
test/test_3d.jl:97-163)."""
import numpy as np
import pytest

import oracle as O
from pointsets import DISTRIBUTIONS, mixed, uniform
from treecheck import check_coverage, check_tree_invariants

BOX = (np.zeros(3), 1.0)


@pytest.mark.parametrize("name", list(DISTRIBUTIONS))
def test_tree_invariants(name):
    rng = np.random.default_rng(3)
    for n, ppl in ((300, 4), (3000, 30)):
        pts = DISTRIBUTIONS[name](rng, n)
        t = O.build_tree(pts, ppl, *BOX)
        check_tree_invariants(t, pts, locate=t.locate)


@pytest.mark.parametrize("name", list(DISTRIBUTIONS))
def test_coverage(name):
    rng = np.random.default_rng(4)
    pts = DISTRIBUTIONS[name](rng, 400)
    t = O.build_tree(pts, 4, *BOX)
    check_coverage(t, t.lists)


def test_coverage_small_ppl_many_ghosts():
    rng = np.random.default_rng(5)
    t = O.build_tree(mixed(rng, 800), 3, *BOX)
    assert t.ghost.sum() > 0
    check_coverage(t, t.lists)
    t = O.build_tree(uniform(rng, 12), 12, *BOX)
    check_coverage(t, t.lists)


def test_tree_closed_forms_and_guards():
    rng = np.random.default_rng(6)
    t = O.build_tree(uniform(rng, 10), 10, *BOX)
    assert t.nlevels == 1 and t.nnodes == 1
    oc = np.array([[0.25 + 0.5 * a, 0.25 + 0.5 * b, 0.25 + 0.5 * c] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
    pts = np.repeat(oc, 6, axis=0) + 0.01 * rng.random((48, 3))
    t = O.build_tree(pts, 6, *BOX)
    assert t.nlevels == 2 and t.nnodes == 9 and t.ghost.sum() == 0
    check_tree_invariants(t, pts, locate=t.locate)
    t3 = O.build_tree(mixed(rng, 3000), 8, *BOX)
    assert t3.ghost.sum() > 0 and t3.nlevels >= 4
    with pytest.raises(O.OracleError, match="max_levels"):
        O.build_tree(np.full((20, 3), 0.3), 5, *BOX)
    with pytest.raises(O.OracleError, match="outside"):
        O.build_tree(uniform(rng, 20) + 1.0, 5, *BOX)
    a = O.build_tree(mixed(np.random.default_rng(9), 2000), 10, *BOX)
    b = O.build_tree(mixed(np.random.default_rng(9), 2000), 10, *BOX)
    assert np.array_equal(a.perm, b.perm)                      # deterministic


def test_direct_sums_against_numpy():
    rng = np.random.default_rng(7)
    s, t = rng.random((300, 3)), rng.random((50, 3))
    tau = rng.standard_normal(300)
    d = t[:, None, :] - s[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    u = (1 / (4 * np.pi * r)) @ tau
    g = np.einsum("jik,ji,i->jk", d, -1 / (4 * np.pi * r ** 3), tau)
    assert np.allclose(O.direct(s, tau, t), u, rtol=1e-13, atol=0)
    assert np.allclose(O.direct(s, tau, t, grad=True), g, rtol=1e-12, atol=1e-14 * np.abs(g).max())
    ds = s[:, None, :] - s[None, :, :]
    rs = np.sqrt((ds ** 2).sum(-1))
    np.fill_diagonal(rs, np.inf)
    assert np.allclose(O.direct_self(s, tau), (1 / (4 * np.pi * rs)) @ tau, rtol=1e-12)


@pytest.fixture(scope="module")
def ops8():
    from ffmm import operators
    return operators.level_operators(1.0, p=8)


def test_oracle_fmm_against_direct(ops8):
    """Laplace cube p = 8 end to end: max-relative < 1e-6 at the sources and at
    targets, potentials and gradients (test/test_3d.jl:155-186); dense and spectral
    translations agree to rounding (test/test_cubeki.jl:43-68)."""
    rng = np.random.default_rng(8)
    N = 3000
    pts = mixed(rng, N)
    tau = rng.standard_normal(N)
    f = O.FMM(ops8, pts, 30, *BOX)
    assert f.tree.nlevels >= 5
    targets = uniform(rng, 500)
    u, v = f.matvec(tau, targets=targets)
    rows = np.arange(0, N, 7)
    ud = O.direct_self(pts, tau, rows=rows)
    assert np.max(np.abs(u[rows] - ud)) / np.max(np.abs(ud)) < 1e-6
    vd = O.direct(pts, tau, targets)
    assert np.max(np.abs(v - vd)) / np.max(np.abs(vd)) < 1e-6
    g = f.evaluate(f.Lc, tau[f.tree.perm], targets, grad=True)
    gd = O.direct(pts, tau, targets, grad=True)
    assert np.max(np.linalg.norm(g - gd, axis=1)) / np.max(np.linalg.norm(gd, axis=1)) < 1e-5
    u2 = f.matvec(tau, m2l="spectral")
    assert np.max(np.abs(u2 - u)) / np.max(np.abs(u)) < 1e-12
    t2 = rng.standard_normal(N)
    assert np.allclose(f.matvec(tau + 2 * t2), u + 2 * f.matvec(t2), rtol=0, atol=1e-12 * np.abs(u).max())


def test_oracle_charged_sphere(ops8):
    """a^2/r outside a unit-charged sphere, a inside (test/test_3d.jl:240-251)."""
    c, a = np.array([0.5, 0.5, 0.5]), 0.3
    x, wg = np.polynomial.legendre.leggauss(40)
    ph = 2 * np.pi * np.arange(80) / 80
    st = np.sqrt(1 - x ** 2)
    nrm = np.stack([np.outer(st, np.cos(ph)), np.outer(st, np.sin(ph)), np.outer(x, np.ones(80))], -1).reshape(-1, 3)
    pts = c + a * nrm
    wts = (a * a * np.outer(wg, np.full(80, 2 * np.pi / 80))).ravel()
    f = O.FMM(ops8, pts, 40, *BOX)
    rng = np.random.default_rng(11)
    cand = rng.random((3000, 3))
    ext = cand[np.linalg.norm(cand - c, axis=1) > 0.45][:200]
    _, v = f.matvec(wts, targets=ext)
    ex = a * a / np.linalg.norm(ext - c, axis=1)
    assert np.max(np.abs(v - ex)) / np.max(np.abs(ex)) < 1e-6
