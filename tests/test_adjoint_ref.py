"""The FP64 references of oracle/adjoint_ref.py pinned on the CPU: central
finite differences of its own FP64 solve, the symmetry of the clamped
5-point operator, agreement of its solve with the oracle's restatement run
to convergence; for the 13-point bilaplacian, exact refill of cubics,
agreement with an independent per-pixel assembly, and its gradient against
the exact change of a linear loss."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from helpers import ROOT, oracle_solve

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import adjoint_ref as ar  # noqa: E402


def problem(seed, h=12, w=14):
    rng = np.random.default_rng(seed)
    g = rng.random((h, w)) * 50
    g[3:9, 4:11] = np.nan
    g[0, 0:3] = np.nan
    g[h - 1, w - 2] = np.nan
    f = rng.standard_normal(g.shape)
    t = rng.random(g.shape) * 50
    return g, f, t


def loss(g, f, t):
    u = ar.solve(g, f)
    hole = np.isnan(g)
    return np.sum((u - t)[hole] ** 2)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_central_differences(seed):
    g, f, t = problem(seed)
    hole = np.isnan(g)
    u = ar.solve(g, f)
    v = np.where(hole, 2 * (u - t), 0)
    gg, gf = ar.gradient(g, f, v)
    eps = 1e-3
    known = list(zip(*np.nonzero(~hole)))
    rng = np.random.default_rng(seed)
    for k in rng.choice(len(known), 12, replace=False):
        a, b = known[k]
        g1, g2 = g.copy(), g.copy()
        g1[a, b] += eps
        g2[a, b] -= eps
        fd = (loss(g1, f, t) - loss(g2, f, t)) / (2 * eps)
        assert abs(fd - gg[a, b]) <= 1e-6 * max(1, abs(fd))
    holes = list(zip(*np.nonzero(hole)))
    for k in rng.choice(len(holes), 8, replace=False):
        a, b = holes[k]
        f1, f2 = f.copy(), f.copy()
        f1[a, b] += eps
        f2[a, b] -= eps
        fd = (loss(g, f1, t) - loss(g, f2, t)) / (2 * eps)
        assert abs(fd - gf[a, b]) <= 1e-6 * max(1, abs(fd))
    assert np.all(gg[hole] == 0) and np.all(gf[~hole] == 0)


def test_operator_is_symmetric():
    # <A x, y> = <x, A y> for the clamped 5-point masked operator: the reason
    # the adjoint is another masked Laplacian solve
    g, _, _ = problem(3, 17, 19)
    g[:, -1] = np.nan            # holes on the image border (clamped reads)
    Ahh, _, _, _ = ar.masked_operator(g)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal((2, Ahh.shape[0]))
    assert abs((Ahh @ x) @ y - x @ (Ahh @ y)) < 1e-9 * np.abs(x).sum()
    assert (Ahh - Ahh.T).count_nonzero() == 0


def test_fp64_solve_matches_converged_oracle():
    g, f, _ = problem(5)
    u = ar.solve(g, f)
    o, _ = oracle_solve(g.astype(np.float32), f.astype(np.float32), niter=0,
                        scale=1, cgit=-1, cg_eps=1e-9)
    assert np.max(np.abs(o - u)) < 1e-3


def test_bilaplacian_fp64_solve_refills_cubics():
    # the 13-point bilaplacian annihilates cubic polynomials away from the
    # border, so holes at least two pixels inside refill the cubic exactly
    h, w = 23, 29
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    cub = 0.01 * x ** 3 - 0.02 * x * y * y + 0.5 * x * y - 3 * y + 7
    g = cub.copy()
    g[4:12, 5:20] = np.nan
    g[15:19, 3:9] = np.nan
    u = ar.solve_bilap(g)
    assert np.max(np.abs(u - cub)) < 1e-8 * np.max(np.abs(cub))


def test_bilaplacian_operator_matches_independent_assembly():
    # the vectorised assembly equals the per-pixel one of the oracle tests,
    # right-hand side included, on holes that touch the border
    from test_oracle_math import _dense_system
    rng = np.random.default_rng(8)
    g = (rng.random((14, 16)) * 40).astype(np.float32)
    g[3:10, 4:12] = np.nan
    g[0:2, 13:16] = np.nan
    f = rng.standard_normal((14, 16)).astype(np.float32)
    A, b, unk = _dense_system(g, f, 1)
    Ahh, Ahk, ih, ik = ar.masked_bilaplacian(g.astype(np.float64))
    assert np.array_equal(ih, unk)
    assert abs(A - Ahh).max() == 0
    b2 = -Ahk @ g.astype(np.float64).ravel()[ik] - f.astype(np.float64).ravel()[ih]
    assert np.max(np.abs(b - b2)) < 1e-9 * np.max(np.abs(b))
    u = ar.solve_bilap(g, f)
    assert np.max(np.abs(u.ravel()[ih] - spla.spsolve(A.tocsc(), b))) < 1e-8 * np.max(np.abs(u))


def test_bilaplacian_gradient_is_exact_and_needs_the_transpose():
    """gradient(bilap=True) against the change of the loss: the fill is
    affine in (g, f), so <grad, d> equals L(g + d) - L(g) for a linear loss,
    on holes that touch the border.  The same formula with A_hh in place of
    its transpose fails this check there (A_hh is not symmetric)."""
    rng = np.random.default_rng(12)
    g = rng.random((13, 15)) * 40
    g[2:9, 3:11] = np.nan
    g[0:2, 12:15] = np.nan
    g[10:13, 0:2] = np.nan
    f = rng.standard_normal(g.shape)
    v = rng.standard_normal(g.shape)
    hole = np.isnan(g)
    gg, gf = ar.gradient(g, f, v, bilap=True)
    assert np.all(gg[hole] == 0) and np.all(gf[~hole] == 0)
    dg = np.where(hole, 0, rng.standard_normal(g.shape))
    df = rng.standard_normal(g.shape)
    du = ar.solve_bilap(g + dg, f + df) - ar.solve_bilap(g, f)
    lhs = np.sum(gg * dg) + np.sum(gf * df)
    assert np.isclose(lhs, np.sum(v * du), rtol=1e-9)
    # the symmetric shortcut (lambda = A_hh^-1 v) gives a different gradient
    Ahh, Ahk, ih, ik = ar.masked_bilaplacian(g)
    assert (Ahh - Ahh.T).count_nonzero() > 0
    lam_wrong = spla.spsolve(Ahh.tocsc(), v.ravel()[ih])
    gf_wrong = np.zeros(g.size)
    gf_wrong[ih] = -lam_wrong
    lhs_wrong = np.sum(gg * dg) + np.sum(gf_wrong.reshape(g.shape) * df)
    assert not np.isclose(lhs_wrong, np.sum(v * du), rtol=1e-6)
    # interior holes only: the operator is symmetric there and both agree
    g2 = rng.random((15, 15)) * 40
    g2[5:10, 5:10] = np.nan
    Ahh2, _, _, _ = ar.masked_bilaplacian(g2)
    assert (Ahh2 - Ahh2.T).count_nonzero() == 0


# ------------------------------------------------- simpois's finite sweeps

@pytest.mark.parametrize("h,w,niter,scale,withf,tstep", [
    (12, 14, 3, 99, True, 0.25), (9, 13, 5, 2, True, 0.25), (17, 10, 4, 99, False, 0.25),
    (7, 5, 1, 99, True, 0.25), (16, 16, 6, 3, True, 0.4), (1, 9, 3, 99, True, 0.25)])
def test_simpois64_matches_the_oracle(h, w, niter, scale, withf, tstep):
    """The FP64 restatement of -c 0 simpois (sweeps, pyramid, level output)
    against the oracle (bit-exact with the reference) within float rounding:
    a wrong weight, index or term would differ at the data's scale."""
    from helpers import oracle_solve
    rng = np.random.default_rng(h * 7 + w)
    g = (rng.standard_normal((h, w)) * 10).astype(np.float32)
    g[rng.random((h, w)) < 0.4] = np.nan
    f = rng.standard_normal((h, w)).astype(np.float32) if withf else None
    u64 = ar.simpois64(g, f, tstep, niter, scale)
    u32, _ = oracle_solve(g, f, tstep, niter, scale, 0)
    assert np.max(np.abs(u64 - u32.reshape(h, w))) <= 1e-6 * np.max(np.abs(u64))


def test_sweeps_fd_gradient_is_the_transpose():
    """sweeps_gradient_fd against an independent check: the map is affine,
    so for random directions <grad, d> equals the change of the loss, and
    the loss is linear in v (the gradient of a sum is the sum)."""
    rng = np.random.default_rng(9)
    g = rng.standard_normal((8, 11)) * 5
    g[rng.random(g.shape) < 0.4] = np.nan
    f = rng.standard_normal(g.shape)
    v1, v2 = rng.standard_normal(g.shape), rng.standard_normal(g.shape)
    kw = dict(tstep=0.25, niter=3, scale=99)
    gg1, gf1 = ar.sweeps_gradient_fd(g, f, v1, **kw)
    gg2, gf2 = ar.sweeps_gradient_fd(g, f, v2, **kw)
    gg12, gf12 = ar.sweeps_gradient_fd(g, f, v1 + v2, **kw)
    assert np.allclose(gg12, gg1 + gg2, atol=1e-9) and np.allclose(gf12, gf1 + gf2, atol=1e-9)
    dg = np.where(np.isnan(g), 0, rng.standard_normal(g.shape))
    df = rng.standard_normal(g.shape)
    u0 = ar.simpois64(g, f, **kw)
    u1 = ar.simpois64(g + dg, f + df, **kw)
    assert np.isclose(np.sum(gg1 * dg) + np.sum(gf1 * df), np.sum(v1 * (u1 - u0)),
                      rtol=1e-9)
