"""Closed-form pins of the oracle, independent of the reference code.

Each fact is chosen so that a plausible restatement mistake breaks it:
a wrong neighbour, sign, weight, association or schedule.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import oracle, oracle_solve, ptr, bits_equal


def _gs(u, g, f, t, n):
    u = np.ascontiguousarray(u, np.float32).copy()
    h, w = u.shape
    oracle().orc_gs_sweeps(ptr(u), ptr(np.ascontiguousarray(g, np.float32)),
                           ptr(f), w, h, t, n)
    return u


def test_linear_ramp_is_refilled_exactly():
    # discrete-harmonic data: interior holes come back exactly (P2 of the
    # survey: the 4x5 ramp refills to 7, 11, 12)
    g = np.arange(20, dtype=np.float32).reshape(4, 5)
    want = g.copy()
    g[1, 2] = g[2, 1] = g[2, 3] = np.nan
    out, _ = oracle_solve(g)
    assert np.array_equal(out, want)


def test_single_hole_one_sweep_is_neighbour_mean():
    rng = np.random.default_rng(0)
    g = rng.random((5, 6)).astype(np.float32) * 10
    x0 = np.float32(3.5)
    g[2, 3] = np.nan
    u = np.where(np.isnan(g), x0, g).astype(np.float32)
    out = _gs(u, g, None, 0.25, 1)
    nb = [g[2, 4], g[3, 3], g[2, 2], g[1, 3]]
    r = np.float32(0)
    for v in nb:                      # E, S, W, N, each (v - x)
        r = np.float32(r + np.float32(v - x0))
    want = np.float32(x0 + np.float32(0.25) * r)
    assert out[2, 3] == want
    assert abs(float(want) - float(np.mean(nb))) < 1e-5


def test_maximum_principle_and_known_pixels():
    rng = np.random.default_rng(1)
    g = (rng.random((40, 50)) * 200 - 50).astype(np.float32)
    holes = rng.random(g.shape) < 0.4
    g[holes] = np.nan
    out, _ = oracle_solve(g, niter=30)
    known = ~holes
    assert np.array_equal(out[known], g[known])
    lo, hi = np.nanmin(g), np.nanmax(g)
    assert out.min() >= lo and out.max() <= hi


def test_all_nan_gives_zeros_and_no_nan_is_identity():
    out, _ = oracle_solve(np.full((9, 7), np.nan, np.float32))
    assert np.array_equal(out, np.zeros((9, 7), np.float32))
    g = np.random.default_rng(2).random((9, 7)).astype(np.float32)
    out, _ = oracle_solve(g)
    assert np.array_equal(out, g)


def test_poisson_converges_to_data_laplacian():
    # g = v with holes, f = clamped 5-point Laplacian of v: the fixed point
    # of the Poisson sweep is v itself
    rng = np.random.default_rng(3)
    h, w = 20, 24
    y, x = np.mgrid[0:h, 0:w]
    v = (np.sin(x / 4.0) * np.cos(y / 5.0) * 30).astype(np.float32)
    vp = np.pad(v, 1, mode="edge")
    f = (vp[1:-1, 2:] + vp[2:, 1:-1] + vp[1:-1, :-2] + vp[:-2, 1:-1]
         - 4 * v).astype(np.float32)
    g = v.copy()
    g[5:14, 6:17] = np.nan
    out, _ = oracle_solve(g, f, niter=3000, scale=1)
    assert np.max(np.abs(out - v)) < 2e-3


def test_biharmonic_annihilates_cubics():
    # the 13-point bilaplacian vanishes on cubic polynomials, so holes at
    # least 2 px from the border refill with the cubic
    h, w = 24, 26
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    v = (0.002 * x ** 3 - 0.01 * x * y ** 2 + 0.3 * y + 5).astype(np.float32)
    g = v.copy()
    g[8:15, 9:16] = np.nan
    out, _ = oracle_solve(g, tstep=-0.08, niter=4000, scale=1)
    assert np.max(np.abs(out - v)) < 2e-3 * np.max(np.abs(v))


@pytest.mark.parametrize("t,R", [(0.25, 1), (0.37, 1), (-0.08, 2), (-1000.01, 3)])
def test_hyperplane_schedule_equals_raster_order(t, R):
    # updates with equal i + R j + (R^2+1) k, in any order, reproduce raster
    # Gauss-Seidel bit for bit: the schedule the GPU kernels rely on
    rng = np.random.default_rng(R)
    h, w = 11, 13
    g = (rng.random((h, w)) * 50).astype(np.float32)
    g[rng.random((h, w)) < 0.5] = np.nan
    f = rng.standard_normal((h, w)).astype(np.float32)
    u0 = np.where(np.isnan(g), rng.random((h, w)) * 50, g).astype(np.float32)
    ref = _gs(u0, g, f, t, 5)
    for seed in (1, 2, 3):
        u = u0.copy()
        oracle().orc_gs_sweeps_hyperplane(ptr(u), ptr(g), ptr(f), w, h, t, 5, seed)
        assert bits_equal(u, ref)


def test_hyperplane_with_wrong_slope_breaks_order():
    # the same test with s = i + j + k (slope 1 in k) must fail for the
    # 5-point stencil, or the previous test would pin nothing
    rng = np.random.default_rng(9)
    h, w = 9, 10
    g = (rng.random((h, w)) * 50).astype(np.float32)
    g[rng.random((h, w)) < 0.7] = np.nan
    u0 = np.where(np.isnan(g), rng.random((h, w)) * 50, g).astype(np.float32)
    ref = _gs(u0, g, None, 0.25, 4)
    # emulate the wrong schedule in numpy, scrambled inside each hyperplane
    u = u0.copy()
    order = rng.permutation(4 * h * w)
    pts = sorted(((i + j + k, order[(k * h + j) * w + i], j, i) for k in range(4)
                  for j in range(h) for i in range(w)))
    for _, _, j, i in pts:
        if not np.isnan(g[j, i]):
            continue
        x = u[j, i]
        r = np.float32(0)
        for di, dj in ((1, 0), (0, 1), (-1, 0), (0, -1)):
            a, b = min(max(i + di, 0), w - 1), min(max(j + dj, 0), h - 1)
            if np.isfinite(u[b, a]):
                r = np.float32(r + np.float32(u[b, a] - x))
        u[j, i] = np.float32(x + np.float32(0.25) * r)
    assert not bits_equal(u, ref)


def test_restriction_weights_and_nan_rule():
    a = np.array([[1, 2, 3], [4, np.nan, 6], [np.inf, 8, 9]], np.float32)
    out = np.zeros((2, 2), np.float32)
    oracle().orc_restrict(ptr(out), 2, 2, ptr(a), 3, 3)
    # odd sizes duplicate the last row / column (clamped reads)
    assert out[0, 0] == np.float32((1 + 2 + 4) / np.float32(3))
    assert out[0, 1] == np.float32((3 + 3 + 6 + 6) / 4)
    assert out[1, 0] == np.float32(8 / 1)       # inf and nan dropped
    assert out[1, 1] == np.float32(9)
    nan = np.full((2, 2), np.nan, np.float32)
    o2 = np.zeros((1, 1), np.float32)
    oracle().orc_restrict(ptr(o2), 1, 1, ptr(nan), 2, 2)
    assert np.isnan(o2[0, 0])


def test_prolongation_weights():
    rng = np.random.default_rng(4)
    c = rng.random((5, 6)).astype(np.float32)
    out = np.zeros((10, 12), np.float32)
    oracle().orc_prolong(ptr(out), 12, 10, ptr(c), 6, 5)
    cp = np.pad(c.astype(np.float64), 1, mode="edge")
    # fine pixel (2m+1, 2n+1) = 9/16 c[m,n] + 3/16 right + 3/16 below + 1/16
    for j in range(1, 9):
        for i in range(1, 11):
            y, x = (j - 0.5) / 2, (i - 0.5) / 2
            iy, ix = int(np.floor(y)), int(np.floor(x))
            fy, fx = y - iy, x - ix
            want = ((1 - fx) * (1 - fy) * cp[iy + 1, ix + 1] + fx * (1 - fy) * cp[iy + 1, ix + 2]
                    + (1 - fx) * fy * cp[iy + 2, ix + 1] + fx * fy * cp[iy + 2, ix + 2])
            assert abs(out[j, i] - want) < 1e-6
    # weights are exactly 9/16, 3/16, 3/16, 1/16 at an interior odd pixel
    assert abs(out[3, 3] - (9 * c[1, 1] + 3 * c[1, 2] + 3 * c[2, 1] + c[2, 2]) / 16) < 1e-6


def _dense_system(g, f, bilap):
    """FP64 masked system of cleant_cgpois.c:225-236 built from scratch."""
    h, w = g.shape
    unk = np.flatnonzero(np.isnan(g).ravel())
    slot = -np.ones(h * w, int)
    slot[unk] = np.arange(len(unk))
    if bilap:
        st = [(0, 0, 20), (1, 0, -8), (0, 1, -8), (-1, 0, -8), (0, -1, -8),
              (1, 1, 2), (1, -1, 2), (-1, 1, 2), (-1, -1, 2), (2, 0, 1),
              (0, 2, 1), (-2, 0, 1), (0, -2, 1)]
        sgn = 1
    else:
        st = [(0, 0, -4), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1)]
        sgn = -1
    A = sp.lil_matrix((len(unk), len(unk)))
    b = np.zeros(len(unk))
    gf = g.astype(np.float64).ravel()
    for p, q in enumerate(unk):
        i, j = q % w, q // w
        bk = 0.0
        for di, dj, cf in st:
            a, bb = min(max(i + di, 0), w - 1), min(max(j + dj, 0), h - 1)
            k = bb * w + a
            if slot[k] >= 0:
                A[p, slot[k]] += sgn * cf
            else:
                bk += sgn * cf * gf[k]
        b[p] = (-float(f.ravel()[q]) if f is not None else 0.0) - bk
    return A.tocsr(), b, unk


@pytest.mark.parametrize("bilap", [0, 1])
def test_cg_converges_to_dense_fp64_solve(bilap):
    rng = np.random.default_rng(5 + bilap)
    h, w = 14, 16
    g = (rng.random((h, w)) * 40).astype(np.float32)
    g[3:10, 4:12] = np.nan
    f = rng.standard_normal((h, w)).astype(np.float32)
    A, b, unk = _dense_system(g, f, bilap)
    x = spla.spsolve(A, b)
    out = np.zeros_like(g)
    init = np.where(np.isnan(g), 0, g).astype(np.float32)
    its = oracle().orc_cg_extend(ptr(out), ptr(g), ptr(f), w, h, ptr(init),
                                 2000, 1e-9 if bilap else 1e-6, bilap, None, None)
    assert its > 0
    got = out.ravel()[unk].astype(np.float64)
    assert np.max(np.abs(got - x)) < 1e-4 * max(1, np.max(np.abs(x)))


def test_cg_zero_initial_residual_gives_nan():
    # exact harmonic data: r0 = 0, alpha = 0/0 (minicg.c:40-42)
    g = np.arange(20, dtype=np.float32).reshape(4, 5)
    g[1, 2] = np.nan
    out, tr = oracle_solve(g, cgit=3, scale=1, trace_cap=10)
    assert np.isnan(out[1, 2]) and len(tr) == 3
