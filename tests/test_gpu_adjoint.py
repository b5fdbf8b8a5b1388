"""GPU gradient of the converged Poisson fill (imscript_amd.autograd) against
the FP64 CPU reference gradient and FP64 central finite differences
(eps 1e-3), bar 1e-3 relative."""
import os
import sys

import numpy as np
import pytest

from helpers import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import adjoint_ref as ar  # noqa: E402

pytestmark = pytest.mark.gpu


def problem(seed, b=3, h=40, w=48):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    g = np.empty((b, h, w), np.float32)
    holes = np.zeros((b, h, w), bool)
    for k in range(b):
        g[k] = 60 * np.sin(x / rng.uniform(4, 12)) * np.cos(y / rng.uniform(4, 12)) + 100
        for _ in range(4):
            cy, cx, r = rng.uniform(0, h), rng.uniform(0, w), rng.uniform(3, 9)
            holes[k] |= (y - cy) ** 2 + (x - cx) ** 2 < r * r
    g[holes] = np.nan
    f = rng.normal(0, 1, (b, h, w)).astype(np.float32)
    t = (rng.random((b, h, w)) * 200).astype(np.float32)
    return g, f, t


def gpu_grad(g, f, t, tol=1e-7):  # multigrid CG reaches it in float
    import torch
    from imscript_amd.autograd import poisson_fill
    dg = torch.from_numpy(g).cuda().requires_grad_(True)
    df = torch.from_numpy(f).cuda().requires_grad_(True)
    dt = torch.from_numpy(t).cuda()
    hole = torch.isnan(dg.detach())
    u = poisson_fill(dg, df, tol=tol)
    loss = ((u - dt) ** 2)[hole].sum()
    loss.backward()
    gg = dg.grad.cpu().numpy()
    gf = df.grad.cpu().numpy()
    return u.detach().cpu().numpy(), gg, gf


def rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


def test_forward_and_gradient_match_fp64_reference():
    g, f, t = problem(1)
    print("adjoint vs FP64 reference", g.shape, flush=True)
    u, gg, gf = gpu_grad(g, f, t)
    for k in range(g.shape[0]):
        uref = ar.solve(g[k], f[k])
        assert rel_err(u[k], uref) < 1e-5
        hole = np.isnan(g[k])
        v = np.where(hole, 2 * (uref - t[k]), 0)
        rg, rf = ar.gradient(g[k], f[k], v)
        assert rel_err(gg[k], rg) < 1e-3
        assert rel_err(gf[k], rf) < 1e-3


def test_gradient_matches_fp64_finite_differences():
    g, f, t = problem(2, b=1, h=24, w=28)
    print("adjoint vs FD", g.shape, flush=True)
    _, gg, gf = gpu_grad(g, f, t)
    g64, f64, t64 = g[0].astype(np.float64), f[0].astype(np.float64), t[0]
    hole = np.isnan(g64)

    def L(gx, fx):
        return np.sum((ar.solve(gx, fx) - t64)[hole] ** 2)

    eps = 1e-3
    rng = np.random.default_rng(3)
    # known pixels next to holes carry the adjoint flux
    ring = (~hole) & (np.roll(hole, 1, 0) | np.roll(hole, -1, 0)
                      | np.roll(hole, 1, 1) | np.roll(hole, -1, 1))
    cand = list(zip(*np.nonzero(ring)))
    for k in rng.choice(len(cand), 10, replace=False):
        a, b = cand[k]
        g1, g2 = g64.copy(), g64.copy()
        g1[a, b] += eps
        g2[a, b] -= eps
        fd = (L(g1, f64) - L(g2, f64)) / (2 * eps)
        assert abs(gg[0, a, b] - fd) <= 1e-3 * max(abs(fd), 1.0)
    holes = list(zip(*np.nonzero(hole)))
    for k in rng.choice(len(holes), 10, replace=False):
        a, b = holes[k]
        f1, f2 = f64.copy(), f64.copy()
        f1[a, b] += eps
        f2[a, b] -= eps
        fd = (L(g64, f1) - L(g64, f2)) / (2 * eps)
        assert abs(gf[0, a, b] - fd) <= 1e-3 * max(abs(fd), 1.0)
