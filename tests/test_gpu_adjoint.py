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


# ------------------------------------------------- simpois's finite sweeps

def sweep_problem(seed, b, h, w, frac=0.4):
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal((b, h, w)) * 10).astype(np.float32)
    g[rng.random((b, h, w)) < frac] = np.nan
    f = rng.standard_normal((b, h, w)).astype(np.float32)
    v = rng.standard_normal((b, h, w)).astype(np.float32)
    return g, f, v


def sweep_grads(g, f, v, niter, scale, tstep=0.25):
    """(u, dL/dg, dL/df) of L = sum(v * simpois_fill(g, f)) on the GPU"""
    import torch
    from imscript_amd.autograd import simpois_fill
    dg = torch.from_numpy(g).cuda().requires_grad_(True)
    df = None if f is None else torch.from_numpy(f).cuda().requires_grad_(True)
    u = simpois_fill(dg, df, niter=niter, scale=scale, tstep=tstep)
    (u * torch.from_numpy(v).cuda()).sum().backward()
    return (u.detach().cpu().numpy(), dg.grad.cpu().numpy(),
            None if f is None else df.grad.cpu().numpy())


@pytest.mark.parametrize("h,w,niter,scale,withf,tstep", [
    (12, 14, 3, 99, True, 0.25), (9, 13, 5, 2, True, 0.25), (17, 10, 4, 99, False, 0.25),
    (7, 5, 1, 99, True, 0.25), (16, 16, 6, 3, True, 0.4), (11, 20, 2, 1, True, 0.25)])
def test_sweep_gradient_matches_fp64_finite_differences(h, w, niter, scale, withf, tstep):
    """The gradient of simpois's own -c 0 output (niter sweeps at every
    pyramid level) against FP64 central differences (eps 1e-3) of the FP64
    restatement of the same map (oracle/adjoint_ref.simpois64, itself within
    float rounding of the bit-exact oracle), two planes, bar 1e-3 relative;
    the forward output is the reference's bit for bit (same float32 kernels
    as simpois)."""
    from helpers import bits_equal, oracle_solve
    g, f, v = sweep_problem(h * 31 + w, 2, h, w)
    if not withf:
        f = None
    print(f"sweep adjoint {h}x{w} n={niter} s={scale} f={withf} t={tstep}", flush=True)
    u, gg, gf = sweep_grads(g, f, v, niter, scale, tstep)
    for k in range(2):
        fk = None if f is None else f[k]
        want, _ = oracle_solve(g[k], fk, tstep, niter, scale, 0)
        assert bits_equal(u[k], want.reshape(h, w))
        rg, rf = ar.sweeps_gradient_fd(g[k], fk, v[k], tstep=tstep, niter=niter,
                                       scale=scale)
        assert rel_err(gg[k], rg) < 1e-3, k
        assert not np.any(gg[k][np.isnan(g[k])])    # 0 at the holes
        if f is not None:
            assert rel_err(gf[k], rf) < 1e-3, k


def test_sweep_gradient_dot_product_at_scale():
    """The map is affine in (g, f), so for any direction (dg on the known
    pixels, df everywhere) <grad, d> = sum(v * (u(g + dg, f + df) - u(g, f))):
    checked at 2 x 600 x 700 with 30 sweeps on 10 levels, and the gradient
    is the same on a second call (cached scratch)."""
    import torch
    from imscript_amd.autograd import simpois_fill
    g, f, v = sweep_problem(5, 2, 600, 700, 0.5)
    rng = np.random.default_rng(6)
    dg = np.where(np.isnan(g), 0, rng.standard_normal(g.shape)).astype(np.float32)
    df = rng.standard_normal(g.shape).astype(np.float32)
    print("sweep adjoint dot product 2x600x700 n=30", flush=True)
    u, gg, gf = sweep_grads(g, f, v, 30, 99)
    u1 = simpois_fill(torch.from_numpy(g + dg).cuda(), torch.from_numpy(f + df).cuda(),
                      niter=30, scale=99).cpu().numpy()
    lhs = float(np.sum(gg.astype(np.float64) * dg) + np.sum(gf.astype(np.float64) * df))
    rhs = float(np.sum(v.astype(np.float64) * (u1.astype(np.float64) - u)))
    print("<grad, d>", lhs, "finite difference", rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-3 * abs(rhs)
    _, gg2, gf2 = sweep_grads(g, f, v, 30, 99)
    assert np.array_equal(gg2.view(np.uint32), gg.view(np.uint32))
    assert np.array_equal(gf2.view(np.uint32), gf.view(np.uint32))


# ------------------------------------------- the converged biharmonic fill

def bih_problem(seed, b=2, h=22, w=26):
    """Smooth data with holes, some of them on the image border, where the
    clamped 13-point operator is not symmetric."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    g = np.empty((b, h, w), np.float32)
    for k in range(b):
        g[k] = 30 * np.sin(x / rng.uniform(5, 9)) * np.cos(y / rng.uniform(5, 9)) + 50
    g[:, 5:13, 6:17] = np.nan
    g[:, 0:3, w - 7:w] = np.nan
    g[b - 1, h - 6:h, 0:4] = np.nan
    f = rng.normal(0, 0.05, (b, h, w)).astype(np.float32)
    v = rng.standard_normal((b, h, w)).astype(np.float32)
    return g, f, v


def bih_grads(g, f, v, tol):
    import torch
    from imscript_amd.autograd import biharmonic_fill
    dg = torch.from_numpy(g).cuda().requires_grad_(True)
    df = torch.from_numpy(f).cuda().requires_grad_(True)
    u = biharmonic_fill(dg, df, tol=tol)
    (u * torch.from_numpy(v).cuda()).sum().backward()
    return u.detach().cpu().numpy(), dg.grad.cpu().numpy(), df.grad.cpu().numpy()


def test_biharmonic_gradient_matches_fp64_reference():
    """Forward against the FP64 direct solve and the gradient against the
    FP64 adjoint (oracle/adjoint_ref.gradient with bilap, pinned by the exact
    change of a linear loss on the CPU), bar 1e-3 relative.  The symmetric
    shortcut (lambda = A_hh^-1 v) misses the reference by more than the bar
    on these border holes, so the transposed operator is what passes."""
    import scipy.sparse.linalg as spla
    g, f, v = bih_problem(3)
    print("biharmonic adjoint", g.shape, flush=True)
    u, gg, gf = bih_grads(g, f, v, 1e-7)
    for k in range(g.shape[0]):
        hole = np.isnan(g[k])
        uref = ar.solve_bilap(g[k], f[k])
        assert rel_err(u[k], uref) < 1e-4, k
        rg, rf = ar.gradient(g[k], f[k], v[k], bilap=True)
        print(k, "grad_g", rel_err(gg[k], rg), "grad_f", rel_err(gf[k], rf), flush=True)
        assert rel_err(gg[k], rg) < 1e-3, k
        assert rel_err(gf[k], rf) < 1e-3, k
        assert np.all(gg[k][hole] == 0) and np.all(gf[k][~hole] == 0)
        Ahh, _, ih, _ = ar.masked_bilaplacian(g[k].astype(np.float64))
        wrong = np.zeros(g[k].size)
        wrong[ih] = -spla.spsolve(Ahh.tocsc(), v[k].astype(np.float64).ravel()[ih])
        assert rel_err(wrong.reshape(g[k].shape), rf) > 1e-2, k


def test_biharmonic_gradient_dot_product():
    """At 2 x 96 x 120 against the change of the loss for a random direction
    (the fill is affine in (g, f)), bar 1e-3 relative."""
    import torch
    from imscript_amd.autograd import biharmonic_fill
    g, f, v = bih_problem(4, 2, 96, 120)
    rng = np.random.default_rng(7)
    dg = np.where(np.isnan(g), 0, rng.standard_normal(g.shape)).astype(np.float32)
    df = (0.05 * rng.standard_normal(g.shape)).astype(np.float32)
    print("biharmonic adjoint dot product 2x96x120", flush=True)
    u, gg, gf = bih_grads(g, f, v, 1e-7)
    u1 = biharmonic_fill(torch.from_numpy(g + dg).cuda(), torch.from_numpy(f + df).cuda(),
                         tol=1e-7).cpu().numpy()
    lhs = float(np.sum(gg.astype(np.float64) * dg) + np.sum(gf.astype(np.float64) * df))
    rhs = float(np.sum(v.astype(np.float64) * (u1.astype(np.float64) - u)))
    print("<grad, d>", lhs, "change of the loss", rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-3 * abs(rhs)


def test_biharmonic_adjoint_batch_of_uneven_images():
    """One sp_adjoint_batch_device call on two images of different sizes and
    plane counts (the transposed solve of a batch), each plane against the
    FP64 adjoint, bar 1e-3 relative."""
    import ctypes
    import torch
    from imscript_amd._lib import lib, check, sp_stats
    from imscript_amd.autograd import sp_adjoint_image, _bih_params
    shapes = [(1, 22, 26), (2, 30, 18)]
    imgs, keep = [], []
    rng = np.random.default_rng(11)
    for pd, h, w in shapes:
        g, _, v = bih_problem(int(rng.integers(100)), pd, h, w)
        dg, dv = torch.from_numpy(g).cuda(), torch.from_numpy(v).cuda()
        gg, gf = torch.empty_like(dg), torch.empty_like(dg)
        keep.append((g, v, dg, dv, gg, gf))
        imgs.append(sp_adjoint_image(w, h, pd, dg.data_ptr(), dv.data_ptr(),
                                     gg.data_ptr(), gf.data_ptr()))
    arr = (sp_adjoint_image * len(imgs))(*imgs)
    p = _bih_params(1e-7, 0, 99)
    st = sp_stats()
    print("biharmonic adjoint batch", shapes, flush=True)
    check(lib.sp_adjoint_batch_device(len(imgs), ctypes.cast(arr, ctypes.c_void_p),
                                      ctypes.byref(p), None, ctypes.byref(st)))
    torch.cuda.synchronize()
    for g, v, _, _, gg, gf in keep:
        gg, gf = gg.cpu().numpy(), gf.cpu().numpy()
        for k in range(g.shape[0]):
            rg, rf = ar.gradient(g[k], None, v[k], bilap=True)
            assert rel_err(gg[k], rg) < 1e-3, (g.shape, k)
            assert rel_err(gf[k], rf) < 1e-3, (g.shape, k)
