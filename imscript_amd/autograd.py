"""Differentiable Poisson hole filling (torch.autograd).

poisson_fill -- the converged masked Poisson fill of simpois (simpois.c:
247-285 with CG to a relative tolerance): holes u_h solve
A u_h = -f_h - A_hk g_k with A = -Laplacian (clamped 5-point,
cleant_cgpois.c:67-89), known pixels are copied from g.  Backward: A is
symmetric, so the adjoint is one more masked Laplacian solve, A lambda =
dL/du on the holes, run on the GPU by the same pyramid and CG
(sp_adjoint_batch_device); then dL/df = -lambda on the holes and
dL/dg = dL/du + (sum of lambda over hole neighbours) on the known pixels.

biharmonic_fill -- the converged biharmonic fill (simpois -t negative, CG to
a relative tolerance): the same system with the reference's 13-point
clamped bilaplacian (cleant_cgpois.c:91-121).  Its clamped taps make A_hh
non-symmetric within two pixels of the border, so the backward solves
A_hh^T lambda = dL/du (the transposed operator in the FP32 CG), then
dL/df = -lambda on the holes and dL/dg = dL/du - A_hk^T lambda on the known
pixels.

simpois_fill -- simpois's own output with -c 0: niter Gauss-Seidel sweeps
at every level of the pyramid (bit-identical to the reference).  Backward:
the exact gradient of that finite-sweep output (sp_adjoint_batch_device
with cgit = 0: transposed sweeps, prolongation and restrictions on the GPU).
"""
import ctypes

import torch

from ._lib import lib, check, default_params, sp_stats


class sp_adjoint_image(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("h", ctypes.c_int), ("pd", ctypes.c_int),
                ("g", ctypes.c_void_p), ("grad_u", ctypes.c_void_p),
                ("grad_g", ctypes.c_void_p), ("grad_f", ctypes.c_void_p)]


def _adjoint(g, grad_u, grad_g, grad_f, params):
    pd, h, w = g.shape
    img = sp_adjoint_image(w, h, pd, g.data_ptr(), grad_u.data_ptr(),
                           grad_g.data_ptr(),
                           None if grad_f is None else grad_f.data_ptr())
    st = sp_stats()
    s = torch.cuda.current_stream()
    check(lib.sp_adjoint_batch_device(1, ctypes.cast(ctypes.byref(img),
                                                     ctypes.c_void_p),
                                      ctypes.byref(params),
                                      ctypes.c_void_p(s.cuda_stream),
                                      ctypes.byref(st)))
    return st


def _params(tol, niter, scale):
    # multigrid-preconditioned FP32 CG: converges to ~1e-7 relative where
    # the plain float CG stalls
    return default_params(tstep=0.25, niter=niter, scale=scale, cgit=-1,
                          cg_fp32=1, cg_rel=tol, cg_mg=1, trace=0)


def _check(g, f):
    """The checks of api.solve_device, before anything reaches a kernel: a
    host tensor's address or a wrong dtype or shape would otherwise be read
    as device float32 planes."""
    if not isinstance(g, torch.Tensor) or g.dtype != torch.float32 or not g.is_cuda:
        raise TypeError("g must be a float32 device tensor")
    if g.dim() != 3:
        raise TypeError("g must have shape (B, H, W)")
    if f is not None:
        if not isinstance(f, torch.Tensor) or f.dtype != torch.float32 or not f.is_cuda:
            raise TypeError("f must be a float32 device tensor")
        if f.device != g.device:
            raise TypeError("f must be on g's device")
        if f.shape != g.shape:
            raise TypeError("f must have g's shape")


class PoissonFill(torch.autograd.Function):
    """u = fill(g, f) for planar float32 device tensors of shape (B, H, W)."""

    @staticmethod
    def forward(ctx, g, f, tol, niter, scale):
        _check(g, f)
        g = g.contiguous()
        f = None if f is None else f.contiguous()
        out = torch.empty_like(g)
        p = _params(tol, niter, scale)
        pd, h, w = g.shape
        st = sp_stats()
        s = torch.cuda.current_stream()
        check(lib.sp_solve_device(out.data_ptr(), g.data_ptr(),
                                  None if f is None else f.data_ptr(), w, h, pd,
                                  ctypes.byref(p), ctypes.c_void_p(s.cuda_stream),
                                  ctypes.byref(st)))
        ctx.save_for_backward(g)
        ctx.has_f = f is not None
        ctx.cfg = (tol, niter, scale)
        ctx.iters = st.cg_iters_finest
        return out

    @staticmethod
    def backward(ctx, grad_u):
        (g,) = ctx.saved_tensors
        grad_u = grad_u.contiguous()
        grad_g = torch.empty_like(g)
        need_f = ctx.has_f and ctx.needs_input_grad[1]
        grad_f = torch.empty_like(g) if need_f else None
        _adjoint(g, grad_u, grad_g, grad_f, _params(*ctx.cfg))
        return grad_g, grad_f, None, None, None


def poisson_fill(g, f=None, tol=1e-6, niter=0, scale=1):
    """Differentiable fill of the NaN pixels of g ((H,W) or (B,H,W) float32
    on the GPU) with optional guidance f; gradients flow to the known pixels
    of g and to f."""
    squeeze = isinstance(g, torch.Tensor) and g.dim() == 2
    if squeeze:
        g = g[None]
        f = f[None] if isinstance(f, torch.Tensor) else f
    _check(g, f)
    u = PoissonFill.apply(g, f, tol, niter, scale)
    return u[0] if squeeze else u


class SimpoisFill(torch.autograd.Function):
    """u = simpois(g, f) with -c 0 for planar float32 device tensors
    (B, H, W): the output of niter Gauss-Seidel sweeps per pyramid level."""

    @staticmethod
    def forward(ctx, g, f, tstep, niter, scale):
        _check(g, f)
        g = g.contiguous()
        f = None if f is None else f.contiguous()
        out = torch.empty_like(g)
        p = default_params(tstep=tstep, niter=niter, scale=scale, cgit=0, trace=0)
        pd, h, w = g.shape
        st = sp_stats()
        s = torch.cuda.current_stream()
        check(lib.sp_solve_device(out.data_ptr(), g.data_ptr(),
                                  None if f is None else f.data_ptr(), w, h, pd,
                                  ctypes.byref(p), ctypes.c_void_p(s.cuda_stream),
                                  ctypes.byref(st)))
        ctx.save_for_backward(g)
        ctx.has_f = f is not None
        ctx.cfg = (tstep, niter, scale)
        return out

    @staticmethod
    def backward(ctx, grad_u):
        (g,) = ctx.saved_tensors
        grad_u = grad_u.contiguous()
        grad_g = torch.empty_like(g)
        need_f = ctx.has_f and ctx.needs_input_grad[1]
        grad_f = torch.empty_like(g) if need_f else None
        tstep, niter, scale = ctx.cfg
        _adjoint(g, grad_u, grad_g, grad_f,
                 default_params(tstep=tstep, niter=niter, scale=scale, cgit=0, trace=0))
        return grad_g, grad_f, None, None, None


def simpois_fill(g, f=None, niter=10, scale=99, tstep=0.25):
    """Differentiable simpois (-c 0): the NaN pixels of g ((H,W) or (B,H,W)
    float32 on the GPU) filled by niter Gauss-Seidel sweeps at every pyramid
    level, with optional data term f (simpois -f).  Gradients flow to the
    known pixels of g and to f; tstep > 0 (Laplace / Poisson)."""
    if not tstep > 0:
        raise ValueError("simpois_fill differentiates the Laplace / Poisson sweeps (tstep > 0)")
    squeeze = isinstance(g, torch.Tensor) and g.dim() == 2
    if squeeze:
        g = g[None]
        f = f[None] if isinstance(f, torch.Tensor) else f
    _check(g, f)
    u = SimpoisFill.apply(g, f, float(tstep), int(niter), int(scale))
    return u[0] if squeeze else u


def _bih_params(tol, niter, scale):
    # FP32 CG (x in FP64) without multigrid, which covers the Poisson system
    return default_params(tstep=-0.08, niter=niter, scale=scale, cgit=-1,
                          cg_fp32=1, cg_rel=tol, cg_mg=0, trace=0)


class BiharmonicFill(torch.autograd.Function):
    """u = biharmonic fill of (g, f) for planar float32 device tensors
    (B, H, W)."""

    @staticmethod
    def forward(ctx, g, f, tol, niter, scale):
        _check(g, f)
        g = g.contiguous()
        f = None if f is None else f.contiguous()
        out = torch.empty_like(g)
        p = _bih_params(tol, niter, scale)
        pd, h, w = g.shape
        st = sp_stats()
        s = torch.cuda.current_stream()
        check(lib.sp_solve_device(out.data_ptr(), g.data_ptr(),
                                  None if f is None else f.data_ptr(), w, h, pd,
                                  ctypes.byref(p), ctypes.c_void_p(s.cuda_stream),
                                  ctypes.byref(st)))
        ctx.save_for_backward(g)
        ctx.has_f = f is not None
        ctx.cfg = (tol, niter, scale)
        return out

    @staticmethod
    def backward(ctx, grad_u):
        (g,) = ctx.saved_tensors
        grad_u = grad_u.contiguous()
        grad_g = torch.empty_like(g)
        need_f = ctx.has_f and ctx.needs_input_grad[1]
        grad_f = torch.empty_like(g) if need_f else None
        tol, _, scale = ctx.cfg
        # the backward's own pyramid starts each level's CG; sweeps of the
        # forward operator have no place in the transposed solve
        _adjoint(g, grad_u, grad_g, grad_f, _bih_params(tol, 0, scale))
        return grad_g, grad_f, None, None, None


def biharmonic_fill(g, f=None, tol=1e-6, niter=0, scale=99):
    """Differentiable biharmonic fill of the NaN pixels of g ((H,W) or
    (B,H,W) float32 on the GPU) with optional data term f (Delta^2 u = -f
    on the holes); gradients flow to the known pixels of g and to f."""
    squeeze = isinstance(g, torch.Tensor) and g.dim() == 2
    if squeeze:
        g = g[None]
        f = f[None] if isinstance(f, torch.Tensor) else f
    _check(g, f)
    u = BiharmonicFill.apply(g, f, float(tol), int(niter), int(scale))
    return u[0] if squeeze else u
