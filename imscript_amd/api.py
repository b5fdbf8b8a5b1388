"""Array-level interface over the C ABI."""
import ctypes

import numpy as np

from ._lib import lib, check, default_params, sp_stats, sp_image, sp_plane_stats


def _f32(a, name):
    a = np.asarray(a)
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous float32 array")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data


def poisson_solver_separable(out, inp, dat, w, h, pd, tstep=0.25, niter=10,
                             scale=99, cgrad=0.0):
    """Same arguments and effect as simpois.c:288-302, on numpy float32
    planar buffers of w*h*pd values (``dat`` may be None).  ``out`` is
    written in place."""
    out = _f32(out, "out")
    inp = _f32(inp, "inp")
    n = w * h * pd
    if out.size != n or inp.size != n or (dat is not None and dat.size != n):
        raise ValueError("buffers must hold w*h*pd floats")
    if dat is not None:
        dat = _f32(dat, "dat")
    if np.shares_memory(out, inp):
        raise ValueError("out must not alias inp")
    lib.poisson_solver_separable(_ptr(out), _ptr(inp), _ptr(dat), w, h, pd,
                                 tstep, int(niter), int(scale), float(cgrad))
    return out


def poisson_rec(u, g, f, w, h, tstep=0.25, niter=10, scale=99, cgit=0):
    """simpois.c:247-285 on one plane."""
    u, g = _f32(u, "u"), _f32(g, "g")
    if f is not None:
        f = _f32(f, "f")
    lib.poisson_rec(_ptr(u), _ptr(g), _ptr(f), w, h, tstep, niter, scale, cgit)
    return u


def _cg(fn, out, inp, dat, w, h, init, maxit, eps):
    out, inp, init = _f32(out, "out"), _f32(inp, "inp"), _f32(init, "init")
    if dat is not None:
        dat = _f32(dat, "dat")
    fn(_ptr(out), _ptr(inp), _ptr(dat), w, h, _ptr(init), maxit, eps)
    return out


def poisson_extension_by_cg(out, inp, dat, w, h, init, maxit, eps):
    """cleant_cgpois.c:266-271."""
    return _cg(lib.poisson_extension_by_cg, out, inp, dat, w, h, init, maxit, eps)


def biharmonic_extension_by_cg(out, inp, dat, w, h, init, maxit, eps):
    """cleant_cgpois.c:273-278."""
    return _cg(lib.biharmonic_extension_by_cg, out, inp, dat, w, h, init,
               maxit, eps)


def simpois(g, f=None, tstep=0.25, niter=10, scale=99, cgit=0, trace=True,
            stats=False, **knobs):
    """Fill the NaN pixels of ``g`` (shape (h, w) or (pd, h, w), planar).

    Returns the filled array (and an sp_stats when ``stats``).  Status codes
    are raised as SimpoisError."""
    g = np.ascontiguousarray(g, dtype=np.float32)
    shape = g.shape
    g3 = g.reshape((1,) + shape) if g.ndim == 2 else g
    pd, h, w = g3.shape
    if f is not None:
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(g3.shape)
    out = np.empty_like(g3)
    p = default_params(tstep=tstep, niter=niter, scale=scale, cgit=cgit,
                       trace=int(bool(trace)), **knobs)
    st = sp_stats()
    check(lib.sp_solve_host(_ptr(out), _ptr(g3), _ptr(f), w, h, pd,
                            ctypes.byref(p), ctypes.byref(st)))
    out = out.reshape(shape)
    return (out, st) if stats else out


def solve_device(g, f=None, out=None, stream=None, tstep=0.25, niter=10,
                 scale=99, cgit=0, trace=False, **knobs):
    """Solve on device tensors (torch, float32, planar (pd,h,w) or (h,w)).
    Returns (out, sp_stats)."""
    import torch
    if g.dtype != torch.float32 or not g.is_cuda or not g.is_contiguous():
        raise TypeError("g must be a contiguous float32 device tensor")
    pd, h, w = (1,) + tuple(g.shape) if g.dim() == 2 else tuple(g.shape)
    if f is not None and (f.shape != g.shape or f.dtype != torch.float32
                          or not f.is_contiguous()):
        raise TypeError("f must match g")
    if out is None:
        out = torch.empty_like(g)
    s = stream if stream is not None else torch.cuda.current_stream()
    p = default_params(tstep=tstep, niter=niter, scale=scale, cgit=cgit,
                       trace=int(bool(trace)), **knobs)
    st = sp_stats()
    check(lib.sp_solve_device(out.data_ptr(), g.data_ptr(),
                              None if f is None else f.data_ptr(), w, h, pd,
                              ctypes.byref(p), ctypes.c_void_p(s.cuda_stream),
                              ctypes.byref(st)))
    return out, st


def residual_device(u, g, f=None, bilap=False, stream=None):
    """FP64 relative residual per plane of the masked Poisson (or
    biharmonic) system, on device tensors."""
    import torch
    pd, h, w = (1,) + tuple(g.shape) if g.dim() == 2 else tuple(g.shape)
    rel = (ctypes.c_double * pd)()
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.sp_residual_device(u.data_ptr(), g.data_ptr(),
                                 None if f is None else f.data_ptr(), w, h, pd,
                                 int(bool(bilap)), rel,
                                 ctypes.c_void_p(s.cuda_stream)))
    return list(rel)


def solve_batch_device(gs, fs=None, outs=None, stream=None, plane_stats=False,
                       **params):
    """Batched solve of device tensors of uneven shapes ((h,w) or (pd,h,w)),
    one pyramid per plane, all planes advancing together.  ``params`` are
    sp_params fields (tstep, niter, scale, cgit, cg_fp32, cg_rel, ...).
    Returns (outs, sp_stats), or (outs, sp_stats, [sp_plane_stats per
    plane]) with ``plane_stats=True`` (sp_solve_batch_device_ex)."""
    import torch
    n = len(gs)
    fs = fs if fs is not None else [None] * n
    if outs is None:
        outs = [torch.empty_like(g) for g in gs]
    arr = (sp_image * n)()
    for k, (g, f, o) in enumerate(zip(gs, fs, outs)):
        if g.dtype != torch.float32 or not g.is_contiguous() or not g.is_cuda:
            raise TypeError("inputs must be contiguous float32 device tensors")
        pd, h, w = (1,) + tuple(g.shape) if g.dim() == 2 else tuple(g.shape)
        arr[k] = sp_image(w, h, pd, g.data_ptr(),
                          None if f is None else f.data_ptr(), o.data_ptr())
    s = stream if stream is not None else torch.cuda.current_stream()
    params.setdefault("trace", 0)
    p = default_params(**params)
    st = sp_stats()
    if not plane_stats:
        check(lib.sp_solve_batch_device(n, arr, ctypes.byref(p),
                                        ctypes.c_void_p(s.cuda_stream),
                                        ctypes.byref(st)))
        return outs, st
    npl = sum(1 if g.dim() == 2 else g.shape[0] for g in gs)
    ps = (sp_plane_stats * npl)()
    check(lib.sp_solve_batch_device_ex(n, arr, ctypes.byref(p),
                                       ctypes.c_void_p(s.cuda_stream),
                                       ctypes.byref(st), ps))
    return outs, st, list(ps)


def last_cg_trace(plane=0, level=0):
    """(unknowns, [r.r after each CG iteration]) of the last solve at one
    plane and pyramid level (0 = finest), or None if it had no such level."""
    n = ctypes.c_int()
    k = lib.sp_last_cg_trace(plane, level, ctypes.byref(n), None, 0)
    if k < 0:
        return None
    buf = (ctypes.c_double * max(k, 1))()
    lib.sp_last_cg_trace(plane, level, ctypes.byref(n), buf, k)
    return n.value, list(buf[:k])


def plan_cache_limit(max_plans=256, max_bytes=16 << 30):
    """Bound the solver's plan cache (least recently used plans go first)."""
    lib.sp_plan_cache_limit(int(max_plans), int(max_bytes))


def plan_cache_clear():
    """Free every cached solve plan."""
    lib.sp_plan_cache_clear()


def set_alloc_cap(nbytes):
    """Cap the device bytes the library holds (0: none); returns the old cap.
    Allocations beyond it fail with SP_ERR_NOMEM (SimpoisError)."""
    return lib.sp_set_alloc_cap(int(nbytes))


def device_bytes():
    """Device bytes the library holds now (cached plans included)."""
    return lib.sp_device_bytes()


def profile(on=True):
    """Enable (and clear) or disable the solver's phase timers."""
    lib.sp_profile(1 if on else 0)


def profile_read(name):
    """(device ms, kernel launches, algorithmic bytes) summed over a phase."""
    ms, n, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
    check(lib.sp_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(n),
                              ctypes.byref(by)))
    return ms.value, n.value, by.value


def autotune_report():
    """Tile shapes chosen by the autotuner in this process, one text line per
    tuned level size (see sp_autotune_report in include/simpois_hip.h)."""
    n = lib.sp_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.sp_autotune_report(buf, n)
    return buf.value.decode()


def cg_persist_counts():
    """(levels, given up): FP64 CG levels solved by the persistent launch in
    this process, and launches that found their work-groups not all resident
    and left the iterations to separate launches."""
    a, b = ctypes.c_long(), ctypes.c_long()
    lib.sp_cg_persist_counts(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def stream_copy(dst, src, stream=None):
    """dst[:] = src with the library's streaming copy kernel (device float32
    tensors of equal size, a multiple of 4 elements)."""
    import torch
    if (dst.dtype != torch.float32 or src.dtype != torch.float32
            or dst.numel() != src.numel() or not dst.is_contiguous()
            or not src.is_contiguous()):
        raise TypeError("dst and src must be contiguous float32 tensors of one size")
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.sp_stream_copy(dst.data_ptr(), src.data_ptr(), src.numel(),
                             ctypes.c_void_p(s.cuda_stream)))


def plambda_lap(x, variant=0, boundary=1, out=None, stream=None):
    """plambda's "x,l" (variant 0) or "x,lr" (1) Laplacian of planar float32
    planes (h, w) or (pd, h, w); boundary = the PLAMBDA_GETPIXEL rule (0 zero,
    1 nearest, 2 reflection, 3 periodic, 4 NaN).  numpy arrays run through
    the host entry, torch device tensors through the device entry."""
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        x3 = x.reshape((1,) + x.shape) if x.ndim == 2 else x
        pd, h, w = x3.shape
        y = np.empty_like(x3)
        check(lib.sp_plambda_lap_host(_ptr(y), _ptr(x3), w, h, pd, variant, boundary))
        return y.reshape(x.shape)
    import torch
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise TypeError("x must be a contiguous float32 device tensor or a numpy array")
    pd, h, w = (1,) + tuple(x.shape) if x.dim() == 2 else tuple(x.shape)
    if out is None:
        out = torch.empty_like(x)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.sp_plambda_lap_device(out.data_ptr(), x.data_ptr(), w, h, pd, variant,
                                    boundary, ctypes.c_void_p(s.cuda_stream)))
    return out


def amle(x, nn=4, niter=100, init=None, trace=False, out=None, stream=None):
    """AMLE hole filling of amle.c on planar float32 planes (h, w) or
    (pd, h, w): the NaN pixels of each plane are filled by niter raster
    sweeps of the AMLE_NN = nn neighbour scheme started from the plane's
    maximum (or from `init` at the holes, as AMLE_INIT does).  trace=True
    also sweeps the lower bound and prints the reference's diagnostic line
    every tenth sweep on stderr.  numpy arrays run through the host entry,
    torch device tensors through the device entry (out may alias x)."""
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        x3 = x.reshape((1,) + x.shape) if x.ndim == 2 else x
        pd, h, w = x3.shape
        i3 = None
        if init is not None:
            i3 = np.ascontiguousarray(init, dtype=np.float32).reshape(x3.shape)
        y = np.empty_like(x3)
        check(lib.sp_amle_host(_ptr(y), _ptr(x3), _ptr(i3), w, h, pd, int(nn),
                               int(niter), int(bool(trace))))
        return y.reshape(x.shape)
    import torch
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise TypeError("x must be a contiguous float32 device tensor or a numpy array")
    pd, h, w = (1,) + tuple(x.shape) if x.dim() == 2 else tuple(x.shape)
    if init is not None and (init.dtype != torch.float32 or not init.is_cuda or
                             not init.is_contiguous() or init.numel() != x.numel()):
        raise TypeError("init must be a contiguous float32 device tensor of x's size")
    if out is None:
        out = torch.empty_like(x)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.sp_amle_device(out.data_ptr(), x.data_ptr(),
                             None if init is None else init.data_ptr(), w, h, pd,
                             int(nn), int(niter), int(bool(trace)),
                             ctypes.c_void_p(s.cuda_stream)))
    return out


def heatd(x, n=1, t=0.25, A=None, out=None, stream=None):
    """heatd's explicit linear diffusion of planar float32 planes (h, w) or
    (pd, h, w): n steps of x + t * L(x) with the 9-point scheme of the
    diffusion-matrix field A ((h, w, 3) interleaved a, b, c; None: the
    isotropic Laplacian), clamped borders.  numpy arrays run through the
    host entry, torch device tensors through the device entry (x is not
    written; out must not alias it)."""
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        x3 = x.reshape((1,) + x.shape) if x.ndim == 2 else x
        pd, h, w = x3.shape
        a = None if A is None else np.ascontiguousarray(A, np.float32).reshape(h, w, 3)
        y = np.empty_like(x3)
        check(lib.sp_heatd_host(_ptr(y), _ptr(x3), _ptr(a), w, h, pd, int(n), float(t)))
        return y.reshape(x.shape)
    import torch
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise TypeError("x must be a contiguous float32 device tensor or a numpy array")
    pd, h, w = (1,) + tuple(x.shape) if x.dim() == 2 else tuple(x.shape)
    if A is not None and (A.dtype != torch.float32 or not A.is_cuda or
                          not A.is_contiguous() or A.numel() != 3 * w * h):
        raise TypeError("A must be a contiguous float32 device tensor of h*w*3 values")
    if out is None:
        out = torch.empty_like(x)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.sp_heatd_device(out.data_ptr(), x.data_ptr(),
                              None if A is None else A.data_ptr(), w, h, pd, int(n),
                              float(t), ctypes.c_void_p(s.cuda_stream)))
    return out
