"""The allocation-free residual, sp_residual_device_noalloc, through the C
ABI: the same FP64 relative residuals, bit for bit, as sp_residual_device
on the same planes, over two calls in a row (the second smaller, so it
reuses the fixed device array the first one filled), for the Poisson and
the biharmonic systems; a call beyond the array's capacity is refused
before any launch."""
import ctypes

import numpy as np
import pytest


def _planes(torch, rng, pd, h, w, frac, with_f):
    g = (rng.standard_normal((pd, h, w)) * 20).astype(np.float32)
    g[rng.random((pd, h, w)) < frac] = np.nan
    u = np.where(np.isnan(g), rng.standard_normal((pd, h, w)), g).astype(np.float32)
    f = rng.standard_normal((pd, h, w)).astype(np.float32) if with_f else None
    dev = lambda a: None if a is None else torch.from_numpy(a).cuda()
    return dev(u), dev(g), dev(f)


def _rel(fn, u, g, f, bilap, stream):
    from imscript_amd import lib
    pd, h, w = g.shape
    rel = (ctypes.c_double * pd)()
    rc = getattr(lib, fn)(u.data_ptr(), g.data_ptr(), None if f is None else f.data_ptr(),
                          w, h, pd, bilap, rel, ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, lib.sp_last_error()
    return np.array(rel[:], np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("bilap", [0, 1])
def test_noalloc_residual_equals_the_allocating_one(bilap):
    import torch
    from imscript_amd import lib
    rng = np.random.default_rng(40 + bilap)
    s = torch.cuda.current_stream()
    before = lib.sp_device_bytes()
    # the first call larger than the second: the second reuses the array
    for pd, h, w, frac, with_f in [(3, 1030, 777, 0.4, True), (2, 64, 50, 0.3, False)]:
        print(f"noalloc residual bilap={bilap} {pd}x{h}x{w}", flush=True)
        u, g, f = _planes(torch, rng, pd, h, w, frac, with_f)
        want = _rel("sp_residual_device", u, g, f, bilap, s)
        got = _rel("sp_residual_device_noalloc", u, g, f, bilap, s)
        assert np.all(np.isfinite(want)) and np.all(want > 0)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    assert lib.sp_device_bytes() == before


def test_noalloc_residual_refuses_what_does_not_fit():
    """a 16384 x 16384 plane (268 M pixels) needs more than the 4 MiB of
    scratch: refused with SP_ERR_ARG on the host before any device call
    (the pointers are never dereferenced; no GPU needed)"""
    from imscript_amd import lib
    rel = (ctypes.c_double * 4)()
    fake = ctypes.c_void_p(1 << 20)
    rc = lib.sp_residual_device_noalloc(fake, fake, None, 16384, 16384, 1, 0, rel, None)
    assert rc == 1
    assert b"scratch" in lib.sp_last_error()
