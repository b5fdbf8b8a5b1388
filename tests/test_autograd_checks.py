"""The differentiable fill refuses inputs the solver cannot read before
anything reaches the library or a kernel (runs without a GPU): host
tensors, other dtypes, mismatched shapes or devices."""
import pytest
import torch

from imscript_amd.autograd import PoissonFill, poisson_fill


def test_host_tensor_is_refused():
    g = torch.zeros(1, 8, 8)
    with pytest.raises(TypeError, match="device"):
        poisson_fill(g)
    with pytest.raises(TypeError, match="device"):
        PoissonFill.apply(g, None, 1e-6, 0, 1)


def test_wrong_dtype_is_refused():
    with pytest.raises(TypeError, match="float32"):
        poisson_fill(torch.zeros(1, 8, 8, dtype=torch.float64))


def test_guidance_shape_and_dtype_are_checked(monkeypatch):
    # a stand-in for a device tensor: the checks look at dtype, device and
    # shape only, and must fail before any use of its memory
    class Fake(torch.Tensor):
        @property
        def is_cuda(self):
            return True

    g = torch.zeros(1, 8, 8).as_subclass(Fake)
    with pytest.raises(TypeError, match="shape"):
        poisson_fill(g, torch.zeros(1, 8, 9).as_subclass(Fake))
    with pytest.raises(TypeError, match="float32"):
        poisson_fill(g, torch.zeros(1, 8, 8, dtype=torch.float16).as_subclass(Fake))
    with pytest.raises(TypeError, match="device tensor"):
        poisson_fill(g, torch.zeros(1, 8, 8))
    with pytest.raises(TypeError, match=r"\(B, H, W\)"):
        PoissonFill.apply(torch.zeros(2, 1, 8, 8).as_subclass(Fake), None, 1e-6, 0, 1)


def test_biharmonic_fill_refuses_host_tensors():
    from imscript_amd.autograd import biharmonic_fill
    with pytest.raises(TypeError, match="device"):
        biharmonic_fill(torch.zeros(1, 8, 8))


def test_adjoint_argument_rules_before_any_device_work():
    """sp_adjoint_batch_device checks its parameters before it touches a
    device: tstep 0 is refused, and so is the biharmonic operator (tstep < 0)
    with -c 0 (the finite-sweep gradient covers the 5-point sweeps)."""
    import ctypes
    from imscript_amd._lib import lib, default_params
    from imscript_amd.autograd import sp_adjoint_image
    img = sp_adjoint_image(8, 8, 1, None, None, None, None)
    for kw, msg in [(dict(tstep=0.0, cgit=-1), b"non-zero"),
                    (dict(tstep=-0.08, cgit=0), b"5-point")]:
        p = default_params(**kw)
        rc = lib.sp_adjoint_batch_device(1, ctypes.cast(ctypes.byref(img), ctypes.c_void_p),
                                         ctypes.byref(p), None, None)
        assert rc != 0 and msg in lib.sp_last_error()
