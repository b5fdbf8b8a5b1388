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
