"""GPU parity of plambda's 3x3 Laplacian operators (the producer of
simpois's guidance field) through the C ABI: bit-identical to the golden
vectors of the reference plambda and to the oracle restatement, every
variant and boundary rule; the solver's fused guidance option equals the
two-step pipeline bit for bit."""
import numpy as np
import pytest

from helpers import bits_equal, load_plambda_golden, oracle_plambda_lap

pytestmark = pytest.mark.gpu

GOLDEN = load_plambda_golden()


def _m():
    import imscript_amd
    return imscript_amd


def test_golden_through_the_host_entry():
    m = _m()
    print(f"plambda golden {len(GOLDEN)} cases", flush=True)
    for x, want, variant, mode in GOLDEN:
        got = m.plambda_lap(x, variant, mode)
        assert bits_equal(got, want), (x.shape, variant, mode)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("mode", range(5))
def test_device_entry_matches_oracle(variant, mode):
    import torch
    m = _m()
    rng = np.random.default_rng(40 + 5 * variant + mode)
    x = (rng.standard_normal((3, 517, 1031)) * 40).astype(np.float32)
    x[0, 5, 7], x[1, 0, 0], x[2, -1, -1], x[0, 300, 1030] = np.nan, np.inf, -np.inf, -0.0
    print(f"plambda device variant={variant} mode={mode} {x.shape}", flush=True)
    got = m.plambda_lap(torch.from_numpy(x).cuda(), variant, mode)
    torch.cuda.synchronize()
    assert bits_equal(got.cpu().numpy(), oracle_plambda_lap(x, variant, mode))


def test_aligned_planes_take_the_vector_path_and_match_oracle():
    # widths divisible by 4 and 16-byte aligned planes: 16-byte row loads in
    # 32-row strips; heights that end a strip early, one- and two-row planes,
    # every boundary rule and both stencils
    import torch
    m = _m()
    rng = np.random.default_rng(48)
    for shape in [(3, 517, 1032), (1, 1, 8), (2, 2, 4), (1, 31, 2048), (2, 33, 12),
                  (1, 65, 1028)]:
        x = (rng.standard_normal(shape) * 40).astype(np.float32)
        x[0, 0, 0], x[-1, -1, -1], x[0, shape[1] // 2, shape[2] // 2] = np.inf, -np.inf, np.nan
        dx = torch.from_numpy(x).cuda()
        for variant in (0, 1):
            for mode in range(5):
                print(f"plambda vector {shape} variant={variant} mode={mode}", flush=True)
                got = m.plambda_lap(dx, variant, mode)
                torch.cuda.synchronize()
                assert bits_equal(got.cpu().numpy(), oracle_plambda_lap(x, variant, mode)), \
                    (shape, variant, mode)


def test_unaligned_planes_and_odd_widths():
    # a view one float into its buffer is not 16-byte aligned, and odd
    # widths leave partial groups: both take the element-wise path
    import torch
    m = _m()
    rng = np.random.default_rng(47)
    for h, w in [(33, 64), (31, 129), (7, 5)]:
        x = (rng.standard_normal(h * w + 1) * 30).astype(np.float32)
        base = torch.from_numpy(x).cuda()
        view = base[1:].view(h, w)
        out = torch.empty(h * w + 1, device="cuda")[1:].view(h, w)
        print(f"plambda unaligned {h}x{w}", flush=True)
        m.plambda_lap(view, 0, 1, out=out)
        torch.cuda.synchronize()
        want = oracle_plambda_lap(x[1:].reshape(h, w), 0, 1)
        assert bits_equal(out.cpu().numpy(), want), (h, w)


def test_bad_arguments_are_refused():
    import torch
    m = _m()
    x = torch.zeros(8, 9, device="cuda")
    with pytest.raises(m.SimpoisError):
        m.plambda_lap(x, 2, 1)
    with pytest.raises(m.SimpoisError):
        m.plambda_lap(x, 0, 5)
    with pytest.raises(m.SimpoisError):
        m.plambda_lap(x, 0, 1, out=x)       # in place is not allowed


@pytest.mark.parametrize("guid,kw", [
    (1, dict(niter=12, scale=99)),                                  # GS pyramid
    (2, dict(niter=0, scale=1, cgit=-1, cg_fp32=1, cg_rel=1e-6, cg_mg=1)),
])
def test_fused_guidance_equals_the_two_step_pipeline(guid, kw):
    import torch
    m = _m()
    from test_gpu_parity import random_case
    g, v = random_case(81, 2, 300, 413, 0.4, True)
    dg, dv = torch.from_numpy(g).cuda(), torch.from_numpy(v).cuda()
    print(f"fused guidance {guid} {kw}", flush=True)
    f = m.plambda_lap(dv, guid - 1, 1)
    two, _ = m.solve_device(dg, f, **kw)
    two = two.clone()
    one, _ = m.solve_device(dg, dv, guidance=guid, **kw)
    torch.cuda.synchronize()
    assert torch.equal(one.view(torch.int32), two.view(torch.int32))
