"""heatd on the GPU (sp_heatd_device / sp_heatd_host, the drop-ins and
main_heatd) against the reference's golden vectors and the oracle."""
import os
import subprocess

import numpy as np
import pytest

import imscript_amd as sp
from helpers import HEATD_BIN, bits_equal, load_heatd_golden, oracle_heatd

pytestmark = pytest.mark.gpu

CASES = load_heatd_golden()


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("k", range(len(CASES)))
def test_host_entry_matches_reference_golden(k):
    c = CASES[k]
    print(f"golden case {k}: {c['x'].shape} n={c['n']} aniso={c['A'] is not None}",
          flush=True)
    y = sp.heatd(c["x"], c["n"], np.float32(c["t"]), c["A"])
    assert bits_equal(y, c["out"])


def field(rng, h, w):
    return np.stack([rng.uniform(0.2, 1.5, (h, w)), rng.uniform(-0.8, 0.8, (h, w)),
                     rng.uniform(0.2, 1.5, (h, w))], -1).astype(np.float32)


@pytest.mark.parametrize("aniso", [False, True])
@pytest.mark.parametrize("shape", [(3, 517, 1032), (2, 301, 777), (1, 5, 3000), (1, 2000, 3)])
def test_device_matches_oracle(shape, aniso):
    rng = np.random.default_rng(shape[1] + aniso)
    x = (rng.standard_normal(shape) * 10).astype(np.float32)
    x.reshape(-1)[rng.choice(x.size, 6, replace=False)] = [np.nan, np.inf, -np.inf,
                                                           -0.0, 0.0, 1e30]
    A = field(rng, *shape[1:]) if aniso else None
    print(f"heatd {shape} aniso={aniso}", flush=True)
    y = host(sp.heatd(dev(x), 4, 0.23, None if A is None else dev(A)))
    assert bits_equal(y, oracle_heatd(x, 4, np.float32(0.23), A)[0])


@pytest.mark.parametrize("T", [1, 2, 3, 4, 5, 6, 7, 8])
def test_multi_step_register_tiles_match_oracle(T, monkeypatch):
    # several steps per launch (isotropic): tiles clamped at every image
    # border, the last tile column / row moved back, a remainder launch;
    # non-finite values on the borders and corners of the image
    monkeypatch.setenv("SIMPOIS_HIP_HEATD_T", str(T))
    rng = np.random.default_rng(T)
    OW, OH = 64 - 2 * T, 48 - 2 * T
    for shape in [(2, 97, 131), (1, OH, OW), (1, 211, 70), (1, 2 * OH + T, 2 * OW + T),
                  (1, OH + T - 1, 3 * OW + 2), (1, 3 * OH - 1, OW + T + 1)]:
        x = (rng.standard_normal(shape) * 10).astype(np.float32)
        p, h, w = shape
        x[0, 0, 0], x[0, -1, -1], x[0, 0, w // 2] = np.inf, -np.inf, np.nan
        x[-1, h // 2, 0], x[-1, 3, -1] = -0.0, 1e30
        x[-1, rng.integers(0, h, 5), rng.integers(0, w, 5)] = np.nan
        n = 2 * T + 3
        print(f"heatd T={T} {shape} n={n}", flush=True)
        y = host(sp.heatd(dev(x), n, 0.21))
        assert bits_equal(y, oracle_heatd(x, n, np.float32(0.21))[0]), shape


@pytest.mark.parametrize("T,rows", [(T, r) for r in (16, 24, 32, 48)
                                    for T in range(1, 9) if r >= 3 * T - 1])
def test_multi_step_field_register_tiles_match_oracle(T, rows, monkeypatch):
    # several steps per launch with a diffusion-matrix field: the same tile
    # geometry as the isotropic kernel with `rows` rows per tile (at least
    # 3T - 1, so that the first and last tile of a column write the rows the
    # others leave); non-finite image values on the borders and corners,
    # zero and -0 shear terms
    monkeypatch.setenv("SIMPOIS_HIP_HEATD_FT", str(T))
    monkeypatch.setenv("SIMPOIS_HIP_HEATD_FR", str(rows))
    rng = np.random.default_rng(100 + 10 * T + rows)
    OW, OH = 64 - 2 * T, rows - 2 * T
    for shape in [(2, 97, 131), (1, OH, OW), (1, 211, 70), (1, 2 * OH + T, 2 * OW + T),
                  (1, OH + T - 1, 3 * OW + 2), (1, 3 * OH - 1, OW + T + 1)]:
        x = (rng.standard_normal(shape) * 10).astype(np.float32)
        p, h, w = shape
        x[0, 0, 0], x[0, -1, -1], x[0, 0, w // 2] = np.inf, -np.inf, np.nan
        x[-1, h // 2, 0], x[-1, 3, -1] = -0.0, 1e30
        x[-1, rng.integers(0, h, 5), rng.integers(0, w, 5)] = np.nan
        A = field(rng, h, w)
        A[0, 0, 1], A[h - 1, w // 3, 1], A[h // 2, w - 1, 1] = 0.0, -0.0, 0.0
        n = 2 * T + 3
        print(f"heatd field T={T} rows={rows} {shape} n={n}", flush=True)
        y = host(sp.heatd(dev(x), n, 0.21, dev(A)))
        assert bits_equal(y, oracle_heatd(x, n, np.float32(0.21), A)[0]), shape


def test_field_steps_per_launch_beyond_the_tile_limit_are_reduced(monkeypatch):
    # 8 steps per launch asked on 16-row tiles (which hold at most 5): the
    # library runs 5 per launch and the result stays the reference's
    monkeypatch.setenv("SIMPOIS_HIP_HEATD_FT", "8")
    monkeypatch.setenv("SIMPOIS_HIP_HEATD_FR", "16")
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((2, 97, 131)) * 10).astype(np.float32)
    A = field(rng, 97, 131)
    y = host(sp.heatd(dev(x), 19, 0.21, dev(A)))
    assert bits_equal(y, oracle_heatd(x, 19, np.float32(0.21), A)[0])


def test_unaligned_views_use_the_element_path():
    import torch
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 64, 257)).astype(np.float32)
    big = torch.zeros(x.size + 1, device="cuda")
    big[1:] = dev(x).reshape(-1)
    view = big[1:].reshape(x.shape)             # 4-byte offset: not 16-byte aligned
    out = torch.empty(x.size + 1, device="cuda")[1:].reshape(x.shape)
    sp.heatd(view, 3, 0.25, out=out)
    assert bits_equal(host(out), oracle_heatd(x, 3, 0.25)[0])


def test_zero_steps_copy_and_bad_arguments():
    x = np.arange(12, dtype=np.float32).reshape(3, 4)
    assert np.array_equal(sp.heatd(x, 0), x)
    y = np.empty_like(x)
    assert sp.lib.sp_heatd_host(y.ctypes.data, x.ctypes.data, None, 4, 3, 1, -1,
                                0.25) == 1
    assert sp.lib.sp_heatd_host(y.ctypes.data, x.ctypes.data, None, 0, 3, 1, 1,
                                0.25) == 1
    assert sp.lib.sp_heatd_host(x.ctypes.data, x.ctypes.data, None, 4, 3, 1, 1,
                                0.25) == 1


def test_drop_ins_leave_x_equal_to_y():
    rng = np.random.default_rng(8)
    x = rng.standard_normal((2, 40, 50)).astype(np.float32)
    A = field(rng, 40, 50)
    want, _ = oracle_heatd(x, 6, 0.2, A)
    xx = x.copy()
    y = np.empty_like(x)
    sp.lib.linear_diffusion_separable(y.ctypes.data, xx.ctypes.data, 50, 40, 2,
                                      A.ctypes.data, 6, 0.2)
    assert bits_equal(y, want) and bits_equal(xx, want)
    one = x[0].copy()
    y1 = np.empty_like(one)
    sp.lib.linear_diffusion_step(y1.ctypes.data, one.ctypes.data, 50, 40, None, 0.2)
    assert bits_equal(y1, oracle_heatd(x[0], 1, 0.2)[0])
    assert np.array_equal(one, x[0])             # the step leaves x alone
    y0 = np.full_like(x, 7.0)
    xx = x.copy()
    sp.lib.linear_diffusion_separable(y0.ctypes.data, xx.ctypes.data, 50, 40, 2,
                                      None, 0, 0.2)
    assert (y0 == 7.0).all() and np.array_equal(xx, x)


def test_cli_matches_golden(tmp_path):
    for k, c in enumerate(CASES):
        if k % 3:
            continue
        np.save(tmp_path / "in.npy", c["xi"])
        args = [HEATD_BIN, "-n", str(c["n"]), "-t", repr(c["t"])]
        if c["A"] is not None:
            np.save(tmp_path / "A.npy", c["A"])
            args += ["-A", str(tmp_path / "A.npy")]
        args += [str(tmp_path / "in.npy"), str(tmp_path / "out.npy")]
        print(f"cli case {k}", flush=True)
        p = subprocess.run(args, capture_output=True, timeout=120)
        assert p.returncode == 0, p.stderr
        y = np.load(tmp_path / "out.npy").reshape(c["xi"].shape)
        assert bits_equal(y.transpose(2, 0, 1), c["out"]), k
