"""AMLE on the GPU (sp_amle_device / sp_amle_host / amle / main_amle)
against the reference's golden vectors and the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import imscript_amd as sp
from helpers import AMLE_BIN, bits_equal, load_amle_golden, oracle_amle

pytestmark = pytest.mark.gpu

CASES = load_amle_golden()


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
    print(f"golden case {k}: {c['x'].shape} nn={c['nn']} niter={c['niter']}", flush=True)
    y = sp.amle(c["x"], c["nn"], c["niter"], init=c["init"])
    assert bits_equal(y, c["out"])


@pytest.mark.parametrize("nn", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 20, 24, 28, 32])
def test_device_matches_oracle_on_every_neighbourhood(nn):
    rng = np.random.default_rng(100 + nn)
    h, w = (211, 307) if nn > 4 else (517, 1031)
    x = (rng.standard_normal((h, w)) * 20).astype(np.float32)
    ties = rng.random(x.shape) < 0.2
    x[ties] = np.round(x[ties])
    x[rng.random(x.shape) < 0.7] = np.nan
    print(f"nn={nn} {h}x{w}", flush=True)
    y = host(sp.amle(dev(x), nn, 37))
    assert bits_equal(y, oracle_amle(x, nn, 37))


@pytest.mark.parametrize("var", range(15))
def test_every_four_neighbour_tile_shape_matches_oracle(var, monkeypatch):
    monkeypatch.setenv("SIMPOIS_HIP_AMLE_VARIANT", str(var))
    rng = np.random.default_rng(var)
    x = rng.standard_normal((300, 410)).astype(np.float32)
    x[rng.random(x.shape) < 0.8] = np.nan
    print(f"variant {var}", flush=True)
    y = host(sp.amle(dev(x), 4, 45))
    assert bits_equal(y, oracle_amle(x, 4, 45))


def test_planes_are_independent_and_init_and_aliasing_work():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((3, 120, 90)).astype(np.float32)
    x[rng.random(x.shape) < 0.6] = np.nan
    init = (rng.standard_normal(x.shape) * 3).astype(np.float32)
    want = np.stack([oracle_amle(x[l], 8, 20, init[l]) for l in range(3)])
    t = dev(x)
    sp.amle(t, 8, 20, init=dev(init), out=t)          # in place
    assert bits_equal(host(t), want)


def test_large_plane_default_parameters():
    # the default AMLE_NN = 4, AMLE_NITER = 100 on a 2048 x 2048 plane with
    # three quarters of it missing
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2048, 2048)).astype(np.float32)
    x[rng.random(x.shape) < 0.75] = np.nan
    print("2048x2048 nn=4 niter=100", flush=True)
    y = host(sp.amle(dev(x)))
    assert bits_equal(y, oracle_amle(x, 4, 100))


def test_zero_sweeps_and_no_holes():
    x = np.array([[1.0, np.nan], [np.nan, 3.0]], np.float32)
    y = sp.amle(x, 4, 0)
    assert y.tolist() == [[1.0, 3.0], [3.0, 3.0]]
    z = np.arange(12, dtype=np.float32).reshape(3, 4)
    assert np.array_equal(sp.amle(z, 4, 7), z)


def _trace_numbers(line):
    # "iter N, e = {A B}, eactus = {C D}" -> [N, A, B, C, D]
    return [float(t) for t in re.findall(r"-?(?:nan|inf|[0-9.]+(?:e[-+]?[0-9]+)?)", line)]


def test_trace_lines_match_reference(capfd):
    c = next(c for c in CASES if c["niter"] == 100)
    capfd.readouterr()
    y = sp.amle(c["x"], c["nn"], c["niter"], trace=True)
    err = capfd.readouterr().err
    assert bits_equal(y, c["out"])
    got = [ln for ln in err.splitlines() if ln.startswith("iter ")]
    ref = c["trace"].splitlines()
    assert len(got) == len(ref) == 10
    # e is a maximum (exact); ea and the sums are float running sums in the
    # reference and double tree sums here
    for a, b in zip(got, ref):
        va, vb = _trace_numbers(a), _trace_numbers(b)
        assert len(va) == len(vb) == 5
        assert np.allclose(va, vb, rtol=1e-4, equal_nan=True), (a, b)


def test_drop_in_and_cli_match_golden(tmp_path):
    for k, c in enumerate(CASES):
        if c["mask"] is None and c["init"] is None and k % 9:
            continue
        np.save(tmp_path / "in.npy", c["raw"])
        env = dict(os.environ, AMLE_NN=str(c["nn"]), AMLE_NITER=str(c["niter"]))
        env.pop("AMLE_INIT", None)
        args = [AMLE_BIN, str(tmp_path / "in.npy")]
        if c["mask"] is not None:
            np.save(tmp_path / "mask.npy", c["mask"])
            args.append(str(tmp_path / "mask.npy"))
        else:
            args.append("")
        if c["init"] is not None:
            np.save(tmp_path / "init.npy", c["init"])
            env["AMLE_INIT"] = str(tmp_path / "init.npy")
        args.append(str(tmp_path / "out.npy"))
        print(f"cli case {k}", flush=True)
        p = subprocess.run(args, env=env, capture_output=True, timeout=120)
        assert p.returncode == 0, p.stderr
        y = np.load(tmp_path / "out.npy").reshape(c["out"].shape)
        assert bits_equal(y, c["out"]), k
        assert b'scanning the environment for "AMLE_NITER"' in p.stderr


def test_drop_in_function_through_the_c_abi(monkeypatch):
    monkeypatch.setenv("SIMPOIS_HIP_QUIET", "1")
    c = CASES[40]
    x = np.ascontiguousarray(c["x"], np.float32)
    y = np.empty_like(x)
    # the smart parameters keep their first values for the process; run the
    # drop-in with whatever they hold and compare with the oracle's run
    sp.lib.amle(y.ctypes.data, x.ctypes.data, x.shape[1], x.shape[0])
    nn = int(os.environ.get("AMLE_NN", "4"))
    niter = int(np.ceil(float(os.environ.get("AMLE_NITER", "100"))))
    assert bits_equal(y, oracle_amle(c["x"], nn, niter))


@pytest.mark.parametrize("nn,niter,w", [(0, 5, 8), (33, 5, 8), (4, -1, 8), (4, 5, 0)])
def test_bad_arguments_are_refused(nn, niter, w):
    x = np.zeros((4, 8), np.float32)
    y = np.empty_like(x)
    rc = sp.lib.sp_amle_host(y.ctypes.data, x.ctypes.data, None, w, 4, 1, nn, niter, 0)
    assert rc == 1
    assert sp.lib.sp_last_error()


def test_back_to_back_calls_on_side_streams():
    """Two calls of one shape into different outputs, queued without a host
    synchronisation on a non-default stream, then a third on another
    stream: the instance tables are rebound in stream order and the shared
    scratch is not reused before the previous call's kernels are done.
    Every output equals the oracle bit for bit."""
    import torch
    rng = np.random.default_rng(77)
    xs = []
    for _ in range(3):
        x = (rng.standard_normal((257, 389)) * 30).astype(np.float32)
        x[rng.random(x.shape) < 0.6] = np.nan
        xs.append(x)
    dx = [dev(x) for x in xs]
    outs = [torch.empty_like(d) for d in dx]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    print("amle on side streams", flush=True)
    with torch.cuda.stream(s1):
        sp.amle(dx[0], 4, 60, out=outs[0])
        sp.amle(dx[1], 4, 60, out=outs[1])
    with torch.cuda.stream(s2):
        sp.amle(dx[2], 4, 60, out=outs[2])
    torch.cuda.synchronize()
    for x, o in zip(xs, outs):
        assert bits_equal(o.cpu().numpy(), oracle_amle(x, 4, 60))
