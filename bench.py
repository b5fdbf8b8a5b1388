"""Benchmark: Poisson hole filling with a guidance field, 4096x4096x3 float32.

Workload (one step = one solve of the whole image):
  per channel a sum of 8 seeded random-direction sinusoids (periods
  16-256 px) scaled to [0, 255] plus N(0, 2) noise; guidance f = clamped
  5-point Laplacian of a second seeded image of the same kind; holes =
  random axis-aligned rectangles with sides 8-512 px until >= 50 % of the
  pixels; simpois's multiscale solve (restrict g and f down the pyramid,
  coarse to fine: prolong, -n Gauss-Seidel sweeps, CG) with CG run to a
  relative residual target at every level.  Inputs are device-resident
  before timing starts.  The FP64 relative residual ||b - A u|| / ||b|| of
  every channel is checked after the timed steps.

Prints one JSON line.  --gpus above 1 is refused: this path runs on one GPU.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


def sinusoid_sum(torch, rng, n, dev, k=8):
    y = torch.arange(n, device=dev, dtype=torch.float32).view(n, 1)
    x = torch.arange(n, device=dev, dtype=torch.float32).view(1, n)
    img = torch.zeros(n, n, device=dev, dtype=torch.float32)
    for _ in range(k):
        th, per, ph = rng.uniform(0, 2 * np.pi), rng.uniform(16, 256), rng.uniform(0, 2 * np.pi)
        img += torch.sin((x * math.cos(th) + y * math.sin(th)) * (2 * math.pi / per) + ph)
    return (img - img.min()) / (img.max() - img.min()) * 255


def rect_holes(rng, n, frac=0.5):
    m = np.zeros((n, n), bool)
    while m.mean() < frac:
        a, b = rng.uniform(8, 512, 2).astype(int)
        y0, x0 = rng.integers(0, n, 2)
        m[y0:y0 + a, x0:x0 + b] = True
    return m


def make_inputs(torch, n, pd, seed, dev):
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    g = torch.stack([sinusoid_sum(torch, rng, n, dev)
                     + 2 * torch.randn(n, n, device=dev, generator=gen)
                     for _ in range(pd)]).contiguous()
    v = torch.stack([sinusoid_sum(torch, rng, n, dev) for _ in range(pd)])
    vp = torch.nn.functional.pad(v[None], (1, 1, 1, 1), mode="replicate")[0]
    f = (vp[:, 1:-1, 2:] + vp[:, 2:, 1:-1] + vp[:, 1:-1, :-2] + vp[:, :-2, 1:-1]
         - 4 * v).contiguous()
    holes = torch.from_numpy(rect_holes(rng, n)).to(dev)
    g[:, holes] = float("nan")
    return g, f, float(holes.float().mean())


def cpu_baseline(g, f, rel, side, params):
    """Reference CPU path on a bounded sample: channel 0, top-left side^2
    crop, same flags and residual target, one core."""
    ref = os.path.join(ROOT, "oracle", "_ref", "simpois")
    gc = np.ascontiguousarray(g[0, :side, :side].cpu().numpy())
    fc = np.ascontiguousarray(f[0, :side, :side].cpu().numpy())
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle, ptr, oracle_solve
    bn = ctypes.c_double()
    u0 = np.where(np.isnan(gc), 0, gc).astype(np.float32)
    oracle().orc_rel_residual(ptr(u0), ptr(gc), ptr(fc), side, side, 0, ctypes.byref(bn))
    eps = rel * bn.value
    if os.path.exists(ref):
        kind = "reference"
        with tempfile.TemporaryDirectory() as td:
            np.save(os.path.join(td, "g.npy"), gc)
            np.save(os.path.join(td, "f.npy"), fc)
            env = dict(os.environ, CG_EPS=repr(eps))
            t0 = time.perf_counter()
            subprocess.run([ref, "-i", os.path.join(td, "g.npy"), "-f",
                            os.path.join(td, "f.npy"), "-n", "10",
                            "-c", "-1", "-o", os.path.join(td, "o.npy")],
                           env=env, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            dt = time.perf_counter() - t0
            out = np.load(os.path.join(td, "o.npy"))[..., 0]
    else:
        kind = "port"
        t0 = time.perf_counter()
        out, _ = oracle_solve(gc, fc, niter=10, cgit=-1, cg_eps=eps)
        dt = time.perf_counter() - t0
    res = oracle().orc_rel_residual(ptr(np.ascontiguousarray(out, np.float32)),
                                    ptr(gc), ptr(fc), side, side, 0, None)
    return {"value": round(side * side / dt / 1e6, 4), "unit": "Mpix/s", "cores": 1,
            "kind": kind, "seconds": round(dt, 3), "rel_residual": res,
            "sample": f"channel 0, top-left {side}x{side} crop, simpois -n 10 "
                      f"-c -1 with CG_EPS = {rel:g} * ||b|| of the crop (the "
                      "reference's pyramid with plain CG at every level), single "
                      "thread, file I/O included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--channels", type=int, default=3)
    ap.add_argument("--niter", type=int, default=None)
    ap.add_argument("--method", default="mgpcg", choices=["mgpcg", "mgpcg1", "cg"],
                    help="mgpcg: pyramid with multigrid-preconditioned CG at "
                         "every level; mgpcg1: the same on the finest level "
                         "only; cg: plain CG at every level (the reference's "
                         "method)")
    ap.add_argument("--target", type=float, default=1e-5)
    ap.add_argument("--cg-rel", type=float, default=5e-6)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-side", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {a.gpus} is not "
              "supported (see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
    import imscript_amd as sp
    dev = "cuda:0"
    g, f, hole_frac = make_inputs(torch, a.size, a.channels, a.seed, dev)
    out = torch.empty_like(g)
    if a.niter is None:
        a.niter = 10 if a.method == "cg" else 0
    params = dict(tstep=0.25, niter=a.niter, scale=1 if a.method == "mgpcg1" else 99,
                  cgit=-1, cg_fp32=1, cg_rel=a.cg_rel, trace=0, cg_chunk=32,
                  cg_mg=0 if a.method == "cg" else 1)
    torch.cuda.synchronize()

    def step():
        return sp.solve_device(g, f, out=out, **params)[1]

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = None
    for _ in range(a.steps):
        stats = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    rel = sp.residual_device(out, g, f)
    ok = max(rel) <= a.target

    # phase timers on one more solve of the same workload
    sp.profile(True)
    step()
    torch.cuda.synchronize()
    cg_ms, cg_launch, cg_bytes = sp.profile_read("cg32")
    gs_ms, gs_launch, gs_bytes = sp.profile_read("gs_tile")
    sp.profile(False)

    mpix = a.size * a.size * a.channels / 1e6
    result = {
        "metric": "solved Mpix/s at rel residual 1e-5 (batched vs plain GPU); "
                  "HBM GB/s % of peak",
        "value": round(mpix / dt, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"Poisson fill with guidance, {a.size}x{a.size}x{a.channels} "
                        "float32, 50% rectangle holes 8-512 px, multiscale "
                        "coarse-to-fine to rel residual 1e-5",
            "method": {
                "cg": f"simpois pyramid, -n {a.niter} Gauss-Seidel sweeps and CG "
                      f"(float vectors, double reductions) to {a.cg_rel:g}*||b|| "
                      "at every level, all channels batched",
                "mgpcg": f"simpois pyramid (-n {a.niter}) with CG preconditioned "
                         "by an aggregation multigrid V-cycle to "
                         f"{a.cg_rel:g}*||b|| at every level, channels batched",
                "mgpcg1": "multigrid-preconditioned CG on the finest level from "
                          f"a zero fill, to {a.cg_rel:g}*||b||, channels batched",
            }[a.method],
            "hole_fraction": round(hole_frac, 4),
            "levels": stats.levels,
            "cg_ms": round(cg_ms, 3),
            "cg_iters_finest_max": stats.cg_iters_finest,
            "rel_residual_max": max(rel),
            "residual_target_met": ok,
        },
    }
    if cg_ms > 0 and a.method == "cg":
        achieved = cg_bytes / (cg_ms * 1e-3) / 1e9
        result["roofline"] = {
            "kernel": "CG iteration (k32_spmv, k32_update, k32_dir + 2 per-plane "
                      "reductions)",
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
            "algorithmic_bytes": cg_bytes, "ms": round(cg_ms, 3),
            "launches": cg_launch}
    if gs_ms > 0:
        result["gs_tile"] = {"ms": round(gs_ms, 3), "launches": gs_launch,
                             "pass_GBps": round(gs_bytes / (gs_ms * 1e-3) / 1e9, 1)}
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(g, f, a.cg_rel, a.cpu_side, params)
    print(json.dumps(result), flush=True)
    if not ok:
        print(f"bench.py: residual target missed: {rel}", file=sys.stderr)


if __name__ == "__main__":
    main()
