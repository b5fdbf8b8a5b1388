"""Benchmark: Poisson hole filling with a guidance field, 4096x4096x3 float32.

Workload (one step = one solve of the whole image):
  per channel a sum of 8 seeded random-direction sinusoids (periods
  16-256 px) scaled to [0, 255] plus N(0, 2) noise; guidance f = clamped
  5-point Laplacian of a second seeded image of the same kind; holes =
  random axis-aligned rectangles with sides 8-512 px until >= 50 % of the
  pixels.  Default method (--method mgpcg1): CG on the full-resolution
  masked system, preconditioned by one multigrid V-cycle per iteration
  (multiscale coarse-to-fine correction over a 2x2 aggregation hierarchy),
  to a relative residual target; --method mgpcg runs simpois's own pyramid
  (restrict g and f, coarse to fine: prolong, -n sweeps, the same CG at
  every level), --method cg the pyramid with plain CG, --method gs the
  reference's default Gauss-Seidel pyramid (parity path, roofline of the
  fused-sweep kernel).  Inputs are device-resident before timing starts.
  The FP64 relative residual ||b - A u|| / ||b|| of every channel is
  checked after the timed steps.

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

# HBM bytes per launch of the roofline kernels at the 4096x4096x3 workload,
# from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over
# `bench.py --steps 1 --warmup 0 --no-cpu --no-configs`, each scaled by what
# the same counter reads for k_stream_copy moving 1 GiB each way in that run
# (FETCH_SIZE x2.000, WRITE_SIZE x1.000).  Summary and commands:
# profiles/hbm_traffic_*.md.  The figures hold for the kernel source they were
# measured on: TRAFFIC_SRC is the SHA-256 prefix of imscript_amd/csrc/sp_cg.hip
# then, and the line reports traffic null when the source differs.
TRAFFIC_PMC = {
    "mg_up0": 633.9e6,
    "mg_down0": 495.3e6,
    "k32_dirspmv": 703.3e6,
}
TRAFFIC_SRC = "53d9b747b6fcb701"


def kernel_source_id(name="sp_cg.hip"):
    import hashlib
    with open(os.path.join(ROOT, "imscript_amd", "csrc", name), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]
# k_plambda_lap on the 4096x4096x3 planes, from the same two passes over
# `bench.py --config lap --no-cpu`, scaled by k_stream_copy moving the same
# 201 MB each way in that run (x2.000, x1.000): 221.0 MB read (the halo rows
# of 32-row strips) + 201.3 MB written.
TRAFFIC_PMC_LAP = 422.3e6
# heatd -n 100 on the 4096x4096x3 planes, per launch of the default
# multi-step kernels (k_heatd_multi<6>; with the field
# k_heatd_multi_field<5, 32>), same passes over `bench.py --config heatd
# [--heatd-field] --no-cpu --steps 1 --warmup 0` and the same scaling by the
# copy in that run: 403.1 + 217.4 MB and 990.4 + 224.1 MB.
TRAFFIC_PMC_HEATD = {"isotropic": 620.5e6, "field": 1214.5e6}


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
    make_inputs.source = v.contiguous()   # the guidance source image itself
    return g, f, float(holes.float().mean())


def cpu_model():
    """The host CPU's model name and logical CPU count (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            names = [ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else "unknown"
    except OSError:
        return "unknown"


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
    single = {"Mpix_s": round(side * side / dt / 1e6, 4), "seconds": round(dt, 3),
              "rel_residual": res}
    if kind != "reference":
        return {"value": single["Mpix_s"], "unit": "Mpix/s", "cores": 1,
                "cpu": cpu_model(), "kind": kind, "single_thread": single,
                "sample": f"channel 0, top-left {side}x{side} crop, CPU port of "
                          f"simpois -n 10 -c -1 with CG_EPS = {rel:g} * ||b|| of "
                          "the crop, single thread"}
    # throughput: P concurrent single-threaded reference processes on P
    # distinct crops (the project's own rule for CPU parallelism)
    P = max(1, min(16, os.cpu_count() or 1))
    pd_, H, W = g.shape
    crops = [(c, y, x) for c in range(pd_) for y in range(0, H - side + 1, side)
             for x in range(0, W - side + 1, side)][:P]
    with tempfile.TemporaryDirectory() as td:
        cmds = []
        for k, (c, y, x) in enumerate(crops):
            gk = np.ascontiguousarray(g[c, y:y + side, x:x + side].cpu().numpy())
            fk = np.ascontiguousarray(f[c, y:y + side, x:x + side].cpu().numpy())
            u0 = np.where(np.isnan(gk), 0, gk).astype(np.float32)
            oracle().orc_rel_residual(ptr(u0), ptr(gk), ptr(fk), side, side, 0,
                                      ctypes.byref(bn))
            gi, fi = os.path.join(td, f"g{k}.npy"), os.path.join(td, f"f{k}.npy")
            np.save(gi, gk)
            np.save(fi, fk)
            cmds.append(([ref, "-i", gi, "-f", fi, "-n", "10", "-c", "-1", "-o",
                          os.path.join(td, f"o{k}.npy")],
                         dict(os.environ, CG_EPS=repr(rel * bn.value), OMP_NUM_THREADS="1")))
        t0 = time.perf_counter()
        procs = [subprocess.Popen(c, env=e, stdout=subprocess.DEVNULL,
                                  stderr=subprocess.DEVNULL) for c, e in cmds]
        rcs = [q.wait() for q in procs]
        dtp = time.perf_counter() - t0
    if any(rcs):
        raise RuntimeError(f"reference processes failed: {rcs}")
    return {"value": round(len(crops) * side * side / dtp / 1e6, 4), "unit": "Mpix/s",
            "cores": len(crops), "cpu": cpu_model(), "kind": kind,
            "seconds": round(dtp, 3),
            "single_thread": single,
            "sample": f"{len(crops)} concurrent single-threaded reference processes, "
                      f"each on a distinct {side}x{side} crop of the input, simpois "
                      f"-n 10 -c -1 with CG_EPS = {rel:g} * ||b|| of its crop (the "
                      "reference's pyramid with plain CG at every level), file I/O "
                      "included; single_thread = channel 0 top-left crop alone"}


def fp64_check(torch, sp, params, seed, side=1024):
    """Error of the measured method against an FP64 direct solve (scipy) of
    a side x side instance of the same generator (one channel)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import adjoint_ref as ar
    g, f, _ = make_inputs(torch, side, 1, seed + 100, "cuda:0")
    if params.get("guidance"):
        f = sp.plambda_lap(make_inputs.source, params["guidance"] - 1, 1)
    out, _ = sp.solve_device(g, f, **{k: v for k, v in params.items() if k != "guidance"})
    torch.cuda.synchronize()
    gh, fh, uh = g[0].cpu().numpy(), f[0].cpu().numpy(), out[0].cpu().numpy()
    t0 = time.perf_counter()
    ex = ar.solve(gh, fh)
    dt = time.perf_counter() - t0
    hole = np.isnan(gh)
    err = np.abs(uh.astype(np.float64) - ex)[hole]
    return {"instance": f"{side}x{side}x1 of the same generator (seed {seed + 100})",
            "max_abs_err": float(err.max()), "rel_max_err": float(err.max() / np.abs(ex).max()),
            "rms_err": float(np.sqrt(np.mean(err ** 2))),
            "rel_residual": sp.residual_device(out, g, f)[0], "fp64_solve_s": round(dt, 2)}


# --------------------------------------------------- CPU legs of C1, C3-C5
#
# The oracle, the reference binary and the FP64 scipy solves are used here,
# on the host, as checkers and timed baselines only; --no-cpu skips them.

def _oracle_imports():
    for d in ("tests", "oracle"):
        if os.path.join(ROOT, d) not in sys.path:
            sys.path.insert(0, os.path.join(ROOT, d))
    from helpers import oracle_solve, bits_equal
    import adjoint_ref as ar
    return oracle_solve, bits_equal, ar


def cpu_leg_lap(v, out, side=2048):
    """The reference plambda on a side x side crop of channel 0 (file I/O
    included, one thread), bit-compared with the GPU result on that crop."""
    oracle_solve, bits_equal, ar = _oracle_imports()
    ref = os.path.join(ROOT, "oracle", "_ref", "plambda")
    if not os.path.exists(ref):
        return None
    x = np.ascontiguousarray(v[0, :side, :side].cpu().numpy())
    with tempfile.TemporaryDirectory() as td:
        np.save(os.path.join(td, "v.npy"), x)
        t0 = time.perf_counter()
        subprocess.run([ref, os.path.join(td, "v.npy"), "x,l", "-o",
                        os.path.join(td, "o.npy")], check=True,
                       env={k: e for k, e in os.environ.items()
                            if k not in ("PLAMBDA_GETPIXEL", "GETPIXEL")})
        dt = time.perf_counter() - t0
        y = np.load(os.path.join(td, "o.npy")).reshape(side, side)
    # the crop's own border is clamped by the reference: compare the interior
    got = out[0, 1:side - 1, 1:side - 1].cpu().numpy()
    return {"value": round(side * side / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1,
            "kind": "reference", "seconds": round(dt, 3),
            "bit_exact_interior": bits_equal(got, y[1:-1, 1:-1]),
            "sample": f"plambda v.npy \"x,l\" on the top-left {side}x{side} crop "
                      "of channel 0, file I/O included"}


def cpu_leg_c1(g, got):
    """C1 against the oracle port (bit-exact, timed), the reference binary
    (bit-exact, timed) and an FP64 direct solve (error)."""
    oracle_solve, bits_equal, ar = _oracle_imports()
    t1 = time.perf_counter()
    want, _ = oracle_solve(g, niter=2000)
    res = {"bit_exact_vs_oracle": bits_equal(got, want),
           "oracle_port_s": round(time.perf_counter() - t1, 3),
           "max_abs_err_vs_fp64_solve": float(np.max(np.abs(got - ar.solve(g))))}
    ref = os.path.join(ROOT, "oracle", "_ref", "simpois")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            np.save(os.path.join(td, "g.npy"), g)
            t1 = time.perf_counter()
            subprocess.run([ref, "-i", os.path.join(td, "g.npy"), "-n", "2000", "-o",
                            os.path.join(td, "o.npy")], check=True)
            res["reference_cpu_s"] = round(time.perf_counter() - t1, 3)
            res["reference_bitexact"] = bits_equal(
                np.load(os.path.join(td, "o.npy"))[..., 0], got)
    return res


def cpu_leg_c3(torch, sp, a, side=512):
    """Error of the C3 method against an FP64 direct solve of the 13-point
    system on a side x side instance of the same generator."""
    _, _, ar = _oracle_imports()
    rng = np.random.default_rng(a.seed + 33)
    g = grf(rng, side)
    g[stripes(rng, side, 0.2)] = np.nan
    dg = torch.from_numpy(g).cuda()
    p = dict(tstep=-0.08, niter=0, scale=99, cgit=-1, cg_eps=0.9e-6 * bnorm_bilap(dg),
             trace=0)
    out, _ = sp.solve_device(dg, **p)
    u = out.cpu().numpy().astype(np.float64)
    t0 = time.perf_counter()
    ex = ar.solve_bilap(g)
    hole = np.isnan(g)
    err = np.abs(u - ex)[hole]
    return {"instance": f"{side}x{side} of the same generator (seed {a.seed + 33})",
            "max_abs_err": float(err.max()), "rel_max_err": float(err.max() / np.abs(ex).max()),
            "rel_residual": sp.residual_device(out, dg, bilap=True)[0],
            "fp64_solve_s": round(time.perf_counter() - t0, 2)}


def cpu_leg_c4(gs, outs, k=8):
    """Error of the batched results against FP64 direct solves of the k
    smallest images."""
    _, _, ar = _oracle_imports()
    order = sorted(range(len(gs)), key=lambda i: gs[i].numel())[:k]
    errs = []
    for i in order:
        g = gs[i].cpu().numpy()
        ex = ar.solve(g)
        hole = np.isnan(g)
        e = np.abs(outs[i].cpu().numpy().astype(np.float64) - ex)[hole]
        errs.append(float(e.max() / np.abs(ex).max()) if e.size else 0.0)
    return {"images": f"the {k} smallest of the batch", "rel_max_err_max": max(errs),
            "rel_max_err_each": [round(e, 9) for e in errs]}


def cpu_leg_c5(torch, poisson_fill, dev):
    """Forward fill and gradient on a small instance against the FP64
    reference (direct solve and adjoint formulas)."""
    _, _, ar = _oracle_imports()
    rs = np.random.default_rng(9)
    gs = sinusoids_np(rs, 48, 56).astype(np.float64)
    gs[10:30, 12:40] = np.nan
    fs, ts = rs.standard_normal(gs.shape), rs.random(gs.shape) * 255
    dgs = torch.tensor(gs, dtype=torch.float32, device=dev).requires_grad_(True)
    dfs = torch.tensor(fs, dtype=torch.float32, device=dev).requires_grad_(True)
    u = poisson_fill(dgs, dfs, tol=1e-7)
    hs = torch.isnan(dgs.detach())
    ((u - torch.tensor(ts, dtype=torch.float32, device=dev)) ** 2)[hs].sum().backward()
    uref = ar.solve(gs, fs)
    v = np.where(np.isnan(gs), 2 * (uref - ts), 0)
    rg, rf = ar.gradient(gs, fs, v)
    uh = u.detach().cpu().numpy().astype(np.float64)
    return {"fwd_rel_err_vs_fp64": float(np.max(np.abs(uh - uref)) / np.max(np.abs(uref))),
            "grad_rel_err_vs_fp64_g": float(np.max(np.abs(dgs.grad.cpu().numpy() - rg))
                                            / np.max(np.abs(rg))),
            "grad_rel_err_vs_fp64_f": float(np.max(np.abs(dfs.grad.cpu().numpy() - rf))
                                            / np.max(np.abs(rf)))}


# ------------------------------------------------------------ other configs

def emit(d):
    print(json.dumps(d), flush=True)


def sinusoids_np(rng, h, w, k=8, pmin=16, pmax=256):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w))
    for _ in range(k):
        th, per, ph = rng.uniform(0, 2 * np.pi), rng.uniform(pmin, pmax), rng.uniform(0, 2 * np.pi)
        img += np.sin(2 * np.pi * (x * np.cos(th) + y * np.sin(th)) / per + ph)
    img -= img.min()
    return (255 * img / img.max()).astype(np.float32)


def c1_input(seed, n=512):
    """The 512x512 Laplace workload: a sum of 8 seeded sinusoids in [0, 255]
    with 40 discs (radii 5-40 px, scaled together until 29-31 % of the
    pixels are covered) set to NaN.  Returns (g, hole fraction)."""
    rng = np.random.default_rng(seed)
    g = sinusoids_np(rng, n, n)
    cy, cx = rng.uniform(0, n, 40), rng.uniform(0, n, 40)
    r0 = rng.uniform(5, 40, 40)
    y, x = np.mgrid[0:n, 0:n]
    lo, hi = 0.5, 2.0   # scale the radii until 29-31 % of the pixels are holes
    for _ in range(40):
        sc = 0.5 * (lo + hi)
        m = np.zeros((n, n), bool)
        for k in range(40):
            m |= (y - cy[k]) ** 2 + (x - cx[k]) ** 2 < (sc * r0[k]) ** 2
        if m.mean() < 0.29:
            lo = sc
        elif m.mean() > 0.31:
            hi = sc
        else:
            break
    g[m] = np.nan
    return g, float(m.mean())


def run_c1(a):
    """512x512 Laplace fill, 30% disc holes, simpois -n 2000 (bit-exact
    Gauss-Seidel on every pyramid level)."""
    import torch
    import imscript_amd as sp
    n = 512
    g, frac = c1_input(a.seed, n)
    dg = torch.from_numpy(g).cuda()
    out = torch.empty_like(dg)
    p = dict(tstep=0.25, niter=2000, scale=99, cgit=0, trace=0,
             autotune=0 if a.no_autotune else 1)
    sp.solve_device(dg, out=out, **p)   # warmup (and tile-shape tuning)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sp.solve_device(dg, out=out, **p)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    got = out.cpu().numpy()
    res = {"config": "c1", "workload": "512x512 Laplace, 30% disc holes, simpois -n 2000",
           "hole_fraction": round(frac, 4), "gpu_ms": round(dt * 1e3, 2),
           "gpu_Mpix_s": round(n * n / dt / 1e6, 3),
           "rel_residual": sp.residual_device(out, dg)[0]}
    if not a.no_cpu:
        res.update(cpu_leg_c1(g, got))
    return res


def grf(rng, n, power=3.0):
    """Gaussian random field with |k|^-power spectrum, scaled to [0, 255]."""
    w = rng.standard_normal((n, n))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-power / 2)
    amp[0, 0] = 0
    v = np.fft.irfft2(np.fft.rfft2(w) * amp, s=(n, n))
    v -= v.min()
    return (255 * v / v.max()).astype(np.float32)


def stripes(rng, n, frac):
    y, x = np.mgrid[0:n, 0:n].astype(np.float32)
    m = np.zeros((n, n), bool)
    while m.mean() < frac:
        wid = rng.uniform(2, 16)
        th = rng.uniform(0, np.pi)
        cy, cx = rng.uniform(0, n, 2)
        ln = rng.uniform(128, 1024)
        dx, dy = np.cos(th), np.sin(th)
        along = (x - cx) * dx + (y - cy) * dy
        across = -(x - cx) * dy + (y - cy) * dx
        m |= (np.abs(across) < wid / 2) & (np.abs(along) < ln / 2)
    return m


def run_c3(a):
    """2048x2048 biharmonic fill of a k^-3 Gaussian random field, 20% thin
    stripes; FP64 CG on the masked bilaplacian at every level to 1e-6*||b||
    (simpois -t -0.08 -n 0 -c -1 with CG_EPS)."""
    import torch
    import imscript_amd as sp
    rng = np.random.default_rng(a.seed + 3)
    n = a.c3_size
    g = grf(rng, n)
    m = stripes(rng, n, 0.2)
    g[m] = np.nan
    dg = torch.from_numpy(g).cuda()
    bn = bnorm_bilap(dg)
    out = torch.empty_like(dg)
    # the CG stops on its own recursive residual evaluated with the float
    # operator; a 10% margin keeps the FP64 residual under 1e-6
    p = dict(tstep=-0.08, niter=0, scale=99, cgit=-1, cg_eps=0.9e-6 * bn, trace=0)
    sp.solve_device(dg, out=out, **p)   # warmup (plan)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, st = sp.solve_device(dg, out=out, **p)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    rel = sp.residual_device(out, dg, bilap=True)[0]
    res = {"config": "c3", "workload": f"{n}x{n} biharmonic, GRF k^-3, 20% stripes, "
           "FP64 CG at every level to 1e-6*||b||", "hole_fraction": round(float(m.mean()), 4),
           "gpu_s": round(dt, 4), "gpu_Mpix_s": round(n * n / dt / 1e6, 3),
           "cg_iters_finest": st.cg_iters_finest, "rel_residual": rel}
    if not a.no_cpu:
        res["fp64_check"] = cpu_leg_c3(torch, sp, a)
    return res


def bnorm_bilap(dg):
    """||b|| of the biharmonic masked system, b = -A(known pixels)
    (cleant_cgpois.c:225-236 with the 13-point operator of :91-121)."""
    g = dg.cpu().numpy().astype(np.float64)
    h, w = g.shape
    st = [(0, 0, 20), (1, 0, -8), (0, 1, -8), (-1, 0, -8), (0, -1, -8), (1, 1, 2),
          (1, -1, 2), (-1, 1, 2), (-1, -1, 2), (2, 0, 1), (0, 2, 1), (-2, 0, 1), (0, -2, 1)]
    known = np.where(np.isnan(g), 0.0, g)
    yy, xx = np.mgrid[0:h, 0:w]
    acc = np.zeros((h, w))
    for di, dj, c in st:
        acc += c * known[np.clip(yy + dj, 0, h - 1), np.clip(xx + di, 0, w - 1)]
    b = -acc[np.isnan(g)]
    return float(np.sqrt(np.sum(b * b)))


def c4_images(torch, rng, count, dev):
    gs, fr = [], []
    amin, amax = np.log(128 * 128), np.log(2048 * 2048)
    for _ in range(count):
        area = np.exp(rng.uniform(amin, amax))
        asp = rng.uniform(1, 4)
        h = max(8, int(round(np.sqrt(area / asp))))
        w = max(8, int(round(area / h)))
        if rng.random() < 0.5:
            h, w = w, h
        yy = torch.arange(h, device=dev, dtype=torch.float32).view(h, 1)
        xx = torch.arange(w, device=dev, dtype=torch.float32).view(1, w)
        img = torch.zeros(h, w, device=dev)
        for _k in range(8):
            th, per, ph = rng.uniform(0, 2 * np.pi), rng.uniform(16, 256), rng.uniform(0, 2 * np.pi)
            img += torch.sin((xx * math.cos(th) + yy * math.sin(th)) * (2 * math.pi / per) + ph)
        img = (img - img.min()) / (img.max() - img.min()) * 255
        frac = rng.uniform(0.10, 0.60)
        m = torch.zeros(h, w, dtype=torch.bool, device=dev)
        while m.float().mean().item() < frac:
            if rng.random() < 0.5:
                r = rng.uniform(3, max(4, min(h, w) / 6))
                cy, cx = rng.uniform(0, h), rng.uniform(0, w)
                m |= (yy - cy) ** 2 + (xx - cx) ** 2 < r * r
            else:
                a_, b_ = rng.uniform(4, max(5, min(h, w) / 4), 2).astype(int)
                y0, x0 = rng.integers(0, h), rng.integers(0, w)
                m[y0:y0 + a_, x0:x0 + b_] = True
        img[m] = float("nan")
        gs.append(img.contiguous())
        fr.append(float(m.float().mean()))
    return gs, fr


def run_c4(a):
    """256 grayscale images, areas log-uniform 128^2..2048^2, aspect 1:1-1:4,
    10-60% holes; each solved to rel residual 1e-5: one batched call versus
    the plain per-image loop."""
    import torch
    import imscript_amd as sp
    rng = np.random.default_rng(a.seed + 4)
    gs, fr = c4_images(torch, rng, a.c4_count, "cuda:0")
    mpix = sum(g.numel() for g in gs) / 1e6
    p = dict(tstep=0.25, niter=0, scale=1, cgit=-1, cg_fp32=1, cg_rel=a.cg_rel,
             cg_mg=1)
    outs = [torch.empty_like(g) for g in gs]
    sp.solve_batch_device(gs, outs=outs, **p)       # warm-up (plan + graphs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, st = sp.solve_batch_device(gs, outs=outs, **p)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / a.steps
    _, _, ps = sp.solve_batch_device(gs, outs=outs, plane_stats=True, **p)
    relb = [q.rel_residual for q in ps]
    status = sorted(set(sp.PLANE_STATUS_NAMES[q.status] for q in ps))
    # plain loop: one solve call per image, each with its own cached plan
    # (the cache is widened to hold all of them), warmed once like the
    # batch and timed over the same number of steps
    sp.plan_cache_limit(4 * len(gs) + 8, 128 << 30)
    outl = [torch.empty_like(g) for g in gs]
    for g, o in zip(gs, outl):
        sp.solve_device(g, out=o, **p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for g, o in zip(gs, outl):
            sp.solve_device(g, out=o, **p)
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / a.steps
    rell = [sp.residual_device(o, g)[0] for o, g in zip(outl, gs)]
    same = all(torch.equal(x.view(torch.int32), y.view(torch.int32))
               for x, y in zip(outs, outl))
    del outl
    sp.plan_cache_clear()
    sp.plan_cache_limit()
    res = {"config": "c4", "workload": f"{len(gs)} images, {mpix:.1f} Mpx total, "
           "log-uniform 128^2-2048^2, aspect 1-4, 10-60% holes, rel residual 1e-5",
           "method": "multigrid-preconditioned FP32 CG per image",
           "batched_s": round(tb, 4), "batched_Mpix_s": round(mpix / tb, 1),
           "loop_s": round(tl, 4), "loop_Mpix_s": round(mpix / tl, 1),
           "speedup_batched_vs_loop": round(tl / tb, 2),
           "loop": "one sp_solve_device call per image, plans cached (warm), "
                   f"timed over {a.steps} passes like the batch",
           "loop_bits_equal_batch": bool(same),
           "plane_status": status,
           "max_rel_residual_batched": max(relb), "max_rel_residual_loop": max(rell),
           "hole_fraction_mean": round(float(np.mean(fr)), 3)}
    if not a.no_cpu:
        res["fp64_check"] = cpu_leg_c4(gs, outs)
    return res


def c5_inputs(torch, seed, B, n, dev):
    """The differentiable-fill workload: B seeded sinusoid images n x n in
    [0, 255] with random discs (radii 5-60 px) set to NaN until 35 % of each
    image; guidance f ~ N(0, 1); target ~ U[0, 255].  Returns (g, f, target)."""
    rng = np.random.default_rng(seed + 5)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    yy = torch.arange(n, device=dev, dtype=torch.float32).view(n, 1)
    xx = torch.arange(n, device=dev, dtype=torch.float32).view(1, n)
    g = torch.empty(B, n, n, device=dev)
    for k in range(B):
        img = torch.zeros(n, n, device=dev)
        for _ in range(8):
            th, per, ph = rng.uniform(0, 2 * np.pi), rng.uniform(16, 256), rng.uniform(0, 2 * np.pi)
            img += torch.sin((xx * math.cos(th) + yy * math.sin(th)) * (2 * math.pi / per) + ph)
        m = torch.zeros(n, n, dtype=torch.bool, device=dev)
        while m.float().mean().item() < 0.35:
            r = rng.uniform(5, 60)
            cy, cx = rng.uniform(0, n, 2)
            m |= (yy - cy) ** 2 + (xx - cx) ** 2 < r * r
        img = (img - img.min()) / (img.max() - img.min()) * 255
        img[m] = float("nan")
        g[k] = img
    f = torch.randn(B, n, n, device=dev, generator=gen)
    target = torch.rand(B, n, n, device=dev, generator=gen) * 255
    return g, f, target


def run_c5(a):
    """Batch 64 x 1024x1024, 35% disc holes: forward Poisson fill, loss on
    the filled pixels, backward by the adjoint solve."""
    import torch
    from imscript_amd.autograd import poisson_fill
    B, n = a.c5_batch, a.c5_size
    dev = "cuda:0"
    g, f, target = c5_inputs(torch, a.seed, B, n, dev)
    # loss on the filled pixels through a 0/1 weight (the same loss and
    # gradient as indexing by the mask, without its device-to-host sync)
    holef = torch.isnan(g).to(torch.float32)

    def step():
        gg = g.clone().requires_grad_(True)
        ff = f.clone().requires_grad_(True)
        u = poisson_fill(gg, ff, tol=a.cg_rel)
        loss = (((u - target) ** 2) * holef).sum()
        loss.backward()
        return gg.grad, ff.grad

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    res = {"config": "c5", "workload": f"batch {B} x {n}x{n}, 35% disc holes, "
           "forward fill + loss + adjoint backward",
           "step_s": round(dt, 4), "Mpix_s_fwd_bwd": round(B * n * n / dt / 1e6, 1)}
    if not a.no_cpu:
        res.update(cpu_leg_c5(torch, poisson_fill, dev))
    return res


def run_lap(a):
    """plambda's x,l Laplacian (the guidance-field producer) on the C2 source
    images, 4096x4096x3 planar float32: throughput, roofline of the kernel
    (8 B per pixel: read the image once, write the result once) and the
    reference plambda on a crop of one channel."""
    import torch
    import imscript_amd as sp
    make_inputs(torch, a.size, a.channels, a.seed, "cuda:0")
    v = make_inputs.source
    out = torch.empty_like(v)
    # one launch is about 0.1 ms: warm up and time many of them
    for _ in range(max(20, a.warmup)):
        sp.plambda_lap(v, 0, 1, out=out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(200, a.steps)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        sp.plambda_lap(v, 0, 1, out=out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    px = v.numel()
    gbps = 8.0 * px / (ms * 1e-3) / 1e9
    res = {"metric": "plambda x,l Laplacian, Mpix/s (guidance-field producer)",
           "value": round(px / 1e6 / (ms * 1e-3), 1), "unit": "Mpix/s", "n_gpus": 1,
           "steps": reps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"plambda \"x,l\" of {a.size}x{a.size}x{a.channels} "
                                  "float32 planar source images, nearest-pixel rule"},
           "roofline": {"kernel": "k_plambda_lap<0> (4 pixels per thread, 32-row strips)",
                        "bound": "hbm",
                        "achieved": round(gbps, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                        "frac": round(gbps / PEAK_HBM_GBPS, 4),
                        "traffic": (TRAFFIC_PMC_LAP if a.size == 4096 and a.channels == 3
                                    else None),
                        "algorithmic_bytes_per_launch": 8.0 * px,
                        "avg_launch_us": round(ms * 1e3, 2)}}
    # the plain copy kernel over the same bytes (one read and one write of
    # the planes): the measured ceiling at this size
    res["stream_copy_same_bytes"] = stream_copy_ceiling(torch, sp, nbytes=4 * px, reps=200)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_leg_lap(v, out)
    emit(res)


def cpu_leg_amle(x, sp, nn, niter, side):
    """The reference amle on a side x side crop of plane 0 (file I/O
    included, one thread), bit-compared with the GPU fill of the same crop."""
    ref = os.path.join(ROOT, "oracle", "_ref", "amle")
    if not os.path.exists(ref):
        return None
    _, bits_equal, _ = _oracle_imports()
    crop = np.ascontiguousarray(x[0, :side, :side].cpu().numpy())
    with tempfile.TemporaryDirectory() as td:
        np.save(os.path.join(td, "x.npy"), crop)
        env = dict(os.environ, AMLE_NN=str(nn), AMLE_NITER=str(niter))
        env.pop("AMLE_INIT", None)
        t0 = time.perf_counter()
        subprocess.run([ref, os.path.join(td, "x.npy"), "", os.path.join(td, "o.npy")],
                       check=True, env=env, stderr=subprocess.DEVNULL)
        dt = time.perf_counter() - t0
        y = np.load(os.path.join(td, "o.npy")).reshape(side, side)
    got = sp.amle(crop, nn, niter)
    return {"value": round(side * side / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1,
            "kind": "reference", "seconds": round(dt, 3),
            "bit_exact": bits_equal(got, y),
            "sample": f"amle x.npy on the top-left {side}x{side} crop of plane 0 "
                      f"(AMLE_NN={nn}, AMLE_NITER={niter}), file I/O included; "
                      "the reference also sweeps its lower bound"}


def run_amle(a):
    """AMLE hole filling (amle.c) of the C2 boundary images, 4096x4096 per
    plane, the reference's defaults AMLE_NN = 4 and AMLE_NITER = 100: fill
    throughput, the tile kernel's HBM roofline (12 B per pixel per pass of
    BT fused sweeps) and the reference amle on a crop."""
    import torch
    import imscript_amd as sp
    nn, niter = a.amle_nn, a.niter if a.niter is not None else 100
    g, _, hole_frac = make_inputs(torch, a.size, a.channels, a.seed, "cuda:0")
    g = g[:a.amle_planes].contiguous()
    out = torch.empty_like(g)
    for _ in range(max(1, a.warmup)):
        sp.amle(g, nn, niter, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sp.amle(g, nn, niter, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    sp.profile(True)
    sp.amle(g, nn, niter, out=out)
    torch.cuda.synchronize()
    t_ms, t_launch, t_bytes = sp.profile_read("amle_tile")
    p_ms, _, _ = sp.profile_read("amle_prep")
    sp.profile(False)
    px = g.numel()
    achieved = t_bytes / (t_ms * 1e-3) / 1e9
    eff = 12.0 * niter * px / (t_ms * 1e-3) / 1e9
    var = int(os.environ.get("SIMPOIS_HIP_AMLE_VARIANT", "6" if a.amle_planes == 1 else "0"))
    r1 = ["32x64x16", "32x128x16", "64x64x16", "16x128x32", "32x64x8",
          "64x32x16", "32x32x16", "16x64x16", "32x64x32", "64x64x8"]
    shape = (r1[var if 0 <= var < 10 else 0] if nn <= 4
             else "32x32x8" if nn <= 8 else "32x32x4")
    res = {"metric": "AMLE hole filling, Mpix/s (amle, AMLE_NN=%d, AMLE_NITER=%d)"
                     % (nn, niter),
           "value": round(px / 1e6 / dt, 1), "unit": "Mpix/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"amle of {a.amle_planes} plane(s) of "
                                  f"{a.size}x{a.size} float32, sum of sinusoids "
                                  f"with {hole_frac:.0%} rectangle holes 8-512 px, "
                                  f"{niter} raster sweeps of the {nn}-neighbour "
                                  "update from the maximum",
                      "output": "bit-identical to the reference amle (tsup); the "
                                "lower bound tinf, which feeds only the stderr "
                                "trace, is not swept"},
           "roofline": {"kernel": (("k_amle4_tile" if nn == 4 else "k_amle_tile") +
                                   f" (tile {shape} = BIxBJxBT: BT raster sweeps fused "
                                   "per pass, one launch per wavefront of skewed tiles)"
                                   if not gs_flow() else
                                   f"k_amle_flow<{'4-neighbour' if nn == 4 else 'table'}> "
                                   f"(tile {shape} = BIxBJxBT: BT raster sweeps fused per "
                                   "pass; one launch per pass, each skewed tile started "
                                   "once its predecessors have stored)"),
                        "bound": "hbm", "achieved": round(achieved, 1),
                        "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                        "algorithmic_bytes": t_bytes, "ms": round(t_ms, 3),
                        "launches": t_launch, "prep_ms": round(p_ms, 3),
                        "sweep_GBps_effective": round(eff, 1),
                        "sweep_frac_effective": round(eff / PEAK_HBM_GBPS, 4)}}
    # with the reference's trace (tinf swept too, a line every tenth sweep)
    import contextlib
    with open(os.devnull, "w") as dn, contextlib.redirect_stderr(dn):
        fd = os.dup(2)
        os.dup2(dn.fileno(), 2)
        try:
            t0 = time.perf_counter()
            sp.amle(g, nn, niter, out=out, trace=True)
            torch.cuda.synchronize()
            tt = time.perf_counter() - t0
        finally:
            os.dup2(fd, 2)
            os.close(fd)
    res["with_trace"] = {"ms": round(tt * 1e3, 3),
                         "Mpix_s": round(px / 1e6 / tt, 1),
                         "what": "tsup and tinf swept, diagnostics every tenth sweep "
                                 "printed as the reference does"}
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_leg_amle(g, sp, nn, niter, a.amle_cpu_side)
    emit(res)


def cpu_leg_heatd(x, sp, n, t, side):
    """The reference heatd on a side x side crop of plane 0 (file I/O
    included, one thread), bit-compared with the GPU run on the same crop."""
    ref = os.path.join(ROOT, "oracle", "_ref", "heatd")
    if not os.path.exists(ref):
        return None
    _, bits_equal, _ = _oracle_imports()
    crop = np.ascontiguousarray(x[0, :side, :side].cpu().numpy())
    with tempfile.TemporaryDirectory() as td:
        np.save(os.path.join(td, "x.npy"), crop)
        t0 = time.perf_counter()
        subprocess.run([ref, "-n", str(n), "-t", repr(t), os.path.join(td, "x.npy"),
                        os.path.join(td, "o.npy")], check=True)
        dt = time.perf_counter() - t0
        y = np.load(os.path.join(td, "o.npy")).reshape(side, side)
    got = sp.heatd(crop, n, t)
    return {"value": round(side * side / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1,
            "kind": "reference", "seconds": round(dt, 3), "bit_exact": bits_equal(got, y),
            "sample": f"heatd -n {n} -t {t} on the top-left {side}x{side} crop of "
                      "plane 0, file I/O included"}


def run_heatd(a):
    """heatd's explicit diffusion (heatd.c) of the C2 source images,
    4096x4096x3 planar float32, n steps: throughput of the whole diffusion,
    the HBM roofline of the kernel that runs (8 B per pixel per launch: read
    the plane, write the next; 12 B more with a diffusion-matrix field; the
    multi-step kernels run several steps per launch) and the reference
    heatd on a crop."""
    import torch
    import imscript_amd as sp
    n = a.niter if a.niter is not None else 100
    t = 0.25
    make_inputs(torch, a.size, a.channels, a.seed, "cuda:0")
    v = make_inputs.source
    A = None
    if a.heatd_field:
        rng = torch.Generator(device="cuda")
        rng.manual_seed(a.seed)
        A = torch.stack([0.2 + torch.rand(a.size, a.size, device="cuda", generator=rng),
                         torch.rand(a.size, a.size, device="cuda", generator=rng) - 0.5,
                         0.2 + torch.rand(a.size, a.size, device="cuda", generator=rng)],
                        -1).contiguous()
    out = torch.empty_like(v)
    for _ in range(max(1, a.warmup)):
        sp.heatd(v, n, t, A, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sp.heatd(v, n, t, A, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    sp.profile(True)
    sp.heatd(v, n, t, A, out=out)
    torch.cuda.synchronize()
    k_ms, k_launch, k_bytes = sp.profile_read("heatd_multi")
    multi = k_launch > 0
    if not multi:
        k_ms, k_launch, k_bytes = sp.profile_read("heatd_step")
    sp.profile(False)
    px = v.numel()
    per = k_bytes / max(k_launch, 1)
    achieved = per / (k_ms * 1e-3 / max(k_launch, 1)) / 1e9
    steps_per_launch = n / max(k_launch, 1)
    # the bytes a one-step-per-launch kernel would move for the n steps
    eff = (8.0 + (12.0 if A is not None else 0.0)) * px * n / (k_ms * 1e-3) / 1e9
    res = {"metric": "heatd diffusion, Mpix/s (n steps per pixel)",
           "value": round(px / 1e6 / dt, 1), "unit": "Mpix/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"heatd -n {n} -t {t}{' -A field' if A is not None else ''} "
                                  f"of {a.size}x{a.size}x{a.channels} float32 planar "
                                  "source images (sum of sinusoids)",
                      "Mpix_steps_per_s": round(px * n / 1e6 / dt, 1)},
           "roofline": {"kernel": ("k_heatd_multi%s<T> (%.3g steps per launch on "
                                   "register tiles of 64 columns per wave)"
                                   % ("_field" if A is not None else "", steps_per_launch))
                        if multi
                        else "k_heatd_step<%s> (4 pixels per thread, one launch per step)"
                        % ("field" if A is not None else "isotropic"),
                        "bound": "hbm", "achieved": round(achieved, 1),
                        "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBPS, 4),
                        "traffic": (TRAFFIC_PMC_HEATD["field" if A is not None else "isotropic"]
                                    if multi and a.size == 4096 and a.channels == 3
                                    and not os.environ.get("SIMPOIS_HIP_HEATD_T")
                                    and not os.environ.get("SIMPOIS_HIP_HEATD_FT")
                                    and not os.environ.get("SIMPOIS_HIP_HEATD_FR") else None),
                        "algorithmic_bytes_per_launch": per,
                        "avg_launch_us": round(k_ms * 1e3 / max(k_launch, 1), 2),
                        "launches": k_launch,
                        "step_GBps_effective": round(eff, 1),
                        "step_frac_effective": round(eff / PEAK_HBM_GBPS, 4)}}
    # the plain copy kernel over the plane bytes (one read, one write): the
    # measured ceiling at this size and the counters' known-size reference
    res["stream_copy_same_bytes"] = stream_copy_ceiling(torch, sp, nbytes=4 * px, reps=20)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_leg_heatd(v, sp, n, t, a.heatd_cpu_side)
    emit(res)


def stream_copy_ceiling(torch, sp, nbytes=1 << 30, reps=10):
    """Bandwidth of the library's plain copy kernel over nbytes (read +
    write counted), timed with events on the current stream: the measured
    ceiling next to the 8 TB/s nominal peak."""
    n = nbytes // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(2):
        sp.stream_copy(dst, src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        sp.stream_copy(dst, src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ok = bool(torch.equal(dst[:1024], src[:1024]))
    del src, dst
    gbps = 2.0 * nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "k_stream_copy (16-byte loads/stores, grid-stride)",
            "bytes_per_launch": 2 * nbytes, "avg_launch_us": round(ms * 1e3, 2),
            "GBps": round(gbps, 1), "frac_of_peak": round(gbps / PEAK_HBM_GBPS, 4),
            "copy_checked": ok}


def gs_flow():
    """Whether the library runs skewed tiles in dependence-ordered launches
    (SIMPOIS_HIP_GS_FLOW, read when it loads; on unless set to 0)."""
    return os.environ.get("SIMPOIS_HIP_GS_FLOW", "1") != "0"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--channels", type=int, default=3)
    ap.add_argument("--niter", type=int, default=None)
    ap.add_argument("--cg-rel-coarse", type=float, default=0.0,
                    help="CG tolerance of the pyramid levels above the finest")
    ap.add_argument("--scale", type=int, default=None,
                    help="pyramid levels (simpois -s); default 1 for mgpcg1, 99 otherwise")
    ap.add_argument("--method", default="mgpcg1", choices=["mgpcg", "mgpcg1", "cg", "gs"],
                    help="mgpcg: pyramid with multigrid-preconditioned CG at "
                         "every level; mgpcg1: the same on the finest level "
                         "only; cg: plain CG at every level (the reference's "
                         "method); gs: the reference's default, -n Gauss-Seidel "
                         "sweeps at every level of the pyramid and no CG (the "
                         "fused-sweep kernel's roofline; the residual is "
                         "reported, not required)")
    ap.add_argument("--target", type=float, default=1e-5)
    ap.add_argument("--cg-rel", type=float, default=5e-6)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-side", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="built-in Gauss-Seidel tile shapes instead of the "
                         "autotuner's choice (tuned during the warmup solve)")
    ap.add_argument("--guidance", action="store_true",
                    help="C2 from the source image: the solver computes the "
                         "data term as plambda's x,l Laplacian of it inside "
                         "the solve (plambda v x,l | simpois -f, fused)")
    ap.add_argument("--config", default="c2",
                    choices=["c1", "c2", "c3", "c4", "c5", "lap", "amle", "heatd"],
                    help="c2 (default) is the headline; the others print their "
                         "own JSON line")
    ap.add_argument("--heatd-field", action="store_true",
                    help="heatd with a per-pixel diffusion-matrix field (-A)")
    ap.add_argument("--heatd-cpu-side", type=int, default=2048)
    ap.add_argument("--amle-nn", type=int, default=4)
    ap.add_argument("--amle-planes", type=int, default=1)
    ap.add_argument("--amle-cpu-side", type=int, default=2048)
    ap.add_argument("--c3-size", type=int, default=2048)
    ap.add_argument("--c4-count", type=int, default=256)
    ap.add_argument("--c5-batch", type=int, default=64)
    ap.add_argument("--c5-size", type=int, default=1024)
    ap.add_argument("--no-configs", action="store_true",
                    help="the 4096x4096x3 line alone, without the summaries of "
                         "the other workloads (c1, c3, c4, c5, gs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {a.gpus} is not "
              "supported (a batch would shard by image and a large image by "
              "row slabs, one process per GPU; nothing here measures more than "
              "one)", file=sys.stderr)
        sys.exit(2)
    if a.config in ("c1", "c3", "c4", "c5"):
        emit({"c1": run_c1, "c3": run_c3, "c4": run_c4, "c5": run_c5}[a.config](a))
        return
    if a.config != "c2":
        {"lap": run_lap, "amle": run_amle, "heatd": run_heatd}[a.config](a)
        return
    result, ok, rel = run_c2(a, a.method, full=True)
    if not a.no_configs and a.method == "mgpcg1":
        # the other workloads, each timed over the same steps after its own
        # warm-up, summarized in this one line
        cfg = {}
        for name, fn in (("c1", run_c1), ("c3", run_c3), ("c4", run_c4), ("c5", run_c5)):
            t0 = time.perf_counter()
            cfg[name] = fn(a)
            cfg[name]["wall_s"] = round(time.perf_counter() - t0, 1)
            print(f"bench.py: {name} done in {cfg[name]['wall_s']} s", file=sys.stderr,
                  flush=True)
        t0 = time.perf_counter()
        gs_line, _, _ = run_c2(a, "gs", full=False)
        cfg["gs"] = {k: gs_line[k] for k in ("metric", "value", "unit", "ms_per_step")}
        cfg["gs"]["roofline"] = gs_line.get("roofline")
        cfg["gs"]["workload"] = "the 4096x4096x3 input, simpois -n 256 (parity Gauss-Seidel)"
        cfg["gs"]["wall_s"] = round(time.perf_counter() - t0, 1)
        result["configs"] = cfg
    print(json.dumps(result), flush=True)
    if not ok:
        print(f"bench.py: residual target missed: {rel}", file=sys.stderr)


def run_c2(a, method, full):
    """The 4096x4096x3 Poisson fill with guidance by `method`; full adds the
    copy ceiling, the host-buffer leg, the CPU baseline and the FP64 check.
    Returns (JSON dict, residual target met, residuals)."""
    import torch
    import imscript_amd as sp
    a = argparse.Namespace(**vars(a))
    a.method = method
    dev = "cuda:0"
    g, f, hole_frac = make_inputs(torch, a.size, a.channels, a.seed, dev)
    out = torch.empty_like(g)
    if a.niter is None:
        a.niter = {"cg": 10, "gs": 256}.get(a.method, 0)
    scale = a.scale if a.scale is not None else (1 if a.method == "mgpcg1" else 99)
    params = dict(tstep=0.25, niter=a.niter, scale=scale,
                  cgit=0 if a.method == "gs" else -1, cg_fp32=1, cg_rel=a.cg_rel,
                  trace=0, cg_chunk=32, cg_mg=0 if a.method == "cg" else 1,
                  autotune=0 if a.no_autotune else 1, cg_rel_coarse=a.cg_rel_coarse)
    torch.cuda.synchronize()

    f_term = f
    if a.guidance:
        params["guidance"] = 1
        f = make_inputs.source
        f_term = sp.plambda_lap(f, 0, 1)   # the data term the solve computes

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
    rel = sp.residual_device(out, g, f_term)
    ok = max(rel) <= a.target

    # phase timers on one more solve of the same workload
    sp.profile(True)
    step()
    torch.cuda.synchronize()
    cg_ms, cg_launch, cg_bytes = sp.profile_read("cg32")
    gs_ms, gs_launch, gs_bytes = sp.profile_read("gs_tile")
    kern = {k: sp.profile_read(k) for k in ("mg_up0", "mg_down0", "k32_dirspmv")}
    sp.profile(False)
    # the fused direction + product kernel per unknown: reads z, p_old and
    # the code, writes p_new and q (17 B); every other pass also reads x
    # (double) and the p of two passes before and writes x (20 B): 27 B on
    # average
    ms_, n_, _ = kern["k32_dirspmv"]
    kern["k32_dirspmv"] = (ms_, n_, 27.0 * hole_frac * a.size * a.size * a.channels * n_)

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
                "mgpcg1": "CG preconditioned by a multigrid V-cycle (multiscale "
                          "coarse-to-fine correction over a 2x2 aggregation "
                          "hierarchy) on the full-resolution system from a zero "
                          f"fill, to {a.cg_rel:g}*||b||, channels batched",
                "gs": f"simpois pyramid with -n {a.niter} parity-exact "
                      "lexicographic Gauss-Seidel sweeps at every level, no CG "
                      "(bit-identical to the reference's default path)",
            }[a.method],
            "hole_fraction": round(hole_frac, 4),
            "data_term": ("plambda x,l of the source image, computed inside the "
                          "solve" if a.guidance else "f given"),
            "levels": stats.levels,
            "cg_ms": round(cg_ms, 3),
            "cg_iters_finest_max": stats.cg_iters_finest,
            "rel_residual_max": max(rel),
            "residual_target_met": ok,
        },
    }
    if a.method == "gs":
        # the residual is a property of -n, not a target of this line
        result["metric"] = ("fused-sweep Gauss-Seidel solves, Mpix/s at -n "
                            f"{a.niter} (parity path); HBM GB/s % of peak")
        ok = True
    if cg_ms > 0 and a.method == "cg":
        achieved = cg_bytes / (cg_ms * 1e-3) / 1e9
        result["roofline"] = {
            "kernel": "CG iteration (k32_update, k32_dirspmv + 2 per-plane "
                      "reductions)",
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
            "algorithmic_bytes": cg_bytes, "ms": round(cg_ms, 3),
            "launches": cg_launch}
    if a.method != "cg" and kern["mg_up0"][0] > 0:
        rep = {}
        for k, (ms, n, by) in kern.items():
            rep[k] = {"avg_us": round(ms * 1e3 / max(n, 1), 2), "launches": n,
                      "GBps": round(by / (ms * 1e-3) / 1e9, 1)}
        dom = max(kern, key=lambda k: kern[k][0])
        ms, n, by = kern[dom]
        achieved = by / (ms * 1e-3) / 1e9
        result["roofline"] = {
            "kernel": {"mg_up0": "k_mg_up0 (fused level-0 coarse correction + "
                                 "10 red-black half sweeps + r.z)",
                       "mg_down0": "k_mg_down0 (fused level-0 10 red-black half "
                                   "sweeps + residual + restriction)",
                       "k32_dirspmv": "k32_dirspmv (fused CG direction + "
                                      "operator product + p.q)"}[dom],
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4),
            "traffic": (TRAFFIC_PMC.get(dom) if a.size == 4096 and a.channels == 3
                        and kernel_source_id() == TRAFFIC_SRC else None),
            "traffic_source": f"rocprofv3 --pmc on sp_cg.hip {TRAFFIC_SRC} "
                              f"(this build: {kernel_source_id()})",
            "algorithmic_bytes_per_launch": by / max(n, 1),
            "avg_launch_us": round(ms * 1e3 / max(n, 1), 2), "launches": n}
        result["kernels"] = rep
    if gs_ms > 0:
        result["gs_tile"] = {"ms": round(gs_ms, 3), "launches": gs_launch,
                             "pass_GBps": round(gs_bytes / (gs_ms * 1e-3) / 1e9, 1)}
    if gs_ms > 0 and a.method == "gs":
        # per-pass compulsory traffic: a pass of BT fused sweeps reads u and
        # the coefficient plane and writes u once per pixel (12 B); the
        # effective bandwidth counts those 12 B for every sweep of every
        # level swept in tiles (levels of more than 16384 pixels; smaller
        # ones run whole in one work-group)
        achieved = gs_bytes / (gs_ms * 1e-3) / 1e9
        tiled_px, w = 0, a.size
        for _ in range(stats.levels):
            if w * w > 16384:
                tiled_px += w * w * a.channels
            w = (w + 1) // 2
        eff = 12.0 * a.niter * tiled_px / (gs_ms * 1e-3) / 1e9
        tune = sp.autotune_report().strip().splitlines()
        finest = max(tune, key=lambda ln: int(ln.split("px=")[1].split()[0])) if tune else ""
        shape = ([c[1:].split(":")[0] for c in finest.split("choice=")[1].split(",")
                  if c.startswith("*")] or ["32x128x16"])[0] if finest else "32x128x16"
        result["roofline"] = {
            "kernel": (f"k_gs_flow<1> (tile {shape} = BIxBJxBT: BT lexicographic "
                       "sweeps fused per pass; one launch per level, each skewed "
                       "tile started once its predecessors have stored)" if gs_flow() else
                       f"k_gs_tile<1> (tile {shape} = BIxBJxBT: BT lexicographic "
                       "sweeps fused per pass, one launch per wavefront of skewed "
                       "tiles)"),
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
            "algorithmic_bytes": gs_bytes, "ms": round(gs_ms, 3),
            "launches": gs_launch,
            "sweep_GBps_effective": round(eff, 1),
            "sweep_frac_effective": round(eff / PEAK_HBM_GBPS, 4)}
        result["autotune"] = tune
    if not full:
        del g, f, out
        return result, ok, rel
    result["stream_copy"] = stream_copy_ceiling(torch, sp)
    # the drop-in boundary takes host buffers: one solve through it, copies
    # to and from the device included (reported beside `value`, never as it)
    gh, fh = g.cpu().numpy(), f.cpu().numpy()
    hp = {k: v for k, v in params.items()}
    sp.simpois(gh, fh, **hp)
    t0 = time.perf_counter()
    sp.simpois(gh, fh, **hp)
    th = time.perf_counter() - t0
    result["host_buffers"] = {"Mpix_s": round(mpix / th, 3), "ms": round(th * 1e3, 2),
                              "path": "sp_solve_host (the drop-in's host-buffer entry), "
                                      "H2D and D2H copies included"}
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(g, f_term, a.cg_rel, a.cpu_side, params)
        result["config"]["fp64_check"] = fp64_check(torch, sp, params, a.seed)
    del g, f, out
    return result, ok, rel


if __name__ == "__main__":
    main()
