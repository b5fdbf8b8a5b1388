"""Write the golden fixtures of the simpois path.

Runs the reference program built from the reference sources by
``oracle/Makefile`` (``oracle/_ref/simpois``, plain ``cc -O3`` as in the
reference's top Makefile) over a matrix of seeded inputs and flags, and
stores inputs, flags, environment, exit code, output file bytes and the
CG trace (``minicg.c:46``) of each case in ``tests/golden/cases.npz``.

Needs the reference tree, so it runs where that tree is mounted; the GPU
tests only read the fixture file it writes.

    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "simpois")


def sinusoids(rng, h, w, k=4):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w))
    for _ in range(k):
        th = rng.uniform(0, 2 * np.pi)
        per = rng.uniform(4, 32)
        ph = rng.uniform(0, 2 * np.pi)
        img += np.sin(2 * np.pi * (x * np.cos(th) + y * np.sin(th)) / per + ph)
    img -= img.min()
    return (255 * img / max(img.max(), 1e-9)).astype(np.float32)


def discs(rng, h, w, n, rmin, rmax):
    y, x = np.mgrid[0:h, 0:w]
    m = np.zeros((h, w), bool)
    for _ in range(n):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        r = rng.uniform(rmin, rmax)
        m |= (y - cy) ** 2 + (x - cx) ** 2 < r * r
    return m


def make_cases():
    rng = np.random.default_rng(20241202)
    cases = []

    def add(name, g, flags, f=None, mask=None, env=None, fdash=False):
        cases.append(dict(name=name, g=g, f=f, mask=mask, flags=flags,
                          env=env or {}, fdash=fdash))

    def planes(pd, h, w, frac=0.3, border=True):
        g = np.stack([sinusoids(rng, h, w) for _ in range(pd)], axis=-1)
        holes = discs(rng, h, w, max(1, (h * w) // 120), 1, max(2, min(h, w) / 6))
        holes |= rng.random((h, w)) < frac * 0.2
        if border:
            holes[0, : w // 3] = True
            holes[:, -1] = True
        g[holes] = np.nan
        return g

    g1 = planes(1, 48, 40)
    add("default", g1, [])
    add("n1", g1, ["-n", "1"])
    add("n300", planes(1, 24, 20), ["-n", "300"])
    add("s1", g1, ["-s", "1"])
    add("s3", g1, ["-s", "3", "-n", "25"])
    add("t04", g1, ["-t", "0.4", "-n", "12"])
    add("t01_n2p7", g1, ["-t", "0.1", "-n", "2.7"])
    add("biharm", g1, ["-t", "-0.08"])
    add("biharm_n30", planes(1, 37, 29), ["-t", "-0.08", "-n", "30"])
    add("triharm", planes(1, 26, 22), ["-t", "-1000.01", "-n", "6"])
    add("odd_13x9", planes(1, 9, 13), ["-n", "7"])
    add("one_pixel", np.full((1, 1, 1), np.nan, np.float32), [])
    add("col_1x7", planes(1, 7, 1, border=False), ["-n", "5"])
    add("row_7x1", planes(1, 1, 7, border=False), ["-n", "5"])
    add("rgb", planes(3, 30, 34), ["-n", "9"])
    gp = planes(1, 40, 44)
    fp = rng.normal(0, 2, gp.shape).astype(np.float32)
    add("poisson", gp, ["-n", "15"], f=fp)
    add("poisson_rgb", planes(3, 22, 26), ["-t", "0.3", "-n", "8"],
        f=rng.normal(0, 1, (22, 26, 3)).astype(np.float32))
    add("poisson_biharm", gp, ["-t", "-0.06", "-n", "8"], f=fp)
    add("pmsfac", gp, ["-n", "6"], f=fp, env={"PMSFAC": "2.5"})
    add("ponlit", planes(1, 32, 32), ["-n", "20"], env={"PONLIT": "16"})
    gm = np.stack([sinusoids(rng, 36, 30)] * 2, axis=-1)
    mk = (discs(rng, 36, 30, 6, 2, 6)).astype(np.float32)[..., None]
    add("mask", gm.copy(), ["-n", "8"], mask=mk)
    gi = planes(1, 20, 24)
    gi[3, 4, 0] = np.inf
    gi[10, 20, 0] = -np.inf
    add("inf_pixels", gi, ["-n", "6"])
    add("all_nan", np.full((12, 10, 1), np.nan, np.float32), [])
    add("no_nan", sinusoids(rng, 15, 11)[..., None], ["-n", "4"])
    gc = planes(1, 33, 29)
    add("cg_poisson", gc, ["-n", "3", "-c", "40"])
    add("cg_poisson_f", gp, ["-n", "2", "-c", "25"], f=fp)
    add("cg_biharm", gc, ["-t", "-0.08", "-n", "3", "-c", "30"])
    add("cg_all", planes(1, 21, 19), ["-n", "0", "-c", "-1"])
    add("cg_eps", gc, ["-n", "2", "-c", "-1"], env={"CG_EPS": "1e-3"})
    # exact integer ramp: zero initial residual => 0/0 in CG (minicg.c:42)
    ramp = (np.arange(20, dtype=np.float32).reshape(4, 5))[..., None].copy()
    ramp[1, 2, 0] = ramp[2, 1, 0] = ramp[2, 3, 0] = np.nan
    add("ramp_default", ramp.copy(), [])
    add("ramp_cg_nan", ramp.copy(), ["-c", "10"])
    add("fdash", None, ["-n", "5"], f=rng.normal(0, 1, (16, 14, 1)).astype(np.float32),
        fdash=True)
    return cases


def run_case(c, tmp):
    ip = os.path.join(tmp, "in.npy")
    op = os.path.join(tmp, "out.npy")
    args = [REF]
    for q in os.listdir(tmp):
        os.unlink(os.path.join(tmp, q))
    stdin = None
    if c["g"] is not None:
        np.save(ip, c["g"])
        args += ["-i", ip]
    if c["f"] is not None:
        fpth = os.path.join(tmp, "f.npy")
        np.save(fpth, c["f"])
        if c["fdash"]:   # "-f -" with the default "-i -": f from stdin
            stdin = open(fpth, "rb").read()
            args += ["-f", "-"]
        else:
            args += ["-f", fpth]
    if c["mask"] is not None:
        mp = os.path.join(tmp, "m.npy")
        np.save(mp, c["mask"])
        args += ["-m", mp]
    args += c["flags"] + ["-o", op]
    env = dict(os.environ)
    for k in ("PMSFAC", "PONLIT", "CG_EPS"):
        env.pop(k, None)
    env.update(c["env"])
    p = subprocess.run(args, env=env, input=stdin, capture_output=True,
                       timeout=600)
    out = open(op, "rb").read() if os.path.exists(op) else b""
    return p.returncode, out, p.stderr.decode()


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    store = {}
    meta = []
    with tempfile.TemporaryDirectory() as tmp:
        for k, c in enumerate(make_cases()):
            rc, out, err = run_case(c, tmp)
            trace = [float(l.split("rr_new=")[1]) for l in err.splitlines()
                     if "rr_new=" in l]
            pre = f"c{k}_"
            if c["g"] is not None:
                store[pre + "g"] = c["g"]
            if c["f"] is not None:
                store[pre + "f"] = c["f"]
            if c["mask"] is not None:
                store[pre + "mask"] = c["mask"]
            store[pre + "out_bytes"] = np.frombuffer(out, np.uint8)
            store[pre + "trace"] = np.array(trace, np.float64)
            meta.append(dict(name=c["name"], flags=c["flags"], env=c["env"],
                             rc=rc, fdash=c["fdash"], has_g=c["g"] is not None,
                             has_f=c["f"] is not None,
                             has_mask=c["mask"] is not None,
                             trace_lines=len(trace)))
            print(f"{c['name']:16s} rc={rc} out={len(out)}B cg_lines={len(trace)}")
    store["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "cases.npz"), **store)
    print("wrote", os.path.join(HERE, "cases.npz"))


if __name__ == "__main__":
    main()
