"""Write the golden fixtures of amle (AMLE hole filling).

Runs the reference amle built from the reference sources by
``oracle/Makefile`` (``oracle/_ref/amle``, plain ``cc -O3``) as
``amle in.npy <mask> out.npy`` with ``AMLE_NN``, ``AMLE_NITER`` and
``AMLE_INIT`` set per case, over seeded single-channel inputs that include
degenerate sizes, images without holes or without known pixels, +-Inf and
signed zeros among the known values, an initial image, and a separate
mask file.  Stores input, mask, initial image, parameters, output and the
diagnostic lines printed on stderr in ``tests/golden/amle.npz``.

Needs the reference tree, so it runs where that tree is mounted; the tests
only read the fixture file it writes.

    python tests/golden/make_golden_amle.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "amle")

SHAPES = [(1, 1), (1, 9), (7, 1), (2, 2), (9, 13), (23, 31), (40, 32)]
NNS = [4, 6, 8, 16, 24, 32]


def image(rng, h, w, holes, special):
    x = (rng.standard_normal((h, w)) * 30).astype(np.float32)
    ints = rng.random((h, w)) < 0.2
    x[ints] = np.round(x[ints])
    x[rng.random((h, w)) < holes] = np.nan
    if special:
        known = np.flatnonzero(~np.isnan(x))
        k = rng.choice(known, size=min(2, known.size), replace=False)
        for t, v in zip(k, [-0.0, 0.0]):
            x.flat[t] = v
    return x


def run(td, x, nn, niter, init=None, mask=None):
    np.save(os.path.join(td, "in.npy"), x)
    env = dict(os.environ, AMLE_NN=str(nn), AMLE_NITER=str(niter))
    env.pop("AMLE_INIT", None)
    if init is not None:
        np.save(os.path.join(td, "init.npy"), init)
        env["AMLE_INIT"] = os.path.join(td, "init.npy")
    mname = ""
    if mask is not None:
        mname = os.path.join(td, "mask.npy")
        np.save(mname, mask)
    p = subprocess.run([REF, os.path.join(td, "in.npy"), mname,
                        os.path.join(td, "out.npy")], env=env, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                       text=True)
    lines = [ln for ln in p.stderr.splitlines() if ln.startswith("iter ")]
    y = np.load(os.path.join(td, "out.npy")).reshape(x.shape).astype(np.float32)
    return y, "\n".join(lines)


def main():
    if not os.path.exists(REF):
        sys.exit(f"build the reference first: make -C oracle ref ({REF} missing)")
    rng = np.random.default_rng(77)
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for (h, w) in SHAPES:
            for nn in NNS:
                for special in (False, True):
                    x = image(rng, h, w, 0.5, special)
                    niter = int(rng.choice([1, 3, 21]))
                    y, tr = run(td, x, nn, niter)
                    cases.append((x, None, None, nn, niter, y, tr))
        # edge cases
        x = np.full((5, 6), np.nan, np.float32)                  # no known pixel
        cases.append((x, None, None, 4, 3) + run(td, x, 4, 3))
        x = image(rng, 6, 5, 0.0, False)                         # no hole
        cases.append((x, None, None, 8, 3) + run(td, x, 8, 3))
        x = image(rng, 12, 14, 0.4, False)                       # infinities
        x[2, 3] = np.inf
        x[8, 1] = -np.inf
        x[0, 0] = np.nan
        for nn in (4, 8, 32):
            cases.append((x, None, None, nn, 5) + run(td, x, nn, 5))
        x = image(rng, 16, 20, 0.7, True)                        # AMLE_INIT
        init = (rng.standard_normal(x.shape) * 10).astype(np.float32)
        for nn in (4, 16):
            cases.append((x, None, init, nn, 7) + run(td, x, nn, 7, init=init))
        x = image(rng, 18, 15, 0.0, False)                       # mask file
        mask = np.where(rng.random(x.shape) < 0.5, 1.0, 0.0).astype(np.float32)
        mask[0, :3] = [-1.0, np.nan, 0.5]
        for nn in (4, 24):
            cases.append((x, mask, None, nn, 9) + run(td, x, nn, 9, mask=mask))
        x = image(rng, 33, 47, 0.8, True)                        # default NITER
        cases.append((x, None, None, 4, 100) + run(td, x, 4, 100))
    arrays = {}
    meta = []
    for k, (x, mask, init, nn, niter, y, tr) in enumerate(cases):
        arrays[f"in{k}"] = x
        arrays[f"out{k}"] = y
        if mask is not None:
            arrays[f"mask{k}"] = mask
        if init is not None:
            arrays[f"init{k}"] = init
        arrays[f"trace{k}"] = np.array(tr)
        meta.append((nn, niter))
    arrays["meta"] = np.array(meta, dtype=np.int32)     # (AMLE_NN, AMLE_NITER)
    np.savez_compressed(os.path.join(HERE, "amle.npz"), **arrays)
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()
