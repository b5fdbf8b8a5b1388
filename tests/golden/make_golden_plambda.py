"""Write the golden fixtures of plambda's 3x3 Laplacian operators.

Runs the reference plambda built from the reference sources by
``oracle/Makefile`` (``oracle/_ref/plambda``, plain ``cc -O3``) as
``plambda in.npy "x,l"`` (and ``"x,lr"``) under each ``PLAMBDA_GETPIXEL``
boundary rule, over seeded inputs that include NaN, +-Inf, signed zeros,
degenerate sizes and several channels, and stores inputs, operator, rule
and the output in ``tests/golden/plambda_lap.npz``.  plambda keeps
channels interleaved (h, w, pd); the fixture stores them that way.

Needs the reference tree, so it runs where that tree is mounted; the tests
only read the fixture file it writes.

    python tests/golden/make_golden_plambda.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "plambda")

SHAPES = [(1, 1, 1), (1, 7, 1), (6, 1, 1), (2, 2, 1), (9, 13, 1), (17, 11, 3),
          (32, 40, 1)]


def image(rng, h, w, pd, special):
    x = (rng.standard_normal((h, w, pd)) * 50).astype(np.float32)
    ints = rng.random((h, w, pd)) < 0.2
    x[ints] = np.round(x[ints])                  # exactly representable values
    if special:
        flat = x.reshape(-1)
        k = rng.choice(flat.size, size=min(4, flat.size), replace=False)
        vals = [np.nan, np.inf, -np.inf, -0.0]
        for t, v in zip(k, vals):
            flat[t] = v
    return x


def main():
    if not os.path.exists(REF):
        sys.exit(f"build the reference first: make -C oracle ref ({REF} missing)")
    rng = np.random.default_rng(2024)
    ins, outs, meta = [], [], []
    with tempfile.TemporaryDirectory() as td:
        for shape in SHAPES:
            for special in (False, True):
                x = image(rng, *shape, special)
                for op in ("x,l", "x,lr"):
                    for mode in range(5):
                        np.save(os.path.join(td, "in.npy"), x)
                        env = dict(os.environ, PLAMBDA_GETPIXEL=str(mode))
                        env.pop("GETPIXEL", None)
                        subprocess.run([REF, os.path.join(td, "in.npy"), op, "-o",
                                        os.path.join(td, "out.npy")], env=env,
                                       check=True, stdout=subprocess.DEVNULL,
                                       stderr=subprocess.DEVNULL)
                        y = np.load(os.path.join(td, "out.npy"))
                        ins.append(x)
                        outs.append(y.reshape(x.shape).astype(np.float32))
                        meta.append((op == "x,lr", mode))
    arrays = {}
    for k, (x, y) in enumerate(zip(ins, outs)):
        arrays[f"in{k}"] = x
        arrays[f"out{k}"] = y
    arrays["meta"] = np.array(meta, dtype=np.int32)   # (variant, mode) per case
    np.savez_compressed(os.path.join(HERE, "plambda_lap.npz"), **arrays)
    print(f"{len(ins)} cases")


if __name__ == "__main__":
    main()
