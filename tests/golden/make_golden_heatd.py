"""Write the golden fixtures of heatd (explicit linear diffusion).

Runs the reference heatd built from the reference sources by
``oracle/Makefile`` (``oracle/_ref/heatd``, plain ``cc -O3``) as
``heatd -n N -t T [-A A.npy] in.npy out.npy`` over seeded inputs that
include NaN, +-Inf, signed zeros, degenerate sizes and several channels,
with and without a diffusion-matrix field, and stores input, field,
parameters and output in ``tests/golden/heatd.npz``.  Images are kept
interleaved (h, w, pd) as the files hold them; the field is (h, w, 3).

Needs the reference tree, so it runs where that tree is mounted; the tests
only read the fixture file it writes.

    python tests/golden/make_golden_heatd.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "heatd")

SHAPES = [(1, 1, 1), (1, 7, 1), (6, 1, 1), (2, 2, 1), (9, 13, 1), (17, 11, 3),
          (32, 40, 1), (23, 19, 2)]


def image(rng, h, w, pd, special):
    x = (rng.standard_normal((h, w, pd)) * 20).astype(np.float32)
    ints = rng.random((h, w, pd)) < 0.2
    x[ints] = np.round(x[ints])
    if special:
        flat = x.reshape(-1)
        k = rng.choice(flat.size, size=min(4, flat.size), replace=False)
        for t, v in zip(k, [np.nan, np.inf, -np.inf, -0.0]):
            flat[t] = v
    return x


def field(rng, h, w):
    a = rng.uniform(0.2, 1.5, (h, w))
    c = rng.uniform(0.2, 1.5, (h, w))
    b = rng.uniform(-0.8, 0.8, (h, w))
    A = np.stack([a, b, c], -1).astype(np.float32)
    A[rng.random((h, w)) < 0.1, 1] = 0.0           # some isotropic pixels
    return A


def main():
    if not os.path.exists(REF):
        sys.exit(f"build the reference first: make -C oracle ref ({REF} missing)")
    rng = np.random.default_rng(99)
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for shape in SHAPES:
            for special in (False, True):
                for aniso in (False, True):
                    if aniso and shape[0] == 1:
                        # iio reads a (1, w, 3) NPY as a 3 x w image: the
                        # reference refuses the field there
                        continue
                    x = image(rng, *shape, special)
                    A = field(rng, *shape[:2]) if aniso else None
                    n = int(rng.choice([1, 2, 5]))
                    t = float(rng.choice([0.25, 0.1, 0.37]))
                    np.save(os.path.join(td, "in.npy"), x)
                    args = [REF, "-n", str(n), "-t", repr(t)]
                    if A is not None:
                        np.save(os.path.join(td, "A.npy"), A)
                        args += ["-A", os.path.join(td, "A.npy")]
                    args += [os.path.join(td, "in.npy"), os.path.join(td, "out.npy")]
                    subprocess.run(args, check=True, stdout=subprocess.DEVNULL,
                                   stderr=subprocess.DEVNULL)
                    y = np.load(os.path.join(td, "out.npy")).reshape(x.shape)
                    cases.append((x, A, n, t, y.astype(np.float32)))
    arrays = {}
    meta = []
    for k, (x, A, n, t, y) in enumerate(cases):
        arrays[f"in{k}"] = x
        arrays[f"out{k}"] = y
        if A is not None:
            arrays[f"A{k}"] = A
        meta.append((n, t))
    arrays["meta"] = np.array(meta, dtype=np.float64)   # (n, t) per case
    np.savez_compressed(os.path.join(HERE, "heatd.npz"), **arrays)
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()
