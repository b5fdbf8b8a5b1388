"""Child process of the GPU test of the Gauss-Seidel tile launch forms:
SIMPOIS_HIP_GS_FLOW (0: one launch per wavefront of tiles; otherwise one
dependence-ordered launch per level) is read when the library loads, so the
parent sets it in this process's environment.  Runs Gauss-Seidel solves of
each radius, a batch of planes, a multiscale pyramid and AMLE fills (which
take the same launch forms) against the oracle and prints one line per
case."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np

from helpers import bits_equal, oracle_amle, oracle_solve
from test_gpu_parity import random_case


def main():
    import imscript_amd as m
    cases = [(1, 300, 420, 0.25, 40, 1), (2, 260, 300, 0.25, 40, 99),
             (1, 200, 210, -0.08, 12, 1), (1, 150, 170, -1000.01, 6, 1),
             (3, 700, 530, 0.25, 33, 99)]
    for k, (pd, h, w, t, n, s) in enumerate(cases):
        g, f = random_case(80 + k, pd, h, w, 0.4, True)
        ref, _ = oracle_solve(g, f, t, n, s, 0)
        out = m.simpois(g, f, tstep=t, niter=n, scale=s, cgit=0, trace=False)
        ok = bits_equal(out, ref)
        print(f"case {pd}x{h}x{w} t={t} n={n} s={s}: {'OK' if ok else 'MISMATCH'}",
              flush=True)
        if not ok:
            return 1
    # AMLE on the same tile schedules (4- and 8-neighbour, one and two planes)
    for k, (pd, nn) in enumerate([(1, 4), (2, 4), (1, 8)]):
        rng = np.random.default_rng(90 + k)
        x = (rng.standard_normal((pd, 260, 330)) * 20).astype(np.float32)
        x[rng.random(x.shape) < 0.7] = np.nan
        y = m.amle(x, nn, 30)
        ok = all(bits_equal(y[l], oracle_amle(x[l], nn, 30)) for l in range(pd))
        print(f"amle {pd}x260x330 nn={nn}: {'OK' if ok else 'MISMATCH'}", flush=True)
        if not ok:
            return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
