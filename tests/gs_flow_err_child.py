"""Child process of the GPU test of the dependence-wait error path: with
SIMPOIS_HIP_GS_FLOW_SPIN=0 (read when the library loads) a tile whose
predecessor has not stored yet gives up at once, and the solve must report
the error instead of returning silently; AMLE, whose calls are asynchronous,
reports it on the next call.  Prints one line per check; exit status 0 when
every check saw its error."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np

from test_gpu_parity import random_case


def main():
    import imscript_amd as m
    g, f = random_case(7, 2, 900, 1000, 0.4, True)
    try:
        m.simpois(g, f, tstep=0.25, niter=64, scale=1, cgit=0, trace=False)
        print("gs: no error", flush=True)
        return 1
    except Exception as e:   # the library's error, as the Python layer raises it
        print(f"gs: {type(e).__name__}: {e}", flush=True)
        if "dependence wait" not in str(e):
            return 1
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((900, 1000)) * 20).astype(np.float32)
    x[rng.random(x.shape) < 0.7] = np.nan
    try:
        for _ in range(3):
            m.amle(x, 4, 64)
        print("amle: no error", flush=True)
        return 1
    except Exception as e:
        print(f"amle: {type(e).__name__}: {e}", flush=True)
        if "dependence wait" not in str(e):
            return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
