"""Child process of the GPU test of the FP64 CG's launch forms:
SIMPOIS_HIP_CG_PERSIST (0: three launches per iteration from captured
graphs; otherwise all iterations of a level in one persistent launch) is
read when the library loads, so the parent sets it in this process's
environment.  Runs FP64 CG fills (biharmonic and Poisson, the pyramid with
CG at every level, an iteration-limited solve, two planes) and writes every
output and r.r history to the .npz named by argv[1].

With SIMPOIS_HIP_GS_FLOW_SPIN=0 in the environment and argv[1] == "err",
checks instead that a grid barrier that gives up at once makes the solve
report an error (exit status 0 when it does)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np

from test_gpu_parity import random_case


def cases():
    """(name, g, f, tstep, cgit)"""
    g1, _ = random_case(61, 1, 700, 650, 0.3, False)
    g2, f2 = random_case(62, 1, 600, 710, 0.4, True)
    g3, f3 = random_case(63, 2, 300, 420, 0.5, True)
    return [("bih", g1, None, -0.08, -1), ("pois", g2, f2, 1.0, -1),
            ("lim", g2, f2, 1.0, 37), ("two", g3, f3, -0.08, -1)]


def main():
    import imscript_amd as m
    if sys.argv[1] == "err":
        g, f = random_case(64, 1, 900, 800, 0.4, True)
        try:
            m.simpois(g, f, tstep=-0.08, niter=0, scale=1, cgit=-1, trace=False)
            print("cg: no error", flush=True)
            return 1
        except Exception as e:   # the library's error, as the Python layer raises it
            print(f"cg: {type(e).__name__}: {e}", flush=True)
            return 0 if "barrier wait" in str(e) else 1
        finally:
            print("persistent launches (ran, gave up):", m.cg_persist_counts(), flush=True)
    res = {}
    for name, g, f, t, cgit in cases():
        out = m.simpois(g, f, tstep=t, niter=0, scale=99, cgit=cgit, trace=False)
        res[name] = out
        for l in range(g.shape[0]):
            for lev in range(12):
                tr = m.last_cg_trace(l, lev)
                if tr is None:
                    break
                res[f"{name}_rr_{l}_{lev}"] = np.array(tr[1], np.float64)
        print(f"{name}: {g.shape}", flush=True)
    ran, gave_up = m.cg_persist_counts()
    print(f"persistent launches: ran {ran}, gave up {gave_up}", flush=True)
    res["persist_ran"] = np.array([ran, gave_up], np.int64)
    np.savez(sys.argv[1], **res)
    return 0


if __name__ == "__main__":
    sys.exit(main())
