"""The oracle (oracle/simpois_oracle.c) pinned against the reference.

1. every golden fixture (written by the reference binary) is reproduced bit
   for bit, CG traces included;
2. random cases agree bit for bit with the reference compiled as a library
   (oracle/_ref, built from the reference sources where they are mounted);
3. closed-form facts that a dropped term, a sign error or a transposed index
   would break (see test_oracle_math.py).
"""
import ctypes
import os

import numpy as np
import pytest

from helpers import (REF_SO, bits_equal, expected_planar, flag_values,
                     knobs_of, load_golden, oracle_solve, planar_inputs, ptr,
                     V, I, F)

CASES = load_golden()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_golden(case):
    assert case["rc"] == 0
    g, f = planar_inputs(case)
    t, n, s, c = flag_values(case["flags"])
    out, trace = oracle_solve(g, f, t, n, s, c, trace_cap=100000,
                              **knobs_of(case["env"]))
    assert bits_equal(out, expected_planar(case))
    # the reference prints rr_new with %g (minicg.c:46)
    assert len(trace) == case["trace_lines"]
    for (_, _, rr), want in zip(trace, case["trace"]):
        if np.isnan(want):
            assert np.isnan(rr)
        else:
            assert float(f"{rr:g}") == want


def _ref():
    r = ctypes.CDLL(REF_SO)
    r.poisson_solver_separable.argtypes = [V, V, V, I, I, I, F, I, I, F]
    return r


@pytest.mark.skipif(not os.path.exists(REF_SO),
                    reason="reference library not built (make -C oracle ref)")
@pytest.mark.parametrize("seed,w,h,pd,t,n,s,c,withf", [
    (1, 31, 17, 1, 0.25, 10, 99, 0, False),
    (2, 40, 33, 2, 0.4, 13, 99, 0, True),
    (3, 29, 35, 1, -0.08, 9, 99, 0, False),
    (4, 26, 20, 1, -0.05, 6, 3, 0, True),
    (5, 18, 15, 1, -1000.02, 3, 99, 0, True),
    (6, 37, 30, 1, 0.25, 4, 99, 33, True),
    (7, 37, 30, 1, -0.08, 2, 99, 21, False),
    (8, 64, 1, 1, 0.25, 20, 99, 0, False),
    (9, 1, 50, 1, -0.08, 20, 99, 0, True),
])
def test_oracle_matches_reference_library(seed, w, h, pd, t, n, s, c, withf):
    rng = np.random.default_rng(seed)
    g = (rng.random((pd, h, w)) * 100).astype(np.float32)
    g[rng.random((pd, h, w)) < 0.35] = np.nan
    f = rng.standard_normal((pd, h, w)).astype(np.float32) if withf else None
    ours, _ = oracle_solve(g, f, t, n, s, c)
    ref = np.zeros_like(g)
    fd = os.dup(2)
    nul = os.open(os.devnull, os.O_WRONLY)
    os.dup2(nul, 2)   # the reference prints its CG trace on stderr
    try:
        _ref().poisson_solver_separable(ptr(ref), ptr(g), ptr(f), w, h, pd, t,
                                        n, s, float(c))
    finally:
        os.dup2(fd, 2)
        os.close(fd)
        os.close(nul)
    assert bits_equal(ours, ref)
