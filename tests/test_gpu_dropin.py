"""The drop-in boundary as a reference-side C program uses it: a C caller
(tests/dropin/dropin_caller.c) includes include/simpois_hip.h, links
libsimpois_hip.so and calls poisson_solver_separable (simpois.c:288-302),
poisson_rec (:247-285), poisson_extension_by_cg and
biharmonic_extension_by_cg (cleant_cgpois.c:266-278) on host buffers.

Bars: the reference's golden outputs bit for bit for the Gauss-Seidel cases,
the CG bar (known pixels bit-exact, filled within 1e-4 relative max-abs) for
the CG cases; the single-level CG drop-ins against the oracle's restatement
of cleant_cgpois.c:197-264 on the same inputs and initial guess.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import (DROPIN_CALLER, bits_equal, close_cg, expected_planar,
                     flag_values, knobs_of, load_golden, oracle, planar_inputs,
                     ptr, Knobs)

pytestmark = pytest.mark.gpu

CASES = load_golden()
BY_NAME = {c["name"]: c for c in CASES}


def _env(extra):
    env = {k: v for k, v in os.environ.items()
           if k not in ("PMSFAC", "PONLIT", "CG_EPS", "SIMPOIS_HIP_QUIET")}
    env["SIMPOIS_HIP_QUIET"] = "1"
    env.update(extra)
    return env


def _call(tmp_path, args, env, arrays):
    """Write the arrays as raw float32 files, run the caller, read out."""
    names = []
    for k, a in enumerate(arrays):
        if a is None:
            names.append("-")
            continue
        p = tmp_path / f"a{k}.f32"
        np.ascontiguousarray(a, np.float32).tofile(p)
        names.append(str(p))
    out = tmp_path / "out.f32"
    fn, w, h, pd = args[:4]
    tail = [str(x) for x in args[4:]]
    cmd = [DROPIN_CALLER, fn, str(w), str(h), str(pd), names[0], names[1], str(out)]
    cmd += names[2:] + tail
    print("dropin", " ".join(cmd[1:5]), tail, flush=True)
    p = subprocess.run(cmd, env=env, capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    return np.fromfile(out, np.float32).reshape(pd, h, w)


def test_caller_is_built():
    assert os.path.exists(DROPIN_CALLER), "build the drop-in caller (__graft_entry__.build())"


@pytest.mark.parametrize("case", [c for c in CASES if c["rc"] == 0],
                         ids=[c["name"] for c in CASES if c["rc"] == 0])
def test_separable_on_golden(case, tmp_path):
    g, f = planar_inputs(case)
    pd, h, w = g.shape
    t, n, s, c = flag_values(case["flags"])
    out = _call(tmp_path, ["separable", w, h, pd, repr(float(t)), n, s, c],
                _env(case["env"]), [g, f])
    want = expected_planar(case)
    if c == 0:
        assert bits_equal(out, want)
    else:
        for l in range(pd):
            close_cg(out[l], want[l], g[l])


@pytest.mark.parametrize("name", ["default", "poisson", "biharm", "triharm", "pmsfac",
                                  "cg_poisson", "cg_biharm", "cg_all", "inf_pixels"])
def test_rec_on_golden(name, tmp_path):
    case = BY_NAME[name]
    g, f = planar_inputs(case)
    assert g.shape[0] == 1
    _, h, w = g.shape
    t, n, s, c = flag_values(case["flags"])
    out = _call(tmp_path, ["rec", w, h, 1, repr(float(t)), n, s, c],
                _env(case["env"]), [g, f])
    want = expected_planar(case)
    if c == 0:
        assert bits_equal(out, want)
    else:
        close_cg(out[0], want[0], g[0])


def _oracle_cg(g, f, init, maxit, eps, bilap, cg_eps=-1.0):
    h, w = g.shape
    out = np.empty_like(g)
    kn = Knobs(3.0, 0.0, cg_eps)
    oracle().orc_cg_extend(ptr(out), ptr(g), ptr(f), w, h, ptr(init), int(maxit),
                           float(eps), int(bilap), ctypes.byref(kn), None)
    return out


def _init_guess(g, seed):
    """An initial fill: the golden GS output where there is one, else a
    smooth random field (what poisson_rec hands the CG after its sweeps)."""
    rng = np.random.default_rng(seed)
    h, w = g.shape
    y, x = np.mgrid[0:h, 0:w]
    return (np.nanmean(g) + 5 * np.sin(x / 5.0) * np.cos(y / 7.0)
            + rng.normal(0, 0.5, (h, w))).astype(np.float32)


@pytest.mark.parametrize("name,fn,maxit,eps", [
    ("cg_poisson", "pcg", 40, 1e-6),
    ("cg_poisson_f", "pcg", 25, 1e-6),
    ("cg_all", "pcg", -1, 1e-6),
    ("cg_eps", "pcg", -1, 1e-3),
    ("poisson", "pcg", 200, 1e-4),
    ("cg_biharm", "bcg", 30, 1e-9),
    ("poisson_biharm", "bcg", 300, 1e-9),
    ("no_nan", "pcg", 10, 1e-6),          # no unknowns: cleant_cgpois.c:203-207
    ("inf_pixels", "pcg", 50, 1e-6),
])
def test_cg_extension_drop_ins(name, fn, maxit, eps, tmp_path):
    case = BY_NAME[name]
    g, f = planar_inputs(case)
    g, f = g[0], None if f is None else f[0]
    h, w = g.shape
    init = _init_guess(g, len(name))
    out = _call(tmp_path, [fn, w, h, 1, maxit, repr(eps)], _env({}),
                [g, f, init])[0]
    want = _oracle_cg(g, f, init, maxit, eps, fn == "bcg")
    if not np.isnan(g).any():
        assert bits_equal(out, want)
    else:
        close_cg(out, want, g)


def test_cg_extension_honours_cg_eps(tmp_path):
    """CG_EPS overrides the eps argument (cleant_cgpois.c:240 via smapa.h)."""
    case = BY_NAME["cg_poisson"]
    g, _ = planar_inputs(case)
    g = g[0]
    h, w = g.shape
    init = _init_guess(g, 3)
    out = _call(tmp_path, ["pcg", w, h, 1, -1, "1e-12"], _env({"CG_EPS": "0.5"}),
                [g, None, init])[0]
    want = _oracle_cg(g, None, init, -1, 1e-12, 0, cg_eps=0.5)
    close_cg(out, want, g)
