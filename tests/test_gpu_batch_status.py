"""Per-plane outcome of a batched solve (sp_solve_batch_device_ex), failure
isolation between the planes of a batch, the allocation cap and the plan
cache, through the C ABI.

A plane that cannot be solved (NaN in its data term, nothing to solve, an
exact-zero residual, an iteration limit) must be flagged and must leave the
bits of the other planes exactly as when they are solved alone.
"""
import numpy as np
import pytest

from helpers import bits_equal, oracle_solve

pytestmark = pytest.mark.gpu


def _m():
    import imscript_amd
    return imscript_amd


def _healthy(seed, h, w, frac=0.4):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    g = (100 * np.sin(x / rng.uniform(5, 30)) * np.cos(y / rng.uniform(5, 30))
         + rng.normal(0, 1, (h, w))).astype(np.float32)
    g[rng.random((h, w)) < frac * 0.5] = np.nan
    g[h // 4:h // 2, w // 3:2 * w // 3] = np.nan
    f = rng.normal(0, 1, (h, w)).astype(np.float32)
    return g, f


def _ramp(h=5, w=6):
    """The ramp j - i with holes on its zero diagonal: the initial guess 0
    is the exact fill, so the CG's initial residual is exactly 0
    (minicg.c:40-42 turns that into 0/0 = NaN in the reference)."""
    g = (np.arange(w)[None, :] - np.arange(h)[:, None]).astype(np.float32)
    want = g.copy()
    g[1, 1] = g[2, 2] = g[3, 3] = np.nan
    return g, np.zeros_like(g), want


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bits(a, b):
    import torch
    return bool(torch.equal(a.view(torch.int32), b.view(torch.int32)))


@pytest.mark.parametrize("mg", [0, 1])
def test_mixed_batch_status_and_isolation(mg):
    import torch
    m = _m()
    hg, hf = _healthy(1, 90, 130)
    hg2, hf2 = _healthy(2, 64, 48)
    bad_g, bad_f = _healthy(3, 70, 60)
    bad_f[np.isnan(bad_g)] = np.nan * bad_f[np.isnan(bad_g)]   # NaN data at every hole
    rg, rf, rwant = _ramp()
    ng = np.full((20, 24), np.nan, np.float32)        # no known pixel at all
    nf = np.zeros_like(ng)
    kg, kf = _healthy(4, 16, 16, frac=0.0)
    kg[np.isnan(kg)] = 1.0                            # no hole at all
    planes = [(hg, hf), (bad_g, bad_f), (rg, rf), (ng, nf), (hg2, hf2), (kg, kf)]
    gs = [_dev(g) for g, _ in planes]
    fs = [_dev(f) for _, f in planes]
    kw = dict(tstep=0.25, niter=0, scale=1, cgit=-1, cg_fp32=1, cg_rel=1e-6, cg_mg=mg)
    print(f"mixed batch mg={mg}", flush=True)
    outs, st, ps = m.solve_batch_device(gs, fs, plane_stats=True, **kw)
    torch.cuda.synchronize()
    names = [m.PLANE_STATUS_NAMES[s.status] for s in ps]
    print("statuses", names, [s.cg_iters for s in ps], flush=True)
    assert ps[0].status == m.PLANE_CONVERGED and ps[0].rel_residual <= 1e-5
    assert ps[1].status == m.PLANE_NONFINITE
    assert ps[2].status == m.PLANE_CONVERGED and ps[2].cg_iters == 0
    assert ps[3].status == m.PLANE_CONVERGED           # b = 0: x = 0 solves it
    assert ps[4].status == m.PLANE_CONVERGED and ps[4].rel_residual <= 1e-5
    assert ps[5].status == m.PLANE_CONVERGED and ps[5].holes == 0
    # the ramp is discrete-harmonic: its holes come back exactly
    assert bits_equal(outs[2].cpu().numpy(), rwant)
    assert bool((outs[3] == 0).all())
    assert bits_equal(outs[5].cpu().numpy(), kg)
    # the healthy planes keep their bits when solved alone
    for k in (0, 4):
        alone, _ = m.solve_batch_device([gs[k]], [fs[k]], **kw)
        assert _bits(alone[0], outs[k]), k
    for s, (g, _) in zip(ps, planes):
        assert s.holes == int(np.isnan(g).sum())
        assert s.levels == 1


def test_iteration_limit_status():
    """-c 3: every plane stops at its limit, flagged; a plane's bits do not
    depend on its neighbours in the batch."""
    import torch
    m = _m()
    g1, f1 = _healthy(5, 100, 120)
    g2, f2 = _healthy(6, 50, 70)
    kw = dict(tstep=0.25, niter=0, scale=1, cgit=3, cg_fp32=1, cg_rel=1e-9)
    outs, st, ps = m.solve_batch_device([_dev(g1), _dev(g2)], [_dev(f1), _dev(f2)],
                                        plane_stats=True, **kw)
    assert [s.status for s in ps] == [m.PLANE_MAXIT] * 2
    assert [s.cg_iters for s in ps] == [3, 3]
    alone, _ = m.solve_batch_device([_dev(g2)], [_dev(f2)], **kw)
    torch.cuda.synchronize()
    assert _bits(alone[0], outs[1])


def test_parity_cg_status_and_sweeps():
    """The FP64 parity CG and the sweep-only path report their outcome too."""
    m = _m()
    g, f = _healthy(7, 60, 50)
    outs, _, ps = m.solve_batch_device([_dev(g)], [_dev(f)], plane_stats=True,
                                       niter=5, cgit=400)
    assert ps[0].status == m.PLANE_CONVERGED and ps[0].sweeps == 5
    assert ps[0].levels == 7 and ps[0].cg_iters > 0
    ref, _ = oracle_solve(g, f, niter=5, cgit=400)
    assert abs(ps[0].rel_residual - m.residual_device(_dev(ref), _dev(g), _dev(f))[0]) \
        <= 1e-6
    outs, _, ps = m.solve_batch_device([_dev(g)], [_dev(f)], plane_stats=True, niter=7)
    assert ps[0].status == m.PLANE_SWEPT and ps[0].cg_iters == 0 and ps[0].cg_rr == -1
    assert bits_equal(outs[0].cpu().numpy(), oracle_solve(g, f, niter=7)[0])


def test_allocation_cap_and_cache_eviction():
    """A cap below what a plan needs fails with SP_ERR_NOMEM (after the cache
    has been emptied and the plan tried once more), leaves nothing held, and
    lifting it restores the solve; the plan cache keeps several plans and
    frees them on request."""
    import torch
    m = _m()
    m.plan_cache_clear()
    base = m.device_bytes()
    g, f = _healthy(8, 300, 280)
    dg, df = _dev(g), _dev(f)
    kw = dict(niter=0, scale=1, cgit=-1, cg_fp32=1, cg_rel=1e-6, cg_mg=1)
    old = m.set_alloc_cap(base + 4096)
    try:
        with pytest.raises(m.SimpoisError, match="status 2"):
            m.solve_device(dg, df, **kw)
        assert m.device_bytes() == base
    finally:
        m.set_alloc_cap(old)
    out, _ = m.solve_device(dg, df, **kw)
    assert m.residual_device(out, dg, df)[0] <= 1e-5
    held = m.device_bytes()
    assert held > base
    # a second call signature is cached beside the first
    out2 = torch.empty_like(dg)
    m.solve_device(dg, df, out=out2, **kw)
    assert m.device_bytes() > held
    torch.cuda.synchronize()
    assert _bits(out, out2)
    m.plan_cache_limit(1, 1 << 40)      # keeps only the most recent plan
    assert m.device_bytes() == held     # (both plans have the same size)
    m.plan_cache_limit(256, 16 << 30)
    m.plan_cache_clear()
    assert m.device_bytes() == base
