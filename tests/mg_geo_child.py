"""Child process of the multigrid tile-shape tests: solves one plane with
the multigrid-preconditioned FP32 CG under the coarse tile shapes that
SIMPOIS_HIP_MG_GEO (read once, when the library loads) selects, and saves
the solution, the relative residual and the CG iteration count.

usage: python mg_geo_child.py H W SEED OUT.npz
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def rectangle_holes(h, w, seed, frac=0.5):
    """A smooth field with rectangles of 8..512 pixels a side removed until
    `frac` of the plane is missing, and a small data term f."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    g = (100 * np.sin(x / 37.0) * np.cos(y / 23.0) + 0.01 * x).astype(np.float32)
    holes = np.zeros((h, w), bool)
    while holes.mean() < frac:
        for _ in range(32):
            rh, rw = rng.integers(8, min(512, h) + 1), rng.integers(8, min(512, w) + 1)
            y0, x0 = rng.integers(0, h - rh + 1), rng.integers(0, w - rw + 1)
            holes[y0:y0 + rh, x0:x0 + rw] = True
    g[holes] = np.nan
    f = rng.normal(0, 0.5, (h, w)).astype(np.float32)
    return g, f


def main():
    h, w, seed, out = int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]), sys.argv[4]
    import torch
    import imscript_amd as m
    g, f = rectangle_holes(h, w, seed)
    dg, df = torch.from_numpy(g).cuda(), torch.from_numpy(f).cuda()
    print(f"mg child {h}x{w} seed={seed} geo={os.environ.get('SIMPOIS_HIP_MG_GEO')}",
          flush=True)
    u, st = m.solve_device(dg, df, niter=0, scale=1, cgit=-1, cg_fp32=1,
                           cg_rel=1e-6, cg_mg=1)
    rel = m.residual_device(u, dg, df)
    np.savez(out, u=u.cpu().numpy(), rel=np.array(rel), iters=st.cg_iters_finest)


if __name__ == "__main__":
    main()
