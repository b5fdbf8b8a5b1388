"""FP64 CPU reference of the converged masked Poisson and biharmonic fills
and their gradients.

TEST INFRASTRUCTURE: only tests/ and the benchmark's checks use it.

The system is the reference's CG system (cleant_cgpois.c:197-264 with the
operator of :67-89): unknowns are the NaN pixels of g,
    A u_h = b,  b = -f_h - A_hk g_k,  A = -Laplacian with clamped reads,
assembled here in FP64 from scratch and solved with a sparse direct solver.
The gradient of a loss L(u) is computed with the general adjoint formulas
(no use of symmetry):  lambda = A_hh^-T v_h,  dL/df_h = -lambda,
dL/dg_k = v_k - (A_hk^T lambda)_k.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_ST5 = [(0, 0, -4.0), (1, 0, 1.0), (0, 1, 1.0), (-1, 0, 1.0), (0, -1, 1.0)]
# the reference's 13-point bilaplacian (cleant_cgpois.c:91-121): 20 c - 8 (4
# neighbours) + 2 (diagonals) + (distance-2 points), every read clamped
_ST13 = [(0, 0, 20.0), (1, 0, -8.0), (0, 1, -8.0), (-1, 0, -8.0), (0, -1, -8.0),
         (1, 1, 2.0), (1, -1, 2.0), (-1, 1, 2.0), (-1, -1, 2.0),
         (2, 0, 1.0), (0, 2, 1.0), (-2, 0, 1.0), (0, -2, 1.0)]


def masked_operator(g):
    """(A_hh, A_hk, holes, known) for the -Laplacian with clamped reads."""
    h, w = g.shape
    hole = np.isnan(g).ravel()
    idx_h = np.flatnonzero(hole)
    idx_k = np.flatnonzero(~hole)
    pos_h = -np.ones(h * w, int)
    pos_h[idx_h] = np.arange(len(idx_h))
    pos_k = -np.ones(h * w, int)
    pos_k[idx_k] = np.arange(len(idx_k))
    rh, ch, vh, rk, ck, vk = [], [], [], [], [], []
    for p, q in enumerate(idx_h):
        i, j = q % w, q // w
        for di, dj, c in _ST5:
            a, b = min(max(i + di, 0), w - 1), min(max(j + dj, 0), h - 1)
            t = b * w + a
            if hole[t]:
                rh.append(p); ch.append(pos_h[t]); vh.append(-c)
            else:
                rk.append(p); ck.append(pos_k[t]); vk.append(-c)
    Ahh = sp.csr_matrix((vh, (rh, ch)), shape=(len(idx_h), len(idx_h)))
    Ahk = sp.csr_matrix((vk, (rk, ck)), shape=(len(idx_h), len(idx_k)))
    return Ahh, Ahk, idx_h, idx_k


def solve(g, f=None):
    """Converged fill in FP64 (float64 array, known pixels copied)."""
    g = np.asarray(g, np.float64)
    Ahh, Ahk, ih, ik = masked_operator(g)
    gf = g.ravel()
    b = -Ahk @ gf[ik]
    if f is not None:
        b -= np.asarray(f, np.float64).ravel()[ih]
    u = gf.copy()
    if len(ih):
        u[ih] = spla.spsolve(Ahh.tocsc(), b)
    return u.reshape(g.shape)


def masked_bilaplacian(g):
    """(A_hh, A_hk, holes, known) of the 13-point clamped bilaplacian, the
    operator of the reference's biharmonic CG (sign as there: +20 centre)."""
    h, w = g.shape
    hole = np.isnan(g).ravel()
    idx_h = np.flatnonzero(hole)
    idx_k = np.flatnonzero(~hole)
    pos = -np.ones(h * w, int)
    pos[idx_h] = np.arange(len(idx_h))
    posk = -np.ones(h * w, int)
    posk[idx_k] = np.arange(len(idx_k))
    i, j = idx_h % w, idx_h // w
    rh, ch, vh, rk, ck, vk = [], [], [], [], [], []
    rows = np.arange(len(idx_h))
    for di, dj, c in _ST13:
        t = np.clip(j + dj, 0, h - 1) * w + np.clip(i + di, 0, w - 1)
        th = hole[t]
        rh.append(rows[th]); ch.append(pos[t[th]]); vh.append(np.full(th.sum(), c))
        rk.append(rows[~th]); ck.append(posk[t[~th]]); vk.append(np.full((~th).sum(), c))
    cat = np.concatenate
    Ahh = sp.csr_matrix((cat(vh), (cat(rh), cat(ch))), shape=(len(idx_h), len(idx_h)))
    Ahk = sp.csr_matrix((cat(vk), (cat(rk), cat(ck))), shape=(len(idx_h), len(idx_k)))
    return Ahh, Ahk, idx_h, idx_k


def solve_bilap(g, f=None):
    """Converged biharmonic fill in FP64: the system of the reference's CG,
    A_hh u = -f - A_hk g_k with the 13-point operator (Delta^2 u = -f;
    cleant_cgpois.c:225-236, 273-278); a direct solve, since A is not
    symmetric within two pixels of the border."""
    g = np.asarray(g, np.float64)
    Ahh, Ahk, ih, ik = masked_bilaplacian(g)
    gf = g.ravel()
    b = -Ahk @ gf[ik]
    if f is not None:
        b -= np.asarray(f, np.float64).ravel()[ih]
    u = gf.copy()
    if len(ih):
        u[ih] = spla.spsolve(Ahh.tocsc(), b)
    return u.reshape(g.shape)


def gradient(g, f, v, bilap=False):
    """(dL/dg, dL/df) for dL/du = v, FP64; bilap: the biharmonic fill of
    solve_bilap (its A_hh is not symmetric near the border: the transpose
    below matters there)."""
    g = np.asarray(g, np.float64)
    v = np.asarray(v, np.float64).ravel()
    Ahh, Ahk, ih, ik = (masked_bilaplacian if bilap else masked_operator)(g)
    lam = spla.spsolve(Ahh.T.tocsc(), v[ih]) if len(ih) else np.zeros(0)
    gg = np.zeros(g.size)
    gf = np.zeros(g.size)
    gg[ik] = v[ik] - Ahk.T @ lam
    gf[ih] = -lam
    return gg.reshape(g.shape), gf.reshape(g.shape)


# ------------------------------------------------- simpois's finite sweeps
#
# simpois with -c 0 (simpois.c:247-285): niter raster-order Gauss-Seidel
# sweeps at every level of the 2x pyramid, restated in FP64 for the finite
# differences that check the gradient of the sweeps' own output.  Data are
# finite or NaN (holes).

def _zoom_out(a):
    """NaN-aware 2x2 mean with clamped reads (simpois.c:172-196)."""
    h, w = a.shape
    hs, ws = (h + 1) // 2, (w + 1) // 2
    out = np.empty((hs, ws))
    for J in range(hs):
        for I in range(ws):
            s, c = 0.0, 0
            for j, i in ((2 * J, 2 * I), (2 * J, 2 * I + 1), (2 * J + 1, 2 * I),
                         (2 * J + 1, 2 * I + 1)):
                v = a[min(j, h - 1), min(i, w - 1)]
                if np.isfinite(v):
                    s += v
                    c += 1
            out[J, I] = s / c if c else np.nan
    return out


def _zoom_in(a, w, h):
    """Cell-centred bilinear x2 with clamped reads (simpois.c:199-239)."""
    hs, ws = a.shape
    out = np.empty((h, w))
    for j in range(h):
        y = (j - 0.5) / 2
        iq = int(np.floor(y))
        qy = y - iq
        for i in range(w):
            x = (i - 0.5) / 2
            ip = int(np.floor(x))
            px = x - ip
            at = lambda ii, jj: a[min(max(jj, 0), hs - 1), min(max(ii, 0), ws - 1)]
            out[j, i] = (at(ip, iq) * (1 - px) * (1 - qy) + at(ip + 1, iq) * px * (1 - qy)
                         + at(ip, iq + 1) * (1 - px) * qy + at(ip + 1, iq + 1) * px * qy)
    return out


def _sweep(u, f, hole, t):
    """One raster-order Gauss-Seidel sweep over the holes, in place
    (simpois.c:72-119: the Laplacian skips non-finite and clamped
    neighbours)."""
    h, w = u.shape
    for j, i in zip(*np.nonzero(hole)):
        x = u[j, i]
        r = 0.0
        for a, b in ((i + 1, j), (i, j + 1), (i - 1, j), (i, j - 1)):
            if 0 <= a < w and 0 <= b < h and np.isfinite(u[b, a]):
                r += u[b, a] - x
        u[j, i] = x + t * (r - (f[j, i] if f is not None else 0.0))


def simpois64(g, f=None, tstep=0.25, niter=10, scale=99, pmsfac=3.0, ponlit=0):
    """poisson_rec (simpois.c:247-285) with cgit = 0, in FP64."""
    g = np.asarray(g, np.float64)
    f = None if f is None else np.asarray(f, np.float64)
    h, w = g.shape
    if scale > 1 and (w > 1 or h > 1):
        gs = _zoom_out(g)
        fs = None if f is None else _zoom_out(f) * pmsfac
        us = simpois64(gs, fs, tstep, niter, scale - 1, pmsfac, ponlit)
        init = _zoom_in(us, w, h)
    else:
        init = np.zeros((h, w))
    n = 0 if (ponlit and ponlit != w) else niter
    u = np.where(np.isfinite(g), g, init)
    hole = np.isnan(g)
    for _ in range(n):
        _sweep(u, f, hole, tstep)
    # the level's output (cleant_cgpois.c:203-207, 251-256)
    if hole.any():
        return np.where(hole, u, g)
    return np.where(np.isfinite(g), g, u)


def sweeps_gradient_fd(g, f, v, eps=1e-3, **kw):
    """(dL/dg at the known pixels, dL/df) of L = sum(v * simpois64(g, f)) by
    FP64 central differences; the map is affine in (g, f), so they are exact
    up to rounding."""
    g = np.asarray(g, np.float64)
    f = np.zeros_like(g) if f is None else np.asarray(f, np.float64)
    v = np.asarray(v, np.float64)
    loss = lambda gg, ff: float(np.sum(v * simpois64(gg, ff, **kw)))
    gg = np.zeros_like(g)
    gf = np.zeros_like(g)
    for j, i in zip(*np.nonzero(~np.isnan(g))):
        e = np.zeros_like(g)
        e[j, i] = eps
        gg[j, i] = (loss(g + e, f) - loss(g - e, f)) / (2 * eps)
    for j, i in np.ndindex(*g.shape):
        e = np.zeros_like(g)
        e[j, i] = eps
        gf[j, i] = (loss(g, f + e) - loss(g, f - e)) / (2 * eps)
    return gg, gf
