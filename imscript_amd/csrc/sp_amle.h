// AMLE (absolutely minimizing Lipschitz extension) sweeps on the skewed
// temporal tiles of sp_gs.h; shared by the device kernels and the CPU
// emulator of the tests.
//
// amle (amle.c:67-88, 203-246) updates the holes in raster order, in place:
// the new value of (i, j) is (a*vmin + b*vmax) / (a + b) over the in-image
// neighbours among the first nn entries of the neighbour table
// (amle.c:16-29), vmin / vmax the first strict minimum / maximum in table
// order and a / b the squared lengths of the maximum's / minimum's offsets.
// The neighbourhood of the first nn entries has the dependence pattern of
// a stencil of radius R = amle_radius(nn) in the sense of sp_gs.h
// (nn <= 4: the 5-point cross, s = i + j + 2k; nn <= 8: the 3x3 box;
// nn <= 16: offsets (2,1); nn <= 32: offsets (3,2)): every offset (dx, dy)
// before the pixel in raster order has dx + R*dy < 0, every offset after it
// has dx + R*dy < R*R + 1, |dy| <= R and |dx + (R-1)*dy| <= R*(R-1) (1 for
// R = 1), so the tiles of radius R reproduce the raster order bit for bit.
#pragma once
#include "sp_gs.h"

// offsets and squared lengths in the reference's order (amle.c:16-29)
struct AmleNbr { int dx, dy, w; };
constexpr AmleNbr AMLE_NBR[32] = {
	{+1, 0, 1}, {0, +1, 1}, {-1, 0, 1}, {0, -1, 1},
	{+1, +1, 2}, {-1, -1, 2}, {-1, +1, 2}, {+1, -1, 2},
	{+2, +1, 5}, {+1, +2, 5}, {+2, -1, 5}, {+1, -2, 5},
	{-2, -1, 5}, {-1, -2, 5}, {-2, +1, 5}, {-1, +2, 5},
	{+3, +1, 10}, {+1, +3, 10}, {+3, -1, 10}, {+1, -3, 10},
	{-3, -1, 10}, {-1, -3, 10}, {-3, +1, 10}, {-1, +3, 10},
	{+3, +2, 13}, {+2, +3, 13}, {+3, -2, 13}, {+2, -3, 13},
	{-3, -2, 13}, {-2, -3, 13}, {-3, +2, 13}, {-2, +3, 13},
};
constexpr int AMLE_NN_MAX = 32;

SP_HD int amle_radius(int nn) { return nn <= 4 ? 1 : nn <= 8 ? 2 : nn <= 16 ? 3 : 4; }

// Tile shapes of the AMLE sweeps (radius, variant, BI, BJ, BT): the
// Gauss-Seidel shapes of radius 1 and smaller ones (a single plane has few
// tiles per wavefront), one shape each for the larger neighbourhoods.
#define SP_AMLE_SHAPES(X) \
	X(1, 0, 32, 64, 16) X(1, 1, 32, 128, 16) X(1, 2, 64, 64, 16) \
	X(1, 3, 16, 128, 32) X(1, 4, 32, 64, 8) X(1, 5, 64, 32, 16) \
	X(1, 6, 32, 32, 16) X(1, 7, 16, 64, 16) X(1, 8, 32, 64, 32) \
	X(1, 9, 64, 64, 8) X(1, 10, 32, 32, 8) X(1, 11, 16, 32, 8) \
	X(1, 12, 16, 32, 16) X(1, 13, 32, 16, 8) X(1, 14, 16, 16, 16) \
	X(2, 0, 32, 32, 8) X(3, 0, 32, 32, 4) X(4, 0, 32, 32, 4)

template <class F> inline bool amle_with_shape(int R, int v, F &&f)
{
#define SP_AMLE_CASE(r, var, bi, bj, bt) \
	if (R == r && v == var) { f(Shape<r, bi, bj, bt>{}); return true; }
	SP_AMLE_SHAPES(SP_AMLE_CASE)
#undef SP_AMLE_CASE
	return false;
}

inline int amle_shape_count(int R)
{
	int n = 0;
#define SP_AMLE_COUNT(r, var, bi, bj, bt) n += R == r;
	SP_AMLE_SHAPES(SP_AMLE_COUNT)
#undef SP_AMLE_COUNT
	return n;
}

// With no in-image neighbour the reference divides 0 by 0 (its zeroed
// value / weight arrays); x86 returns its default NaN, sign bit set.
constexpr uint32_t AMLE_DIV0_NAN_BITS = 0xffc00000u;

// New value of hole (i, j) of a W x H plane whose value sits at U[idx]
// (neighbour (dx, dy) at U[idx + dy*RS + dx]), first nn table entries.
SP_HD float amle_value(const float *U, int idx, int i, int j, int W, int H,
		int RS, int nn)
{
	bool any = false;
	float vmin = 0.0f, vmax = 0.0f, wmin = 0.0f, wmax = 0.0f;
#pragma unroll
	for (int p = 0; p < AMLE_NN_MAX; p++) {
		if (p >= nn) break;
		const int dx = AMLE_NBR[p].dx, dy = AMLE_NBR[p].dy;
		const int ii = i + dx, jj = j + dy;
		if (ii < 0 || jj < 0 || ii >= W || jj >= H) continue;
		const float v = U[idx + dy * RS + dx], wt = (float)AMLE_NBR[p].w;
		if (!any) {
			vmin = vmax = v;
			wmin = wmax = wt;
			any = true;
			continue;
		}
		if (v < vmin) { vmin = v; wmin = wt; }
		if (v > vmax) { vmax = v; wmax = wt; }
	}
	if (!any) return sp_float(AMLE_DIV0_NAN_BITS);
	const float p1 = wmax * vmin, p2 = wmin * vmax;
	return (p1 + p2) / (wmax + wmin);
}

// The same for nn = 4 inside an R = 1 tile, where every weight is 1:
// (vmin + vmax) / 2.  Cells outside the image hold NaN (tile_load), which a
// strict comparison never selects, so they drop out of both scans except as
// the scans' first value: that is the east neighbour, outside the image in
// the last column only, where the first in-image neighbour replaces it.  A
// 1 x 1 plane (no neighbour at all) is the caller's case.
SP_HD float amle_value4(const float *U, int idx, int i, int j, int W, int H, int RS)
{
	const float e = U[idx + 1], so = U[idx + RS], w = U[idx - 1], n = U[idx - RS];
	float f = e;
	if (i + 1 >= W) f = j + 1 < H ? so : i > 0 ? w : n;
	float mn = f, mx = f;
	mn = so < mn ? so : mn;
	mx = so > mx ? so : mx;
	mn = w < mn ? w : mn;
	mx = w > mx ? w : mx;
	mn = n < mn ? n : mn;
	mx = n > mx ? n : mx;
	return (mn + mx) * 0.5f;
}

// phase 2 of an AMLE tile (thread tid, in-tile hyperplane sg): as
// tile_step of sp_gs.h with the AMLE update.  The tile was loaded by
// tile_load (for R = 1 cells outside the image hold NaN and the known
// marker; for R > 1 the update reads only in-image neighbours).
template <class S>
SP_HD void amle_tile_step(const TileCtx<S> &c, int tid, int sg, int nn)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	const int a = tid % G::BI, tau = tid / G::BI;
	if (tau >= c.bte) return;
	const int b = R == 1 ? sg - a : sg - a - tau;
	if (b < 0 || b >= G::BJ) return;
	if (R == 1) {
		const int i = c.x0 + a - tau, j = c.y0 + b - tau;
		const int idx = (b - tau + G::BT) * G::LW + (a - tau + G::BT);
		const float cv = c.C[idx], x = c.U[idx];
		const float y = nn != 4 ? amle_value(c.U, idx, i, j, c.W, c.H, G::LW, nn)
			: c.W == 1 && c.H == 1 ? sp_float(AMLE_DIV0_NAN_BITS)
			: amle_value4(c.U, idx, i, j, c.W, c.H, G::LW);
		c.U[idx] = sp_known(cv) ? x : y;
		return;
	}
	const int i = c.x0 + a - G::SH * b, j = c.y0 + b - R * tau;
	if (i < 0 || i >= c.W || j < 0 || j >= c.H) return;
	const int idx = (j - c.jlo) * G::LW + (i + G::SH * j - c.K);
	if (!sp_known(c.C[idx]))
		c.U[idx] = amle_value(c.U, idx, i, j, c.W, c.H, G::LW + G::SH, nn);
}

// The 4-neighbour sweep of a radius-1 tile with its per-thread state kept
// across steps: thread (a, tau) updates the cell at in-tile row b = sg - a,
// whose LDS index and image row advance by one row per step, so a step is
// the update alone.  Same cells, order and results as amle_tile_step with
// nn = 4 (planes larger than 1 x 1).
struct Amle4Thread { int idx, j, i, a; bool on, eb; };

template <class S>
SP_HD Amle4Thread amle4_thread(const TileCtx<S> &c, int tid)
{
	using G = Geo<S>;
	Amle4Thread t;
	const int a = tid % G::BI, tau = tid / G::BI;
	t.a = a;
	t.on = tau < c.bte;
	t.i = c.x0 + a - tau;
	t.j = c.y0 - a - tau;                              // at sg = 0, b = -a
	t.idx = (G::BT - a - tau) * G::LW + (a - tau + G::BT);
	t.eb = t.i + 1 >= c.W;                             // east neighbour outside
	return t;
}

template <class S>
SP_HD void amle4_step(const TileCtx<S> &c, Amle4Thread &t, int sg)
{
	using G = Geo<S>;
	const int b = sg - t.a;
	if (t.on && b >= 0 && b < G::BJ) {
		const float *U = c.U;
		const int q = t.idx;
		const float cv = c.C[q], x = U[q];
		const float e = U[q + 1], so = U[q + G::LW], w = U[q - 1], n = U[q - G::LW];
		float f = e;
		if (t.eb) f = t.j + 1 < c.H ? so : t.i > 0 ? w : n;
		float mn = f, mx = f;
		mn = so < mn ? so : mn;
		mx = so > mx ? so : mx;
		mn = w < mn ? w : mn;
		mx = w > mx ? w : mx;
		mn = n < mn ? n : mn;
		mx = n > mx ? n : mx;
		c.U[q] = sp_known(cv) ? x : (mn + mx) * 0.5f;
	}
	t.idx += G::LW;
	t.j += 1;
}
