// Gauss-Seidel tile geometry and per-thread phases, shared by the device
// kernels and by the CPU emulator used in the tests (the same code, so the
// CPU run checks the device kernels' indexing and schedule).
//
// Gauss-Seidel (simpois.c:93-119) updates the NaN pixels in raster order,
// in place.  The update of pixel (i,j) at sweep k only depends on updates
// with a smaller value of s = i + R*j + (R*R+1)*k (R = stencil radius), so
// all updates on one hyperplane s are independent and hyperplanes executed
// in increasing s reproduce the raster order bit for bit.
//
// Skewed coordinates (i', j') and the tile box [I'BI, +BI) x [J'BJ, +BJ) x
// [T'BT, +BT) of (i', j', k):
//   R = 1: i' = i + k,                 j' = j + k      in-tile step a + b
//   R > 1: i' = i + (R-1) j + R(R-1) k, j' = j + R k   in-tile step a + b + tau
// Every dependence between updates (true, and the anti-dependence of the
// in-place update) has a componentwise non-negative skewed distance, so a
// tile only depends on componentwise-smaller tiles and tiles with equal
// I'+J'+T' are independent.  A tile loads its whole footprint from the
// in-place image, runs its hyperplanes in LDS and writes back the pixels
// it updated; the image then always holds the version each tile needs.
#pragma once
#include <stdint.h>
#include <string.h>
#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SP_HD __host__ __device__ __forceinline__
#else
#define SP_HD inline
#endif

#define SP_KNOWN_BITS 0x7fbadbadu
#define SP_QNAN_BITS 0x7fc00000u

// Gauss-Seidel work item: one plane with its sweep count and effective step.
struct GsInst {
	float *u;
	const float *c;
	int w, h, nsweep;
	float t;
	const int *active;   // optional per-instance flag (nullptr = always on)
};

// Skewed temporal tile: (instance, I', J', T').
struct GsTile { int inst, ti, tj, tt; };

// a member's tile grid for the dependence-ordered launch: flag index of
// tile (I', J', T') = base + (T' nJ + J') nI + I'
struct GsFlowInst { int base, nI, nJ, nT; };

// Tile shape: skewed extents BI x BJ and BT sweeps per tile (the sweeps of
// one pass over the image) for stencil radius R.
template <int R_, int BI_, int BJ_, int BT_> struct Shape {
	static constexpr int R = R_, BI = BI_, BJ = BJ_, BT = BT_;
};

// The compiled shapes (radius, variant, BI, BJ, BT); variant 0 of each
// radius is the default, the others are candidates for the autotuner.
#define SP_GS_SHAPES(X) \
	X(1, 0, 32, 64, 16) X(1, 1, 32, 128, 16) X(1, 2, 64, 64, 16) \
	X(1, 3, 16, 128, 32) X(1, 4, 32, 64, 8) X(1, 5, 64, 32, 16) \
	X(1, 6, 32, 32, 16) X(1, 7, 16, 64, 16) X(1, 8, 32, 64, 32) \
	X(1, 9, 64, 64, 8) X(1, 10, 16, 16, 32) X(1, 11, 16, 32, 32) \
	X(2, 0, 32, 32, 8) X(2, 1, 32, 64, 8) X(2, 2, 32, 32, 16) \
	X(2, 3, 64, 32, 8) \
	X(3, 0, 32, 32, 4) X(3, 1, 32, 64, 4) X(3, 2, 32, 32, 8)

// calls f(Shape<...>{}) for variant v of radius R; false if there is none
template <class F> inline bool sp_with_shape(int R, int v, F &&f)
{
#define SP_SHAPE_CASE(r, var, bi, bj, bt) \
	if (R == r && v == var) { f(Shape<r, bi, bj, bt>{}); return true; }
	SP_GS_SHAPES(SP_SHAPE_CASE)
#undef SP_SHAPE_CASE
	return false;
}

inline int sp_shape_count(int R)
{
	int n = 0;
#define SP_SHAPE_COUNT(r, var, bi, bj, bt) n += R == r;
	SP_GS_SHAPES(SP_SHAPE_COUNT)
#undef SP_SHAPE_COUNT
	return n;
}

SP_HD uint32_t sp_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
SP_HD float sp_float(uint32_t b) { return __builtin_bit_cast(float, b); }
SP_HD bool sp_known(float c) { return sp_bits(c) == SP_KNOWN_BITS; }
SP_HD int sp_clamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// coefficient of a pixel: the data term at a hole (NaN made canonical so it
// can never equal the known marker), the known marker elsewhere
SP_HD float sp_coef(float g, float d)
{
	if (!isnan(g)) return sp_float(SP_KNOWN_BITS);
	return isnan(d) ? sp_float(SP_QNAN_BITS) : d;
}

// Nested clamped Laplacian of depth D centred at offset (CX, CY) inside a
// register window v of side N = 2R+1: -4*c + E + S + W + N, left to right
// (simpois.c:35-68).  Depth 0 is the pixel value itself.
template <int D, int CX, int CY, int N> struct NestedLap {
	static SP_HD float eval(const float (&v)[N][N])
	{
		float r = -4.0f * NestedLap<D - 1, CX, CY, N>::eval(v);
		r = r + NestedLap<D - 1, CX + 1, CY, N>::eval(v);
		r = r + NestedLap<D - 1, CX, CY + 1, N>::eval(v);
		r = r + NestedLap<D - 1, CX - 1, CY, N>::eval(v);
		r = r + NestedLap<D - 1, CX, CY - 1, N>::eval(v);
		return r;
	}
};
template <int CX, int CY, int N> struct NestedLap<0, CX, CY, N> {
	static SP_HD float eval(const float (&v)[N][N]) { return v[CY + N / 2][CX + N / 2]; }
};

// One in-place update of hole pixel (i,j) whose value sits at U[idx];
// neighbour (di,dj) sits at U[idx + dj*RS + di].
//  R = 1: Laplacian that skips non-finite neighbours (simpois.c:72-90).
//         A clamped neighbour is the pixel itself and adds (x - x) = +0 or
//         is skipped, which leaves the sum unchanged, so it is left out.
//  R = 2, 3: nested clamped Laplacians (simpois.c:48-68); every read is of
//         the clamped coordinate, so the window is read with clamping.
template <int R>
SP_HD void gs_update(float *U, float cv, int idx, int i, int j, int W, int H,
		int RS, float t)
{
	if constexpr (R == 1) {
		const float x = U[idx];
		float r = 0.0f;
		if (i + 1 < W) { float v = U[idx + 1]; if (isfinite(v)) r = r + (v - x); }
		if (j + 1 < H) { float v = U[idx + RS]; if (isfinite(v)) r = r + (v - x); }
		if (i > 0) { float v = U[idx - 1]; if (isfinite(v)) r = r + (v - x); }
		if (j > 0) { float v = U[idx - RS]; if (isfinite(v)) r = r + (v - x); }
		U[idx] = x + t * (r - cv);
	} else {
		constexpr int N = 2 * R + 1;
		float v[N][N];
#pragma unroll
		for (int dy = -R; dy <= R; dy++) {
			const int cj = sp_clamp(j + dy, 0, H - 1) - j;
#pragma unroll
			for (int dx = -R; dx <= R; dx++) {
				if ((dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) > R) {
					v[dy + R][dx + R] = 0.0f;
					continue;
				}
				const int ci = sp_clamp(i + dx, 0, W - 1) - i;
				v[dy + R][dx + R] = U[idx + cj * RS + ci];
			}
		}
		const float l = NestedLap<R, 0, 0, N>::eval(v);
		U[idx] = v[R][R] + t * (l - cv);
	}
}

// ------------------------------------------------------------- tiles

template <class S> struct Geo {
	static constexpr int R = S::R, BI = S::BI, BJ = S::BJ, BT = S::BT;
	static constexpr int SH = R - 1;                       // col shear per row
	static constexpr int HC = R == 1 ? 1 : R * (R - 1);    // column halo
	static constexpr int LW0 = R == 1 ? BI + BT + 1 : BI + R * (R - 1) * (BT - 1) + 2 * HC;
	static constexpr int LW = LW0 + (LW0 & 1);             // even: odd diagonal stride
	static constexpr int LH = R == 1 ? BJ + BT + 1 : BJ + R * (BT - 1) + 2 * R;
	static constexpr int NT = BI * BT;                      // threads per tile
	static constexpr int NSTEP = R == 1 ? BI + BJ - 1 : BI + BJ + BT - 2;
	static constexpr int LDS_FLOATS = 2 * LW * LH;          // u and coefficient
	static_assert(NT <= 1024 && NT % 64 == 0, "tile threads");
	static_assert(LDS_FLOATS * 4 <= 65536, "tile footprint exceeds 64 KiB");
};

// placement of one tile: LDS cell (row, col) holds pixel
// j = row + jlo, i = col + K - SH*j
template <class S> struct TileCtx {
	int x0, y0, jlo, K, bte, W, H;
	float t;
	float *U, *C;
};

template <class S>
SP_HD TileCtx<S> tile_ctx(const GsTile &tl, const GsInst &in, float *lds)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	TileCtx<S> c;
	const int k0 = tl.tt * G::BT;
	c.bte = in.nsweep - k0 < G::BT ? in.nsweep - k0 : G::BT;
	if (R == 1) {
		c.x0 = tl.ti * G::BI - k0;
		c.y0 = tl.tj * G::BJ - k0;
		c.jlo = c.y0 - G::BT;
		c.K = c.x0 - G::BT;
	} else {
		c.x0 = tl.ti * G::BI - G::SH * tl.tj * G::BJ;
		c.y0 = tl.tj * G::BJ - R * k0;
		c.jlo = c.y0 - R * (G::BT - 1) - R;
		c.K = c.x0 + G::SH * c.y0 - R * G::SH * (G::BT - 1) - G::HC;
	}
	c.W = in.w;
	c.H = in.h;
	c.t = in.t;
	c.U = lds;
	c.C = lds + G::LW * G::LH;
	return c;
}

// phase 1 (thread tid): copy the tile's footprint into LDS.  For R = 1 a
// cell outside the image holds a NaN value and the known marker: the
// Laplacian skips non-finite neighbours, and leaving out a clamped
// neighbour (the pixel itself, adding x - x = +0) does not change the sum,
// so the update needs no image-bounds tests.
template <class S>
SP_HD void tile_load(const TileCtx<S> &c, const GsInst &in, int tid)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	for (int e = tid; e < G::LW * G::LH; e += G::NT) {
		const int row = e / G::LW, col = e - row * G::LW;
		const int j = row + c.jlo, i = col + c.K - G::SH * j;
		if (i >= 0 && i < c.W && j >= 0 && j < c.H) {
			const size_t q = (size_t)j * c.W + i;
			c.U[e] = in.u[q];
			c.C[e] = in.c[q];
		} else if (R == 1) {
			c.U[e] = sp_float(SP_QNAN_BITS);
			c.C[e] = sp_float(SP_KNOWN_BITS);
		}
	}
}

// R = 1 update of the cell U[idx] inside a loaded tile (every neighbour is
// a footprint cell; outside the image it is NaN), same terms and order as
// gs_update<1>.  A known cell stores its own value back: no other thread
// writes it on this hyperplane, so the store changes nothing, and the reads
// need not wait for the known test.
SP_HD void gs_update_tile1(float *U, const float *C, int idx, int RS, float t)
{
	const float cv = C[idx], x = U[idx];
	const float e = U[idx + 1], so = U[idx + RS], w = U[idx - 1], n = U[idx - RS];
	float r = 0.0f;
	if (isfinite(e)) r = r + (e - x);
	if (isfinite(so)) r = r + (so - x);
	if (isfinite(w)) r = r + (w - x);
	if (isfinite(n)) r = r + (n - x);
	const float y = x + t * (r - cv);
	U[idx] = sp_known(cv) ? x : y;
}

// phase 2 (thread tid, step sg): the update of thread (a, tau) on the
// in-tile hyperplane sg; steps run in increasing sg with a barrier between
template <class S>
SP_HD void tile_step(const TileCtx<S> &c, int tid, int sg)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	const int a = tid % G::BI, tau = tid / G::BI;
	if (tau >= c.bte) return;
	const int b = R == 1 ? sg - a : sg - a - tau;
	if (b < 0 || b >= G::BJ) return;
	if (R == 1) {
		// (i, j) = (x0 + a - tau, y0 + b - tau) sits at row j - jlo =
		// b - tau + BT, column i - K = a - tau + BT; cells outside the
		// image carry the known marker
		const int idx = (b - tau + G::BT) * G::LW + (a - tau + G::BT);
		gs_update_tile1(c.U, c.C, idx, G::LW, c.t);
		return;
	}
	int i, j;
	if (R == 1) { i = c.x0 + a - tau; j = c.y0 + b - tau; }
	else { i = c.x0 + a - G::SH * b; j = c.y0 + b - R * tau; }
	if (i < 0 || i >= c.W || j < 0 || j >= c.H) return;
	const int idx = (j - c.jlo) * G::LW + (i + G::SH * j - c.K);
	const float cv = c.C[idx];
	if (!sp_known(cv))
		gs_update<R>(c.U, cv, idx, i, j, c.W, c.H, G::LW + G::SH, c.t);
}

// The R = 1 sweep with its per-thread state kept across steps: thread
// (a, tau) updates the cell at in-tile row b = sg - a, whose LDS index
// advances by one row per step, so a step is the update alone.  Same
// cells, order and results as tile_step for R = 1.
struct Gs1Thread { int idx, a; bool on; };

template <class S>
SP_HD Gs1Thread gs1_thread(const TileCtx<S> &c, int tid)
{
	using G = Geo<S>;
	Gs1Thread t;
	const int a = tid % G::BI, tau = tid / G::BI;
	t.a = a;
	t.on = tau < c.bte;
	t.idx = (G::BT - a - tau) * G::LW + (a - tau + G::BT);   // b = -a at sg = 0
	return t;
}

template <class S>
SP_HD void gs1_step(const TileCtx<S> &c, Gs1Thread &t, int sg)
{
	using G = Geo<S>;
	const int b = sg - t.a;
	if (t.on && b >= 0 && b < G::BJ) gs_update_tile1(c.U, c.C, t.idx, G::LW, c.t);
	t.idx += G::LW;
}

// does this tile update pixel (i, j) at some sweep?
template <class S>
SP_HD bool tile_owns(const TileCtx<S> &c, int i, int j)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	if (R == 1) {
		// sweeps tt with 0 <= i - x0 + tt < BI and 0 <= j - y0 + tt < BJ
		const int di = i - c.x0, dj = j - c.y0;
		const int lo = di < dj ? -di : -dj;
		int hi = G::BI - di < G::BJ - dj ? G::BI - di : G::BJ - dj;
		if (hi > c.bte) hi = c.bte;
		return (lo > 0 ? lo : 0) < hi;
	}
	for (int tt = 0; tt < c.bte; tt++) {
		int aa, bb;
		if (R == 1) { aa = i - c.x0 + tt; bb = j - c.y0 + tt; }
		else { bb = j - c.y0 + R * tt; aa = i - c.x0 + G::SH * bb; }
		if (aa >= 0 && aa < G::BI && bb >= 0 && bb < G::BJ) return true;
	}
	return false;
}

// phase 3 (thread tid): write back the hole pixels this tile updated
template <class S>
SP_HD void tile_store(const TileCtx<S> &c, const GsInst &in, int tid)
{
	using G = Geo<S>;
	for (int e = tid; e < G::LW * G::LH; e += G::NT) {
		const int row = e / G::LW, col = e - row * G::LW;
		const int j = row + c.jlo, i = col + c.K - G::SH * j;
		if (i < 0 || i >= c.W || j < 0 || j >= c.H || sp_known(c.C[e])) continue;
		if (tile_owns<S>(c, i, j)) in.u[(size_t)j * c.W + i] = c.U[e];
	}
}

// ------------------------------------------ whole level in one work-group

// For levels that fit in LDS: one block runs every sweep of one plane,
// hyperplane by hyperplane.  Step s, thread slot e: (row j, k-offset qk);
// a thread's slots are fixed for the whole run.  For R = 1 the level sits
// in LDS inside a one-cell frame of NaN / known cells (same argument as the
// tiles), so the update needs no bounds tests.
constexpr int SMALL_MAXSLOT = 8;
struct SmallSlots { int n; int j[SMALL_MAXSLOT], qk[SMALL_MAXSLOT]; };

SP_HD int small_slot_count(int R, int W, int H)
{
	return H * ((W - 1) / (R * R + 1) + 1);
}

// LDS cells of one plane (u and the coefficient take this many floats each)
SP_HD int small_cells(int R, int W, int H)
{
	return R == 1 ? (W + 2) * (H + 2) : W * H;
}

template <int R>
SP_HD void small_slots(int W, int H, int tid, int nthreads, SmallSlots &o)
{
	const int slots = small_slot_count(R, W, H);
	o.n = tid < slots ? (slots - tid + nthreads - 1) / nthreads : 0;
	if (o.n > SMALL_MAXSLOT) o.n = SMALL_MAXSLOT;
#pragma unroll
	for (int m = 0; m < SMALL_MAXSLOT; m++) {
		const int e = tid + m * nthreads;
		o.j[m] = e % H;
		o.qk[m] = e / H;
	}
}

template <int R>
SP_HD void small_load(float *U, float *C, const GsInst &in, int tid, int nthreads)
{
	const int W = in.w, H = in.h;
	if (R == 1) {
		const int P = W + 2, n = P * (H + 2);
		for (int e = tid; e < n; e += nthreads) {
			const int r = e / P, col = e - r * P, i = col - 1, j = r - 1;
			if (i >= 0 && i < W && j >= 0 && j < H) {
				U[e] = in.u[j * W + i];
				C[e] = in.c[j * W + i];
			} else {
				U[e] = sp_float(SP_QNAN_BITS);
				C[e] = sp_float(SP_KNOWN_BITS);
			}
		}
	} else {
		for (int e = tid; e < W * H; e += nthreads) {
			U[e] = in.u[e];
			C[e] = in.c[e];
		}
	}
}

template <int R>
SP_HD void small_store(const float *U, const float *C, const GsInst &in, int tid,
		int nthreads)
{
	const int W = in.w, H = in.h;
	for (int q = tid; q < W * H; q += nthreads) {
		const int e = R == 1 ? (q / W + 1) * (W + 2) + q % W + 1 : q;
		if (!sp_known(C[e])) in.u[q] = U[e];
	}
}

template <int R>
SP_HD void small_step(float *U, const float *C, int W, int H, int T, float t,
		int s, const SmallSlots &sl)
{
	constexpr int CT = R * R + 1;
#pragma unroll
	for (int m = 0; m < SMALL_MAXSLOT; m++) {
		if (m >= sl.n) break;
		const int j = sl.j[m];
		const int base = s - R * j;   // = i + CT*k
		if (base < 0) continue;
		int k = base / CT;
		if (k > T - 1) k = T - 1;
		k -= sl.qk[m];
		if (k < 0) continue;
		const int i = base - CT * k;
		if (i >= W) continue;
		if (R == 1) {
			gs_update_tile1(U, C, (j + 1) * (W + 2) + i + 1, W + 2, t);
		} else {
			const int idx = j * W + i;
			const float cv = C[idx];
			if (!sp_known(cv)) gs_update<R>(U, cv, idx, i, j, W, H, W, t);
		}
	}
}

SP_HD long small_nsteps(int R, int W, int H, int T)
{
	return (long)(W - 1) + (long)R * (H - 1) + (long)(R * R + 1) * (T - 1) + 1;
}

// ------------------------------- whole R = 1 level by its unknown pixels
//
// The hyperplane schedule of k_gs_small with the work dealt by pixel: the
// updates on hyperplane s = i + j + 2k all have i + j = s (mod 2), so a
// level's unknown pixels are listed once per parity (raster order) and
// dealt round-robin to the threads, and on step s a thread updates its
// pixels of parity s & 1 whose sweep k = (s - i - j) / 2 lies in [0, T).
// Known pixels take no thread.  A thread keeps its pixels' values and
// coefficients in registers (no other thread writes them) and reads the
// four neighbours from the level's values in LDS, inside a frame of NaN
// (skipped as non-finite, as in the tiles).  Same updates, same order of
// terms, same results as the raster sweeps.
//
// A sweep that leaves every unknown's bits unchanged is a fixed point: each
// later sweep reads the same values and repeats it, so the level's values
// are final from there on and the remaining steps are skipped (gsc_final).
// Coarse pyramid levels reach one long before their last sweep (a level
// without unknowns at its first).

constexpr int GSC_MAXCELLS = 16384;   // pixels of a level: <= 8192 per parity
constexpr int GSC_MAXNT = 1024;

// LDS bytes: the framed values, the list of unknowns, 16 scan words and
// the two change marks
SP_HD long gsc_lds_bytes(int W, int H)
{
	return 4L * ((long)(W + 2) * (H + 2) + (long)W * H + 18);
}

// threads for a level: one per pixel of a parity (64..1024)
SP_HD int gsc_threads(int W, int H)
{
	const int half = (W * H + 1) / 2, th = (half + 63) / 64 * 64;
	return th < 64 ? 64 : (th > GSC_MAXNT ? GSC_MAXNT : th);
}

// pixels per thread and parity the level can need with nt threads
SP_HD int gsc_slots(int W, int H, int nt)
{
	return ((W * H + 1) / 2 + nt - 1) / nt;
}

template <int MAXS> struct GscThread {
	int n[2];                 // pixels of each parity
	// framed LDS index q = (j + 1) (W + 2) + i + 1 | (i + j) << 16 (both
	// below 2^16 for GSC_MAXCELLS pixels)
	unsigned qij[2][MAXS];
	float x[2][MAXS], c[2][MAXS];
};

// phase 1 (thread tid of nt): the framed values into LDS
SP_HD void gsc_load(float *U, const GsInst &in, int tid, int nt)
{
	const int W = in.w, H = in.h, P = W + 2, n = P * (H + 2);
	for (int e = tid; e < n; e += nt) {
		const int r = e / P, col = e - r * P, i = col - 1, j = r - 1;
		U[e] = i >= 0 && i < W && j >= 0 && j < H ? in.u[j * W + i]
			: sp_float(SP_QNAN_BITS);
	}
}

// the raster chunk [q0, q1) of pixels a thread lists
SP_HD void gsc_chunk(int W, int H, int tid, int nt, int &q0, int &q1)
{
	const int n = W * H, ch = (n + nt - 1) / nt;
	q0 = tid * ch < n ? tid * ch : n;
	q1 = q0 + ch < n ? q0 + ch : n;
}

// phase 2: the chunk's unknowns by parity, packed as even | odd << 16
SP_HD int gsc_count(const GsInst &in, int tid, int nt)
{
	int q0, q1;
	gsc_chunk(in.w, in.h, tid, nt, q0, q1);
	int j = q0 / in.w, i = q0 - j * in.w, ce = 0, co = 0;
	for (int q = q0; q < q1; q++) {
		if (!sp_known(in.c[q])) {
			if ((i + j) & 1) co++;
			else ce++;
		}
		if (++i == in.w) { i = 0; j++; }
	}
	return ce | co << 16;
}

// phase 3: the chunk's unknowns into the list, the even parity's first;
// pre = exclusive prefix of the packed counts over the threads, tot = sum
SP_HD void gsc_fill(int *L, const GsInst &in, int tid, int nt, int pre, int tot)
{
	int q0, q1;
	gsc_chunk(in.w, in.h, tid, nt, q0, q1);
	const int P = in.w + 2;
	int oe = pre & 0xffff, oo = (tot & 0xffff) + (pre >> 16);
	int j = q0 / in.w, i = q0 - j * in.w;
	for (int q = q0; q < q1; q++) {
		if (!sp_known(in.c[q])) L[(i + j) & 1 ? oo++ : oe++] = (j + 1) * P + i + 1;
		if (++i == in.w) { i = 0; j++; }
	}
}

// phase 4: the thread's pixels, entries tid, tid + nt, ... of each list
template <int MAXS>
SP_HD void gsc_take(GscThread<MAXS> &t, const float *U, const int *L,
		const GsInst &in, int tid, int nt, int tot)
{
	const int W = in.w, P = W + 2;
	const int nn[2] = {tot & 0xffff, tot >> 16};
#pragma unroll
	for (int p = 0; p < 2; p++) {
		t.n[p] = 0;
#pragma unroll
		for (int m = 0; m < MAXS; m++) {
			const int r = tid + m * nt;
			const bool on = r < nn[p];
			const int q = on ? L[(p ? nn[0] : 0) + r] : P + 1;   // unused: pixel 0
			const int j = q / P - 1, i = q - (j + 1) * P - 1;
			t.qij[p][m] = (unsigned)q | (unsigned)(i + j) << 16;
			t.x[p][m] = U[q];
			t.c[p][m] = in.c[j * W + i];
			t.n[p] += on;
		}
	}
}

// the update of gs_update_tile1 for one listed pixel (value x, coefficient
// cv, never known), same terms in the same order
SP_HD float gsc_update(const float *U, int q, int P, float x, float cv, float t)
{
	const float e = U[q + 1], so = U[q + P], w = U[q - 1], n = U[q - P];
	float r = 0.0f;
	if (isfinite(e)) r = r + (e - x);
	if (isfinite(so)) r = r + (so - x);
	if (isfinite(w)) r = r + (w - x);
	if (isfinite(n)) r = r + (n - x);
	return x + t * (r - cv);
}

// the thread's updates of step s (parity PAR); true when one of them
// changed its pixel's bits
template <int MAXS, int PAR>
SP_HD bool gsc_step_par(GscThread<MAXS> &th, float *U, int P, int s, int T2, float t)
{
	bool ch = false;
#pragma unroll
	for (int m = 0; m < MAXS; m++) {
		if (m >= th.n[PAR]) break;
		const int d = s - (int)(th.qij[PAR][m] >> 16);
		if ((unsigned)d >= (unsigned)T2) continue;
		const int q = (int)(th.qij[PAR][m] & 0xffffu);
		const float y = gsc_update(U, q, P, th.x[PAR][m], th.c[PAR][m], t);
		ch |= sp_bits(y) != sp_bits(th.x[PAR][m]);
		U[q] = y;
		th.x[PAR][m] = y;
	}
	return ch;
}

// phase 5, step s of small_nsteps(1, W, H, T): steps run in increasing s
// with a barrier between them
template <int MAXS>
SP_HD bool gsc_step(GscThread<MAXS> &th, float *U, int W, int T, float t, int s)
{
	if (s & 1) return gsc_step_par<MAXS, 1>(th, U, W + 2, s, 2 * T, t);
	return gsc_step_par<MAXS, 0>(th, U, W + 2, s, 2 * T, t);
}

// true when the level's values are final after step s, L being the last
// step on which an update changed its pixel (-1: none so far): sweep k's
// updates lie on steps [2k, 2k + W + H - 2], so the first sweep wholly
// after L has ended and changed nothing (a fixed point).  Such a sweep is
// always below the sweep count when s is a step of the schedule.
SP_HD bool gsc_final(int L, int W, int H, int s)
{
	return ((L + 2) & ~1) + W + H - 2 <= s;
}

// phase 6: the thread's pixels back to the image
template <int MAXS>
SP_HD void gsc_store(const GscThread<MAXS> &th, const GsInst &in)
{
	const int P = in.w + 2;
#pragma unroll
	for (int p = 0; p < 2; p++)
#pragma unroll
		for (int m = 0; m < MAXS; m++) {
			if (m >= th.n[p]) break;
			const int q = (int)(th.qij[p][m] & 0xffffu);
			const int j = q / P - 1, i = q - (j + 1) * P - 1;
			in.u[j * in.w + i] = th.x[p][m];
		}
}
