// Batched masked conjugate gradient with FP32 vectors and FP64 reductions.
//
// Same system as the reference's CG (cleant_cgpois.c:197-264, minicg.c:24-56):
// unknowns are the NaN pixels of g; A is -Laplacian (Poisson) or the
// 13-point bilaplacian, evaluated in float on float values exactly as the
// reference's operator does (its getter rounds each unknown to float); the
// right-hand side is b = -f - A(known pixels).  The difference from the
// reference is the storage of x, r, p, q in float (the reference keeps them
// in double); x stays in double, and dot products, norms and the stopping
// test are double.
//
// Many planes of uneven size run together: every kernel walks a list of
// work blocks (plane, first pixel), per-block partial sums land in a slot
// array whose slots are contiguous per plane, and one small kernel per
// reduction folds each plane's slots in a fixed order (deterministic).
// Planes that have converged are skipped by all later launches.
#include "sp_internal.h"
#include "sp_cg.h"
#include "../../include/simpois_hip.h"

namespace spk {

constexpr int NT = 256;
constexpr int PPT = 4;
static_assert(NT * PPT == 1024, "work blocks span 1024 pixels");

// Vector kernels below give each thread the 4 consecutive pixels
// start + 4 tid .. + 3 of its work block (plane vectors start 256-byte
// aligned).  Known pixels hold x = r = p = q = z = 0 throughout, so a group
// with at least one unknown is updated whole: the known entries stay 0.
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ unsigned code4(const unsigned char *code, unsigned q4)
{
	return *(const unsigned *)(code + q4);
}

static __device__ __forceinline__ int clampi(int v, int lo, int hi)
{
	return v < lo ? lo : (v > hi ? hi : v);
}

// Work-groups go to the 8 XCDs in turn (work-group b to XCD b % 8), each
// XCD with its own L2.  xcd_block() maps b to a logical block so that every
// XCD takes a contiguous run of logical blocks, in order: neighbouring
// tiles meet in the same L2 and share their halos (the multigrid tiles'
// HBM traffic halves).  A bijection of [0, gridDim.x): results do not
// depend on it.  Only for launches of uniform work per block: the CG
// vector kernels' work follows the holes, and contiguous runs of it
// unbalance the XCDs (measured slower there).
static __device__ __forceinline__ int xcd_block()
{
	const int n = (int)gridDim.x, b = (int)blockIdx.x;
	const int x = b & 7, q = n >> 3, r = n & 7;
	return x * q + (x < r ? x : r) + (b >> 3);
}

// For the CG vector kernels: every XCD takes runs of CH consecutive work
// blocks, the runs dealt to the XCDs in turn, so a block and the blocks of
// the rows above and below it mostly share one L2 while work that follows
// the holes still spreads over all XCDs.  A bijection of [0, gridDim.x)
// (the grid's last partial round of runs keeps its order); the partial-sum
// slot of a block follows the logical index, so sums do not depend on it.
// n = gridDim.x, passed as an argument: it then arrives with the other
// kernel arguments instead of one more dependent load.
constexpr int CG_XCD_RUN = 32;
static __device__ __forceinline__ int xcd_run_block(int n)
{
	const int b = (int)blockIdx.x;
	const int full = n / (8 * CG_XCD_RUN) * (8 * CG_XCD_RUN);
	if (b >= full) return b;
	const int x = b & 7, i = b >> 3;
	return ((i / CG_XCD_RUN) * 8 + x) * CG_XCD_RUN + i % CG_XCD_RUN;
}

template <int N>
static __device__ double bsum(double v, double *red)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) red[wv] = v;
	__syncthreads();
	double r = 0;
	if (threadIdx.x == 0)
		for (int k = 0; k < N / 64; k++) r += red[k];
	__syncthreads();
	return r;
}

// A at (i,j) over float values val(k), reference term order
template <typename V>
static __device__ __forceinline__ float op32(int bilap, int i, int j, int w, int h, V val)
{
	auto p = [&](int a, int b) {
		return val((size_t)clampi(b, 0, h - 1) * w + clampi(a, 0, w - 1));
	};
	if (!bilap) {
		float r = -4.0f * p(i, j);
		r = r + p(i + 1, j);
		r = r + p(i, j + 1);
		r = r + p(i - 1, j);
		r = r + p(i, j - 1);
		return -r;
	}
	float r = 20.0f * p(i, j);
	r = r - 8.0f * p(i + 1, j);
	r = r - 8.0f * p(i, j + 1);
	r = r - 8.0f * p(i - 1, j);
	r = r - 8.0f * p(i, j - 1);
	r = r + 2.0f * p(i + 1, j + 1);
	r = r + 2.0f * p(i + 1, j - 1);
	r = r + 2.0f * p(i - 1, j + 1);
	r = r + 2.0f * p(i - 1, j - 1);
	r = r + p(i + 2, j);
	r = r + p(i, j + 2);
	r = r + p(i - 2, j);
	r = r + p(i, j - 2);
	return r;
}

// The transposed 13-point operator at (i,j): the sum over the pixels
// (m,n) whose clamped taps land on (i,j) of the tap's coefficient times
// val(m,n), i.e. (A^T val)(i,j) over the whole plane.  Such pixels lie
// within two of (i,j) in each direction.  Four or more pixels from the
// border no tap of them clamps and the stencil is symmetric, so A^T = A
// there (op32's sum).  The adjoint of the biharmonic fill solves with it.
template <typename V>
static __device__ float op32T(int i, int j, int w, int h, V val)
{
	if (i >= 4 && j >= 4 && i + 4 < w && j + 4 < h) return op32(1, i, j, w, h, val);
	constexpr int DI[13] = {0, 1, 0, -1, 0, 1, 1, -1, -1, 2, 0, -2, 0};
	constexpr int DJ[13] = {0, 0, 1, 0, -1, 1, -1, 1, -1, 0, 2, 0, -2};
	constexpr float CF[13] = {20.0f, -8.0f, -8.0f, -8.0f, -8.0f, 2.0f, 2.0f, 2.0f, 2.0f,
		1.0f, 1.0f, 1.0f, 1.0f};
	float r = 0.0f;
	for (int n = max(j - 2, 0); n <= min(j + 2, h - 1); n++)
		for (int m = max(i - 2, 0); m <= min(i + 2, w - 1); m++) {
			float c = 0.0f;   // the entry of A in row (m,n), column (i,j)
#pragma unroll
			for (int t = 0; t < 13; t++)
				if (clampi(m + DI[t], 0, w - 1) == i && clampi(n + DJ[t], 0, h - 1) == j)
					c += CF[t];
			if (c != 0.0f) r += c * val((size_t)n * w + m);
		}
	return r;
}

// A (bilap 0 or 1) or A^T (bilap 2, the 13-point operator) at (i,j)
template <typename V>
static __device__ __forceinline__ float op32_any(int bilap, int i, int j, int w, int h, V val)
{
	return bilap == 2 ? op32T(i, j, w, h, val) : op32(bilap, i, j, w, h, val);
}

// pixel q of thread k-th visit inside work block b, or -1 past the plane
static __device__ __forceinline__ long pix(const BlockWork &b, int k, size_t n)
{
	const size_t q = (size_t)b.start + (size_t)k * NT + threadIdx.x;
	return q < n ? (long)q : -1;
}

// Level-0 code bytes of the pixels q4 .. q1-1 (at most 4, from the NaNs of
// g), packed in one word (pixel q4 + k in byte k).  A group inside one row
// of a 16-byte aligned plane (w a multiple of 4) reads its row and the rows
// above and below as one 16-byte load each.
// the code word of the group at (i, j) from its row c, the rows above (up)
// and below (dn) and the pixels left (lf) and right (rt) of it; a
// neighbour outside the plane is never looked at, so those may hold any value
static __device__ __forceinline__ unsigned row_code(f4 c, f4 up, f4 dn, float lf, float rt,
		unsigned i, unsigned j, unsigned w, unsigned h)
{
	const unsigned nrow = (j > 0) + (j + 1 < h);
	unsigned word = 0;
#pragma unroll
	for (int k = 0; k < 4; k++) {
		if (!isnan(c[k])) continue;
		const unsigned ii = i + k;
		unsigned v = 1;
		if (ii + 1 < w && isnan(k < 3 ? c[k + 1] : rt)) v |= 2;
		if (j + 1 < h && isnan(dn[k])) v |= 4;
		if (ii > 0 && isnan(k > 0 ? c[k - 1] : lf)) v |= 8;
		if (j > 0 && isnan(up[k])) v |= 16;
		v |= ((ii > 0) + (ii + 1 < w) + nrow) << 5;
		word |= v << (8 * k);
	}
	return word;
}

static __device__ __forceinline__ unsigned group_code(const Cg32Plane &P, unsigned q4,
		unsigned q1)
{
	const unsigned w = (unsigned)P.w, h = (unsigned)P.h;
	unsigned j = q4 / w, i = q4 - j * w, word = 0;
	if (q1 - q4 == 4 && !(w & 3) && !((uintptr_t)P.g & 15)) {
		const f4 c = *(const f4 *)(P.g + q4);
		const f4 up = j > 0 ? *(const f4 *)(P.g + q4 - w) : c;
		const f4 dn = j + 1 < h ? *(const f4 *)(P.g + q4 + w) : c;
		const float lf = i > 0 ? P.g[q4 - 1] : 0.0f;
		const float rt = i + 4 < w ? P.g[q4 + 4] : 0.0f;
		return row_code(c, up, dn, lf, rt, i, j, w, h);
	}
	for (unsigned q = q4; q < q1; q++) {
		unsigned c = 0;
		if (isnan(P.g[q])) {
			c = 1;
			if (i + 1 < w && isnan(P.g[q + 1])) c |= 2;
			if (j + 1 < h && isnan(P.g[q + w])) c |= 4;
			if (i > 0 && isnan(P.g[q - 1])) c |= 8;
			if (j > 0 && isnan(P.g[q - w])) c |= 16;
			c |= ((i > 0) + (i + 1 < w) + (j > 0) + (j + 1 < h)) << 5;
		}
		word |= c << (8 * (q - q4));
		if (++i == w) { i = 0; j++; }
	}
	return word;
}

// code bytes (above), b, r0 = b - A x0 (x0 = u at holes), x = x0; partial
// r.r and b.b
__global__ __launch_bounds__(NT) void k32_setup(const BlockWork *bw,
		const Cg32Plane *pl, double *part)
{
	__shared__ double red[NT / 64];
	const int blk = blockIdx.x;
	const BlockWork b = bw[blk];
	const Cg32Plane P = pl[b.plane];
	double rr = 0, bb = 0;
	const unsigned n = (unsigned)P.w * (unsigned)P.h, w = (unsigned)P.w;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	const unsigned q1 = q4 + 4 < n ? q4 + 4 : n;
	const bool whole = q4 + 3 < n;
	f4 rv = {0.0f, 0.0f, 0.0f, 0.0f};
	d4 xv = {0.0, 0.0, 0.0, 0.0};
	const float *g = P.g, *u = P.u;
	unsigned cc = 0;
	if (whole && !P.bilap && !(w & 3)
			&& !(((uintptr_t)g | (uintptr_t)u | (uintptr_t)P.f) & 15)) {
		// a group inside one row of 16-byte aligned planes: the clamped
		// 5-point reads of g as one 16-byte load per row and a scalar on
		// each side, all at once; the codes from them; then, for a group
		// with an unknown, the same reads of u and f at once; the same sums
		// in the same order as below.  With u0 only its values at unknowns
		// enter, all 0: u is not read.
		const unsigned h = (unsigned)P.h, j = q4 / w, i = q4 - j * w;
		const unsigned qn = j > 0 ? q4 - w : q4, qs = j + 1 < h ? q4 + w : q4;
		const unsigned qw = i > 0 ? q4 - 1 : q4, qe = i + 4 < w ? q4 + 4 : q4 + 3;
		const f4 z0 = {0.0f, 0.0f, 0.0f, 0.0f};
		const bool ru = !P.u0;
		const f4 gc = gload<f4>(g + q4), gn = gload<f4>(g + qn), gs = gload<f4>(g + qs);
		const float gw = gload<float>(g + qw), ge = gload<float>(g + qe);
		cc = row_code(gc, gn, gs, gw, ge, i, j, w, h);
		if (cc & 0x01010101u) {
			f4 uc = z0, un = z0, us = z0;
			float uw = 0.0f, ue = 0.0f;
			if (ru) {
				uc = gload<f4>(u + q4);
				un = gload<f4>(u + qn);
				us = gload<f4>(u + qs);
				uw = gload<float>(u + qw);
				ue = gload<float>(u + qe);
			}
			const f4 fv = P.f ? gload<f4>(P.f + q4) : z0;
#pragma unroll
			for (int k = 0; k < 4; k++) {
				if (!((cc >> (8 * k)) & 1u)) continue;
				const float gv[4] = {k < 3 ? gc[k + 1] : ge, gs[k], k > 0 ? gc[k - 1] : gw, gn[k]};
				const float uv[4] = {k < 3 ? uc[k + 1] : ue, us[k], k > 0 ? uc[k - 1] : uw, un[k]};
				float r0 = -4.0f * 0.0f, rx = -4.0f * uc[k];
#pragma unroll
				for (int d = 0; d < 4; d++) {
					const bool kn = !isnan(gv[d]);
					r0 = r0 + (kn ? gv[d] : 0.0f);
					rx = rx + (kn ? 0.0f : uv[d]);
				}
				const float bv = (P.f ? -fv[k] : 0.0f) - -r0;
				const float r = bv - -rx;
				rr += (double)r * r;
				bb += (double)bv * bv;
				rv[k] = r;
				xv[k] = (double)uc[k];
			}
		}
	} else {
		cc = q4 < n ? group_code(P, q4, q1) : 0u;
		if (!whole || (cc & 0x01010101u)) {
			// coordinates of the group's first pixel; the others follow along
			// the row (a group may wrap onto the next row)
			unsigned j = q4 / w, i = q4 - j * w;
			for (unsigned q = q4; q < q1; q++) {
				const unsigned k = q - q4;
				const bool hole = (cc >> (8 * k)) & 1u;
				if (!whole) P.code[q] = (unsigned char)(cc >> (8 * k));
				if (hole) {
					float a0, ax;
					if (!P.bilap) {
						// op32's clamped 5-point sum in its term order, both fields
						// (known values, unknowns) from one set of loads
						const unsigned t[4] = {i + 1 < w ? q + 1 : q,
							j + 1 < (unsigned)P.h ? q + w : q, i > 0 ? q - 1 : q,
							j > 0 ? q - w : q};
						float gv[4], uv[4];
#pragma unroll
						for (int d = 0; d < 4; d++) { gv[d] = g[t[d]]; uv[d] = u[t[d]]; }
						float r0 = -4.0f * 0.0f, rx = -4.0f * u[q];
#pragma unroll
						for (int d = 0; d < 4; d++) {
							const bool kn = !isnan(gv[d]);
							r0 = r0 + (kn ? gv[d] : 0.0f);
							rx = rx + (kn ? 0.0f : uv[d]);
						}
						a0 = -r0;
						ax = -rx;
					} else {
						a0 = op32_any(P.bilap, (int)i, (int)j, P.w, P.h, [&](size_t tt) {
							const float v = g[tt]; return isnan(v) ? 0.0f : v; });
						ax = op32_any(P.bilap, (int)i, (int)j, P.w, P.h, [&](size_t tt) {
							return isnan(g[tt]) ? u[tt] : 0.0f; });
					}
					const float bv = (P.f ? -P.f[q] : 0.0f) - a0;
					const float r = bv - ax;
					rr += (double)r * r;
					bb += (double)bv * bv;
					if (whole) { rv[k] = r; xv[k] = (double)u[q]; }
					else { P.r[q] = r; P.p[q] = r; P.q[q] = 0.0f; P.x[q] = (double)u[q]; }
				} else if (!whole) {
					P.r[q] = 0.0f; P.p[q] = 0.0f; P.q[q] = 0.0f; P.x[q] = 0.0;
				}
				if (++i == w) { i = 0; j++; }
			}
		}
	}
	if (whole) {
		// known entries are 0.  r is read everywhere (the multigrid's
		// right-hand side), x only in groups with an unknown.  The
		// biharmonic iteration starts from p = r, and its product reads p
		// at every pixel of the 13-point stencil, unknown or not: p and q
		// everywhere.  (The Poisson iteration writes p and q itself.)
		gstore<unsigned>(P.code + q4, cc);
		gstore<f4>(P.r + q4, rv);
		if (P.bilap) {
			gstore<f4>(P.p + q4, rv);
			gstore<f4>(P.q + q4, f4{0.0f, 0.0f, 0.0f, 0.0f});
		}
		if (cc & 0x01010101u) gstore<d4>(P.x + q4, xv);
	}
	const double s1 = bsum<NT>(rr, red), s2 = bsum<NT>(bb, red);
	if (threadIdx.x == 0) {
		part[2 * blk] = s1;
		part[2 * blk + 1] = s2;
	}
}

// Per-plane folds run one work-group of RT threads per plane: a plane has
// up to ~16K partial sums (one per work block), and the fold is bound by
// the latency of its loads.
constexpr int RT = 1024;

// one block per plane: fold the plane's slots in a fixed order
static __device__ void fold(const double *part, const Cg32Plane &P, int stride,
		int off, double *red, double *out)
{
	// eight loads in flight per thread; slot k of the plane always lands in
	// accumulator ((k / RT) % 8) of thread k % RT, so the order is fixed
	double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	const double *pp = part + (size_t)stride * P.blk0 + off;
	int k = threadIdx.x;
	for (; k + 7 * RT < P.nblk; k += 8 * RT)
#pragma unroll
		for (int u = 0; u < 8; u++) v[u] += pp[(size_t)stride * (k + u * RT)];
#pragma unroll
	for (int u = 0; u < 7; u++)
		if (k + u * RT < P.nblk) v[u] += pp[(size_t)stride * (k + u * RT)];
	const double t = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
	const double s = bsum<RT>(t, red);
	if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(RT) void k32_reduce0(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part, int *nactive)
{
	__shared__ double red[RT / 64];
	__shared__ double rr, bb;
	const Cg32Plane P = pl[blockIdx.x];
	fold(part, P, 2, 0, red, &rr);
	fold(part, P, 2, 1, red, &bb);
	__syncthreads();
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		s.rr = rr;
		s.rz = rr;   // replaced by (r, M r) when preconditioned
		s.bnorm = sqrt(bb);
		s.stop = s.rel ? s.eps * s.bnorm : s.eps;
		s.iters = 0;
		const int n = *P.cnt;
		if (s.maxit_from_count) s.maxit = (int)(float)n;
		// a plane stops before its first iteration when it has nothing to
		// solve, no iteration allowed, a residual that is not finite (NaN
		// or Inf in the data its system reads) or exactly zero (x0 solves
		// it; the reference's 0/0, minicg.c:40-42, is not reproduced here)
		s.status = n <= 0 || rr == 0.0 ? SP_PLANE_CONVERGED
			: !isfinite(rr) ? SP_PLANE_NONFINITE
			: s.maxit <= 0 ? SP_PLANE_MAXIT : 0;
		s.active = s.status == 0;
		if (s.active) atomicAdd(nactive, 1);
	}
}

// The loads of a plane's descriptor go out with those of its scalars: used
// only past a kernel's early exit, the compiler would sink them there and
// the work-group would wait for one more dependent load.
static __device__ __forceinline__ void cg32_early(const Cg32Plane &P)
{
	asm volatile("" :: "s"(P.code), "s"(P.x), "s"(P.r), "s"(P.p), "s"(P.q), "s"(P.p2),
			"s"(P.z), "s"(P.u), "s"(P.w), "s"(P.h));
}

// q = A p, partial p.q
__global__ __launch_bounds__(NT) void k32_spmv(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, double *part, int nblk)
{
	__shared__ double red[NT / 64];
	const int blk = xcd_run_block(nblk);
	const BlockWork b = bw[blk];
	if (!sc[b.plane].active) return;
	const Cg32Plane P = pl[b.plane];
	const float *p = P.p;
	double acc = 0;
	const unsigned n = (unsigned)P.w * (unsigned)P.h, w = (unsigned)P.w;
	if (!P.bilap) {
		// -Laplacian with clamped reads: a clamped neighbour is the pixel
		// itself, p is 0 at known pixels, so A p = d p - (sum of the
		// in-image unknown neighbours), d = in-image neighbours; the code
		// byte gives both, no coordinates needed
		const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
		if (q4 + 3 < n) {
			const unsigned cc = code4(P.code, q4);
			if (cc & 0x01010101u) {
				const f4 pc = *(const f4 *)(p + q4);
				f4 pn, ps;
				if (!(w & 3) && q4 >= w) pn = *(const f4 *)(p + q4 - w);
				else
#pragma unroll
					for (int k = 0; k < 4; k++)
						pn[k] = (cc >> (8 * k)) & 16 ? p[q4 + k - w] : 0.0f;
				if (!(w & 3) && q4 + w + 3 < n) ps = *(const f4 *)(p + q4 + w);
				else
#pragma unroll
					for (int k = 0; k < 4; k++)
						ps[k] = (cc >> (8 * k)) & 4 ? p[q4 + k + w] : 0.0f;
				const float pe = cc & (2u << 24) ? p[q4 + 4] : 0.0f;
				const float pw = cc & 8u ? p[q4 - 1] : 0.0f;
				f4 y;
#pragma unroll
				for (int k = 0; k < 4; k++) {
					const unsigned c = (cc >> (8 * k)) & 255u;
					const float e = k < 3 ? pc[k + 1] : pe, ww = k > 0 ? pc[k - 1] : pw;
					float nb = 0.0f;
					if (c & 2) nb += e;
					if (c & 4) nb += ps[k];
					if (c & 8) nb += ww;
					if (c & 16) nb += pn[k];
					y[k] = (float)(c >> 5) * pc[k] - nb;
					acc += (double)y[k] * pc[k];
				}
				*(f4 *)(P.q + q4) = y;
			}
		} else {
			for (unsigned q = q4; q < n; q++) {
				const unsigned c = P.code[q];
				if (!(c & 1)) continue;
				const float pc = p[q];
				float nb = 0.0f;
				if (c & 2) nb += p[q + 1];
				if (c & 4) nb += p[q + w];
				if (c & 8) nb += p[q - 1];
				if (c & 16) nb += p[q - w];
				const float y = (float)(c >> 5) * pc - nb;
				P.q[q] = y;
				acc += (double)y * pc;
			}
		}
	} else {
		for (int k_ = 0; k_ < PPT; k_++) {
			const unsigned q = (unsigned)b.start + k_ * NT + threadIdx.x;
			if (q >= n) break;
			if (!(P.code[q] & 1)) continue;
			const unsigned j = q / w, i = q - j * w;
			const float y = op32_any(P.bilap, (int)i, (int)j, P.w, P.h,
					[&](size_t k) { return p[k]; });
			P.q[q] = y;
			acc += (double)y * p[q];
		}
	}
	const double s = bsum<NT>(acc, red);
	if (threadIdx.x == 0) part[2 * blk] = s;
}

__global__ __launch_bounds__(RT) void k32_reduce_a(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part, int *nactive)
{
	__shared__ double red[RT / 64];
	__shared__ double pq;
	if (!sc[blockIdx.x].active) {
		if (threadIdx.x == 0) sc[blockIdx.x].upd = sc[blockIdx.x].fin = 0;
		return;
	}
	fold(part, pl[blockIdx.x], 2, 0, red, &pq);
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		s.alpha_prev = s.alpha;
		// breakdown (p.Ap = 0 or not finite once float round-off has taken
		// over): keep the current x and stop this plane
		if (!(pq > 0) || !isfinite(s.rz / pq)) {
			s.alpha = 0.0;
			s.upd = 0;
			s.fin = 1;
			s.active = 0;
			s.status = isfinite(pq) ? SP_PLANE_BREAKDOWN : SP_PLANE_NONFINITE;
			atomicSub(nactive, 1);
		} else {
			s.alpha = s.rz / pq;
			s.upd = 1;
		}
	}
}

// r -= alpha q ; partial r.r  (x += alpha p follows in k32_dir, which
// reads p anyway)
__global__ __launch_bounds__(NT) void k32_update(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, double *part, int nblk)
{
	__shared__ double red[NT / 64];
	const int blk = xcd_run_block(nblk);
	const BlockWork b = bw[blk];
	const Cg32Plane P = pl[b.plane];
	const Cg32Scal &s = sc[b.plane];
	const bool act = s.active;
	const double ad = s.alpha;
	cg32_early(P);
	asm volatile("" :: "s"(ad));
	if (!act) return;
	const float a = (float)ad;
	double acc = 0;
	const unsigned n = (unsigned)P.w * (unsigned)P.h;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	if (q4 + 3 < n) {
		if (gload<unsigned>(P.code + q4) & 0x01010101u) {
			const f4 qq = gload<f4>(P.q + q4);
			f4 r = gload<f4>(P.r + q4);
#pragma unroll
			for (int k = 0; k < 4; k++) {
				r[k] = r[k] - a * qq[k];
				acc += (double)r[k] * r[k];
			}
			gstore<f4>(P.r + q4, r);
		}
	} else {
		for (unsigned q = q4; q < n; q++) {
			if (!(P.code[q] & 1)) continue;
			const float r = P.r[q] - a * P.q[q];
			P.r[q] = r;
			acc += (double)r * r;
		}
	}
	const double t = bsum<NT>(acc, red);
	if (threadIdx.x == 0) part[2 * blk + 1] = t;
}

// rr_new, absolute stop test on sqrt(rr_new), beta
__global__ __launch_bounds__(RT) void k32_reduce_b(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part, int *nactive, double *trace, int precond)
{
	__shared__ double red[RT / 64];
	__shared__ double rr1;
	if (!sc[blockIdx.x].active) return;
	fold(part, pl[blockIdx.x], 2, 1, red, &rr1);
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		if (trace && s.iters < s.trace_cap) trace[s.trace0 + s.iters] = rr1;
		s.iters++;
		if (!precond) {   // z = r: rho = r.r
			s.beta = rr1 / s.rz;
			s.rz = rr1;
		}
		s.rr = rr1;
		const int st = sqrt(rr1) < s.stop ? SP_PLANE_CONVERGED
			: !isfinite(rr1) ? SP_PLANE_NONFINITE
			: s.iters >= s.maxit ? SP_PLANE_MAXIT : 0;
		if (st) {
			s.status = st;
			s.active = 0;
			s.fin = 1;
			atomicSub(nactive, 1);
		}
	}
}

// x += alpha p when this iteration updated r (also in the iteration that
// stops the plane); p = z + beta p while the plane goes on (z = r without
// preconditioner)
__global__ __launch_bounds__(NT) void k32_dir(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, int nblk)
{
	const int blk = xcd_run_block(nblk);
	const BlockWork b = bw[blk];
	const Cg32Scal &s = sc[b.plane];
	const bool upd = s.upd, act = s.active;
	if (!upd && !act) return;
	const Cg32Plane P = pl[b.plane];
	const double ad = s.alpha;
	const float be = (float)s.beta;
	const unsigned n = (unsigned)P.w * (unsigned)P.h;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	if (q4 + 3 < n) {
		if (code4(P.code, q4) & 0x01010101u) {
			f4 pp = *(const f4 *)(P.p + q4);
			if (upd) {
				d4 x = *(const d4 *)(P.x + q4);
#pragma unroll
				for (int k = 0; k < 4; k++) x[k] = x[k] + ad * (double)pp[k];
				*(d4 *)(P.x + q4) = x;
			}
			if (act) {
				const f4 z = *(const f4 *)(P.z + q4);
#pragma unroll
				for (int k = 0; k < 4; k++) pp[k] = z[k] + be * pp[k];
				*(f4 *)(P.p + q4) = pp;
			}
		}
	} else {
		for (unsigned q = q4; q < n; q++) {
			if (!(P.code[q] & 1)) continue;
			if (upd) P.x[q] = P.x[q] + ad * (double)P.p[q];
			if (act) P.p[q] = P.z[q] + be * P.p[q];
		}
	}
}

// Direction and product in one pass (Poisson operator):
//   x += alpha p_old            when this iteration updated r (upd), in
//                               pairs: see below
//   p_new = z + beta p_old      while the plane goes on (first: p_new = z)
//   q = A p_new, partial p_new.q
// p_new of the unknown neighbours is recomputed from their z and p_old
// with the same operations as their owners, so it has the same bits.
// Neighbours read p_old while owners write p_new: p_old and p_new are two
// buffers, p (par 1) or p2 (par 0) receiving p_new.  Every value is taken
// under the unknown bits, so known pixels and groups without unknowns may
// hold anything.
//
// x is read and written every other pass only.  Pass T >= 1 (par 0 for T
// odd, 1 for T even) owes x the term alpha_{T-1} p_{T-1} (p_old).  An odd
// pass leaves it pending; an even pass adds the pending alpha_{T-2} p_{T-2}
// (still in the buffer it is about to overwrite with p_new, read first)
// and then its own term, the same two double additions in the same order
// as one per pass, so x and u keep their bits.  The pass that stops a plane
// adds what is owed and writes u (x is not read after that).
static __device__ __forceinline__ float pnew(float z, float po, float be, bool first)
{
	return first ? z : z + be * po;
}

__global__ __launch_bounds__(NT) void k32_dirspmv(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, double *part, int par, int first,
		int nblk)
{
	__shared__ double red[NT / 64];
	const int blk = xcd_run_block(nblk);
	const BlockWork b = bw[blk];
	// the plane's descriptor loads alongside its scalars
	const Cg32Plane P = pl[b.plane];
	const Cg32Scal &s = sc[b.plane];
	const int su = s.upd, sa = s.active, sf = s.fin;
	const bool upd = !first && su, act = sa, fin = !first && sf;
	cg32_early(P);
	if (!upd && !act && !fin) return;
	const float *po = par ? P.p2 : P.p;
	float *pw_ = par ? P.p : P.p2;
	const float be = (float)s.beta;
	const double ad = s.alpha, ad2 = s.alpha_prev;
	const bool fst = first;
	// even pass: the pending term comes in (pair), x is loaded, stored while
	// the plane goes on; odd pass: x only for the stopping pass
	const bool pair = par && !fst && (upd || fin);
	const bool ldx = pair || fin, stx = pair && !fin;
	double acc = 0;
	const unsigned n = (unsigned)P.w * (unsigned)P.h, w = (unsigned)P.w;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	const f4 z0 = {0.0f, 0.0f, 0.0f, 0.0f};
	if (q4 + 3 < n && !(w & 3)) {
		// rows of 16-byte groups: every load of the group goes out before
		// the first use.  The rows above and below and the east / west
		// pixels are clamped into the plane; their values are taken under
		// the code bits only, which never name a pixel outside it.
		const unsigned cc = gload<unsigned>(P.code + q4);
		if (cc & 0x01010101u) {
			const unsigned qn = q4 >= w ? q4 - w : q4, qs = q4 + w + 3 < n ? q4 + w : q4;
			const unsigned qe = q4 + 4 < n ? q4 + 4 : q4, qw = q4 > 0 ? q4 - 1 : q4;
			f4 poc = z0, on = z0, os = z0, zc = z0, zn = z0, zs = z0, pp2 = z0;
			float oe = 0.0f, ow = 0.0f, ze = 0.0f, zw = 0.0f;
			d4 x = {0.0, 0.0, 0.0, 0.0};
			if (!fst) {
				poc = gload<f4>(po + q4);
				if (act) {
					on = gload<f4>(po + qn);
					os = gload<f4>(po + qs);
					oe = gload<float>(po + qe);
					ow = gload<float>(po + qw);
				}
			}
			if (pair) pp2 = gload<f4>(pw_ + q4);
			if (ldx) x = gload<d4>(P.x + q4);
			if (act) {
				zc = gload<f4>(P.z + q4);
				zn = gload<f4>(P.z + qn);
				zs = gload<f4>(P.z + qs);
				ze = gload<float>(P.z + qe);
				zw = gload<float>(P.z + qw);
			}
			// one straight stretch of arithmetic (the flags select), then
			// the stores: a use of loaded values after a branch would make
			// the compiler wait for every access in flight, stores included
#pragma unroll
			for (int k = 0; k < 4; k++) {
				if (pair && ((cc >> (8 * k)) & 1u)) x[k] = x[k] + ad2 * (double)pp2[k];
				if (upd && ldx && ((cc >> (8 * k)) & 1u)) x[k] = x[k] + ad * (double)poc[k];
			}
			f4 pc, pn, ps;
#pragma unroll
			for (int k = 0; k < 4; k++) {
				const unsigned c = act ? (cc >> (8 * k)) & 255u : 0u;
				pc[k] = c & 1u ? pnew(zc[k], poc[k], be, fst) : 0.0f;
				pn[k] = c & 16u ? pnew(zn[k], on[k], be, fst) : 0.0f;
				ps[k] = c & 4u ? pnew(zs[k], os[k], be, fst) : 0.0f;
			}
			const float pe = act && (cc & (2u << 24)) ? pnew(ze, oe, be, fst) : 0.0f;
			const float pw = act && (cc & 8u) ? pnew(zw, ow, be, fst) : 0.0f;
			f4 y;
#pragma unroll
			for (int k = 0; k < 4; k++) {
				const unsigned c = (cc >> (8 * k)) & 255u;
				const float e = k < 3 ? pc[k + 1] : pe, ww = k > 0 ? pc[k - 1] : pw;
				float nb = 0.0f;
				if (c & 2) nb += e;
				if (c & 4) nb += ps[k];
				if (c & 8) nb += ww;
				if (c & 16) nb += pn[k];
				y[k] = (float)(c >> 5) * pc[k] - nb;
				if (act) acc += (double)y[k] * pc[k];
			}
			if (act) {
				gstore<f4>(pw_ + q4, pc);
				gstore<f4>(P.q + q4, y);
			}
			if (stx) gstore<d4>(P.x + q4, x);
			// the solution: u may be the caller's buffer, 4-byte stores at
			// the unknowns only
			if (fin)
#pragma unroll
				for (int k = 0; k < 4; k++)
					if ((cc >> (8 * k)) & 1u) gstore<float>(P.u + q4 + k, (float)x[k]);
		}
	} else if (q4 + 3 < n) {
		// rows that do not start 16-byte aligned: neighbours one by one
		const unsigned cc = code4(P.code, q4);
		if (cc & 0x01010101u) {
			const f4 poc = fst ? z0 : *(const f4 *)(po + q4);
			if (ldx) {
				d4 x = *(const d4 *)(P.x + q4);
				if (pair) {
					const f4 pp2 = *(const f4 *)(pw_ + q4);   // before p_new replaces it
#pragma unroll
					for (int k = 0; k < 4; k++)
						if ((cc >> (8 * k)) & 1u) x[k] = x[k] + ad2 * (double)pp2[k];
				}
				if (upd) {
#pragma unroll
					for (int k = 0; k < 4; k++)
						if ((cc >> (8 * k)) & 1u) x[k] = x[k] + ad * (double)poc[k];
				}
				if (stx) *(d4 *)(P.x + q4) = x;
				if (fin)
#pragma unroll
					for (int k = 0; k < 4; k++)
						if ((cc >> (8 * k)) & 1u) P.u[q4 + k] = (float)x[k];
			}
			if (act) {
				const f4 zc = *(const f4 *)(P.z + q4);
				f4 pc, pn, ps;
#pragma unroll
				for (int k = 0; k < 4; k++) {
					pc[k] = (cc >> (8 * k)) & 1u ? pnew(zc[k], poc[k], be, fst) : 0.0f;
					pn[k] = (cc >> (8 * k)) & 16u ? pnew(P.z[q4 + k - w],
							fst ? 0.0f : po[q4 + k - w], be, fst) : 0.0f;
					ps[k] = (cc >> (8 * k)) & 4u ? pnew(P.z[q4 + k + w],
							fst ? 0.0f : po[q4 + k + w], be, fst) : 0.0f;
				}
				const float pe = cc & (2u << 24) ? pnew(P.z[q4 + 4],
						fst ? 0.0f : po[q4 + 4], be, fst) : 0.0f;
				const float pw = cc & 8u ? pnew(P.z[q4 - 1],
						fst ? 0.0f : po[q4 - 1], be, fst) : 0.0f;
				f4 y;
#pragma unroll
				for (int k = 0; k < 4; k++) {
					const unsigned c = (cc >> (8 * k)) & 255u;
					const float e = k < 3 ? pc[k + 1] : pe, ww = k > 0 ? pc[k - 1] : pw;
					float nb = 0.0f;
					if (c & 2) nb += e;
					if (c & 4) nb += ps[k];
					if (c & 8) nb += ww;
					if (c & 16) nb += pn[k];
					y[k] = (float)(c >> 5) * pc[k] - nb;
					acc += (double)y[k] * pc[k];
				}
				*(f4 *)(pw_ + q4) = pc;
				*(f4 *)(P.q + q4) = y;
			}
		}
	} else {
		for (unsigned q = q4; q < n; q++) {
			const unsigned c = P.code[q];
			if (!(c & 1)) {
				if (act) { pw_[q] = 0.0f; P.q[q] = 0.0f; }
				continue;
			}
			const float o = fst ? 0.0f : po[q];
			if (ldx) {
				double x = P.x[q];
				if (pair) x = x + ad2 * (double)pw_[q];   // before p_new replaces it
				if (upd) x = x + ad * (double)o;
				if (stx) P.x[q] = x;
				if (fin) P.u[q] = (float)x;
			}
			if (!act) continue;
			const float pc = pnew(P.z[q], o, be, fst);
			float nb = 0.0f;
			if (c & 2) nb += pnew(P.z[q + 1], fst ? 0.0f : po[q + 1], be, fst);
			if (c & 4) nb += pnew(P.z[q + w], fst ? 0.0f : po[q + w], be, fst);
			if (c & 8) nb += pnew(P.z[q - 1], fst ? 0.0f : po[q - 1], be, fst);
			if (c & 16) nb += pnew(P.z[q - w], fst ? 0.0f : po[q - w], be, fst);
			const float y = (float)(c >> 5) * pc - nb;
			pw_[q] = pc;
			P.q[q] = y;
			acc += (double)y * pc;
		}
	}
	const double t = bsum<NT>(acc, red);
	if (threadIdx.x == 0) part[2 * blk] = t;
}

__global__ __launch_bounds__(NT) void k32_finish(const BlockWork *bw,
		const Cg32Plane *pl)
{
	const BlockWork b = bw[blockIdx.x];
	const Cg32Plane P = pl[b.plane];
	const unsigned n = (unsigned)P.w * (unsigned)P.h;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	if (q4 + 3 < n) {
		const unsigned cc = code4(P.code, q4);
		if (!(cc & 0x01010101u)) return;
		// u may be the caller's buffer: 4-byte accesses, known pixels kept
#pragma unroll
		for (int k = 0; k < 4; k++)
			if ((cc >> (8 * k)) & 1u) P.u[q4 + k] = (float)P.x[q4 + k];
	} else {
		for (unsigned q = q4; q < n; q++)
			if (P.code[q] & 1) P.u[q] = (float)P.x[q];
	}
}

// ------------------------------------------------------------ multigrid
//
// Preconditioner M ~ A^-1: one symmetric V-cycle over a hierarchy built by
// aggregation.  Level l+1 cell (I,J) aggregates the active cells of
// (2I..2I+1, 2J..2J+1); the interpolation P is piecewise constant and the
// coarse operator is the exact Galerkin product P^T A P, again a symmetric
// 5-point operator stored as a diagonal D and east/south couplings wE, wS
// (level 0 is the masked -Laplacian itself, read from the mask).  Smoother:
// red-black Gauss-Seidel, red then black before the coarse correction and
// black then red after it, so that M is symmetric; the coarse correction is
// scaled by wc.  The coarsest level runs many such sweeps.

// level-0 coefficients from the code byte: D = in-image neighbours, w = 1
// between two unknowns; levels > 0 from their stored Galerkin coefficients
static __device__ __forceinline__ float diag_at(const MgLevel &L, int i, int j,
		bool l0)
{
	const size_t q = (size_t)j * L.w + i;
	return l0 ? (float)(L.code[q] >> 5) : L.D[q];   // code 0 at known pixels
}

static __device__ __forceinline__ float we_at(const MgLevel &L, int i, int j, bool l0)
{   // coupling between (i,j) and (i+1,j)
	if (i + 1 >= L.w) return 0.0f;
	const size_t q = (size_t)j * L.w + i;
	return l0 ? (float)((L.code[q] >> 1) & 1u) : L.wE[q];
}

static __device__ __forceinline__ float ws_at(const MgLevel &L, int i, int j, bool l0)
{   // coupling between (i,j) and (i,j+1)
	if (j + 1 >= L.h) return 0.0f;
	const size_t q = (size_t)j * L.w + i;
	return l0 ? (float)((L.code[q] >> 2) & 1u) : L.wS[q];
}

// sum of couplings times z over the four neighbours of (i,j)
static __device__ __forceinline__ float offdiag(const MgLevel &L, int i, int j,
		bool l0)
{
	const size_t q = (size_t)j * L.w + i;
	float s = 0.0f;
	if (i + 1 < L.w) s += we_at(L, i, j, l0) * L.z[q + 1];
	if (i > 0) s += we_at(L, i - 1, j, l0) * L.z[q - 1];
	if (j + 1 < L.h) s += ws_at(L, i, j, l0) * L.z[q + L.w];
	if (j > 0) s += ws_at(L, i, j - 1, l0) * L.z[q - L.w];
	return s;
}

// Galerkin coarse operator of the next level (one thread per coarse cell)
__global__ __launch_bounds__(NT) void k_mg_coarsen(const BlockWork *bw,
		const MgLevel *fine, const MgLevel *coarse, int l0)
{
	const BlockWork b = bw[blockIdx.x];
	const MgLevel F = fine[b.plane], C = coarse[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)C.w * C.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const unsigned Ju = (unsigned)q / (unsigned)C.w;
		const int I = (int)((unsigned)q - Ju * (unsigned)C.w), J = (int)Ju;
		float d = 0, inner = 0, e = 0, sth = 0;
		for (int c = 0; c < 4; c++) {
			const int i = 2 * I + (c & 1), j = 2 * J + (c >> 1);
			if (i >= F.w || j >= F.h) continue;
			d += diag_at(F, i, j, l0);
			const float we = we_at(F, i, j, l0), ws = ws_at(F, i, j, l0);
			if (c & 1) e += we; else inner += we;      // east edge leaves the cell
			if (c >> 1) sth += ws; else inner += ws;   // south edge leaves the cell
		}
		C.D[q] = d - 2.0f * inner;
		C.wE[q] = I + 1 < C.w ? e : 0.0f;
		C.wS[q] = J + 1 < C.h ? sth : 0.0f;
		C.z[q] = 0.0f;
	}
}

// Level 0 -> 1 in one pass: the Galerkin coefficients of level 1 straight
// from the code bytes, and their packed words (what k_mg_coarsen and
// k_mg_gcode give, in the same integers).  Coarse cell (I, J) aggregates
// fine cells c0 = (2I, 2J), c1 = (2I+1, 2J), c2 = (2I, 2J+1), c3 = (2I+1,
// 2J+1): D = sum of their diagonals - 2 (E(c0) + S(c0) + S(c1) + E(c2)),
// wE = E(c1) + E(c3), wS = S(c2) + S(c3), with E, S the unknown-neighbour
// bits (0 at the image border).
struct Coarse0 { unsigned d, e, s; };

static __device__ __forceinline__ Coarse0 coarse0_of(unsigned c0, unsigned c1,
		unsigned c2, unsigned c3)
{
	auto E = [](unsigned c) { return (c >> 1) & 1u; };
	auto S = [](unsigned c) { return (c >> 2) & 1u; };
	Coarse0 r;
	r.d = (c0 >> 5) + (c1 >> 5) + (c2 >> 5) + (c3 >> 5)
		- 2u * (E(c0) + S(c0) + S(c1) + E(c2));
	r.e = E(c1) + E(c3);
	r.s = S(c2) + S(c3);
	return r;
}

// fine code byte at (i, j), 0 outside the fine level
static __device__ __forceinline__ unsigned code_at(const MgLevel &F, int i, int j)
{
	return i >= 0 && j >= 0 && i < F.w && j < F.h ? F.code[(size_t)j * F.w + i] : 0u;
}

static __device__ __forceinline__ Coarse0 coarse0_cell(const MgLevel &F, int I, int J)
{
	return coarse0_of(code_at(F, 2 * I, 2 * J), code_at(F, 2 * I + 1, 2 * J),
			code_at(F, 2 * I, 2 * J + 1), code_at(F, 2 * I + 1, 2 * J + 1));
}

static __device__ __forceinline__ void coarse0_store(const MgLevel &C, size_t q,
		int I, int J, Coarse0 v, unsigned we, unsigned no)
{
	const unsigned e = I + 1 < C.w ? v.e : 0u, so = J + 1 < C.h ? v.s : 0u;
	C.D[q] = (float)v.d;
	C.wE[q] = (float)e;
	C.wS[q] = (float)so;
	C.z[q] = 0.0f;
	C.gcode[q] = v.d | e << 11 | so << 16 | we << 21 | no << 26;
}

__global__ __launch_bounds__(NT) void k_mg_coarsen0(const BlockWork *bw,
		const MgLevel *fine, const MgLevel *coarse)
{
	const BlockWork b = bw[blockIdx.x];
	const MgLevel F = fine[b.plane], C = coarse[b.plane];
	const unsigned n = (unsigned)C.w * (unsigned)C.h, cw = (unsigned)C.w;
	const unsigned q4 = (unsigned)b.start + 4 * threadIdx.x;
	if (q4 >= n) return;
	const int J = (int)(q4 / cw), I = (int)(q4 - (unsigned)J * cw);
	if (!(cw & 3) && F.w == 2 * C.w) {
		// four cells of one row: 8 fine columns of fine rows 2J - 1 .. 2J + 1
		// as one 8-byte load each (the fine width is a multiple of 8)
		// (every load at once, rows and columns outside the level clamped to
		// the group's own and their values dropped by selects)
		const unsigned char *r0 = F.code + (size_t)(2 * J) * F.w + 2 * I;
		const bool hs = 2 * J + 1 < F.h, hn = J > 0, hw = I > 0;
		const gu2 a = gload<gu2>(r0);
		const gu2 cl = gload<gu2>(hs ? r0 + F.w : r0), ul = gload<gu2>(hn ? r0 - F.w : r0);
		const unsigned w0 = gload<unsigned char>(hw ? r0 - 1 : r0);
		const unsigned w1 = gload<unsigned char>(hw && hs ? r0 + F.w - 1 : r0);
		const gu2 z2 = {0u, 0u};
		const gu2 c = hs ? cl : z2, u = hn ? ul : z2;
		auto byte = [](gu2 v, int k) { return ((k < 4 ? v.x : v.y) >> (8 * (k & 3))) & 255u; };
		// east coupling of the west neighbour (I - 1, J): fine column 2I - 1
		const unsigned we0 = hw ? ((w0 >> 1) & 1u) + (hs ? (w1 >> 1) & 1u : 0u) : 0u;
		unsigned we = we0;
		float4 D, E, S;
		uint4 G;
#pragma unroll
		for (int k = 0; k < 4; k++) {
			const Coarse0 v = coarse0_of(byte(a, 2 * k), byte(a, 2 * k + 1), byte(c, 2 * k),
					byte(c, 2 * k + 1));
			const unsigned e = I + k + 1 < C.w ? v.e : 0u, so = J + 1 < C.h ? v.s : 0u;
			// south coupling of the north neighbour: fine row 2J - 1
			const unsigned no = ((byte(u, 2 * k) >> 2) & 1u) + ((byte(u, 2 * k + 1) >> 2) & 1u);
			const unsigned g = v.d | e << 11 | so << 16 | we << 21 | no << 26;
			(&D.x)[k] = (float)v.d;
			(&E.x)[k] = (float)e;
			(&S.x)[k] = (float)so;
			(&G.x)[k] = g;
			we = e;
		}
		gstore<f4>(C.D + q4, f4{D.x, D.y, D.z, D.w});
		gstore<f4>(C.wE + q4, f4{E.x, E.y, E.z, E.w});
		gstore<f4>(C.wS + q4, f4{S.x, S.y, S.z, S.w});
		gstore<f4>(C.z + q4, f4{0.0f, 0.0f, 0.0f, 0.0f});
		gstore<gu4>(C.gcode + q4, gu4{G.x, G.y, G.z, G.w});
		return;
	}
	const unsigned q1 = q4 + 4 < n ? q4 + 4 : n;
	int i = I, j = J;
	for (unsigned q = q4; q < q1; q++) {
		const Coarse0 v = coarse0_cell(F, i, j);
		const unsigned we = i > 0 ? coarse0_cell(F, i - 1, j).e : 0u;
		const unsigned no = j > 0 ? coarse0_cell(F, i, j - 1).s : 0u;
		coarse0_store(C, q, i, j, v, we, no);
		if (++i == C.w) { i = 0; j++; }
	}
}

// one red-black half sweep on colour `color`; `first` starts from z = 0
__global__ __launch_bounds__(NT) void k_mg_smooth(const BlockWork *bw,
		const MgLevel *lv, const Cg32Scal *sc, int color, int first, int l0)
{
	const BlockWork b = bw[blockIdx.x];
	if (!sc[b.plane].active) return;
	const MgLevel L = lv[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)L.w * L.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const int i = (int)(q % L.w), j = (int)(q / L.w);
		if (((i + j) & 1) != color) {
			if (first) L.z[q] = 0.0f;
			continue;
		}
		const float d = diag_at(L, i, j, l0);
		if (d == 0.0f) continue;
		const float s = first ? L.b[q] : L.b[q] + offdiag(L, i, j, l0);
		L.z[q] = s / d;
	}
}

// coarse right-hand side: sum of the fine residuals of the children
__global__ __launch_bounds__(NT) void k_mg_restrict(const BlockWork *bw,
		const MgLevel *fine, const MgLevel *coarse, const Cg32Scal *sc, int l0)
{
	const BlockWork b = bw[blockIdx.x];
	if (!sc[b.plane].active) return;
	const MgLevel F = fine[b.plane], C = coarse[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)C.w * C.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const int I = (int)(q % C.w), J = (int)(q / C.w);
		float r = 0;
		for (int c = 0; c < 4; c++) {
			const int i = 2 * I + (c & 1), j = 2 * J + (c >> 1);
			if (i >= F.w || j >= F.h) continue;
			const float d = diag_at(F, i, j, l0);
			if (d == 0.0f) continue;
			const size_t f = (size_t)j * F.w + i;
			r += F.b[f] - (d * F.z[f] - offdiag(F, i, j, l0));
		}
		C.b[q] = r;
		C.z[q] = 0.0f;
	}
}

// z += wc * P z_coarse on the active fine cells
__global__ __launch_bounds__(NT) void k_mg_prolong(const BlockWork *bw,
		const MgLevel *fine, const MgLevel *coarse, const Cg32Scal *sc, float wc,
		int l0)
{
	const BlockWork b = bw[blockIdx.x];
	if (!sc[b.plane].active) return;
	const MgLevel F = fine[b.plane], C = coarse[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)F.w * F.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const int i = (int)(q % F.w), j = (int)(q / F.w);
		if (diag_at(F, i, j, l0) == 0.0f) continue;
		F.z[q] += wc * C.z[(size_t)(j >> 1) * C.w + (i >> 1)];
	}
}

// ------------------------------------------------------ fused V-cycle ends
//
// Level 0 is the masked -Laplacian itself.  A byte per pixel encodes it:
// bit 0 unknown, bits 1-4 the E, S, W, N neighbour is an in-image unknown,
// bits 5-7 the number of in-image neighbours (the diagonal, 2..4).
// Levels 1..MG_G32_MAX pack their integer Galerkin coefficients into 32
// bits: the diagonal in bits 0-10, the E, S, W, N couplings in 5 bits each
// from bit 11 (a cell is active when its diagonal is not 0); the levels
// after them, to MG_FUSE_MAX, into two words: the couplings in a byte
// each, and the diagonal's float bits.
// The kernels below run on tiles of TX x TY cells (even origin) with a halo
// of H = 2*nu: every red-black half sweep shrinks the region where z is
// exact by one cell, and the first half sweep from z = 0 is exact
// everywhere, so after 2*nu half sweeps z is exact on the tile plus one
// cell (down: enough for the residual) or on the tile (up).

// Thread layout: a thread owns the 2x2 blocks (column pair pc, row pair m)
// for m = m0, m0+RG, ...; b and the codes of its cells stay in registers
// across half sweeps, and a block is also one coarse cell (origins even),
// so the restriction and the coarse correction are thread-local.
template <int TX_, int TY_, int NU_, int T_> struct MgGeo {
	static constexpr int TX = TX_, TY = TY_, NU = NU_, HALO = 2 * NU, T = T_;
	static constexpr int RW = TX + 2 * HALO, RH = TY + 2 * HALO;
	static constexpr int PAIRS_X = RW / 2, ROWPAIRS = RH / 2;
	static constexpr int RG = T / PAIRS_X, K = ROWPAIRS / RG;
	// z of the region in LDS with a zero frame.  Each row stores its even
	// columns, one frame cell, then its odd columns (ZS = RW + 1 per row),
	// so the cells a wave touches at once (one colour, consecutive column
	// pairs) are consecutive words: no bank conflicts.  The frame cell is
	// the west neighbour of column 0 and the east neighbour of column RW-1;
	// a frame row lies above and below the region.  Every neighbour read of
	// a region cell stays in bounds and reads 0 outside the region, with no
	// tests.
	static constexpr int HALF = PAIRS_X + 1;   // even column c/2 -> odd column c/2
	static constexpr int ZS = RW + 1, ZOFF = ZS, ZN = (RH + 2) * ZS;
	static_assert(TX % 2 == 0 && TY % 2 == 0, "tile origins must be even");
	static_assert(PAIRS_X * RG == T && RG * K == ROWPAIRS,
			"multigrid tile shape must divide the thread layout");
	static __device__ __forceinline__ int zid(int r, int col)
	{
		return ZOFF + r * ZS + (col & 1) * HALF + (col >> 1);
	}
	// offsets of the east / west neighbours of a cell in an even (par 0) or
	// odd (par 1) column
	static __device__ __forceinline__ int de(int par) { return par ? 1 - HALF : HALF; }
	static __device__ __forceinline__ int dw(int par) { return par ? -HALF : HALF - 1; }
};
template <int F>
using GeoF = MgGeo<MG_FINE_TX[F], MG_FINE_TY[F], MG_FINE_NU[F], MG_FINE_T[F]>;
template <int G>
using GeoC = MgGeo<MG_GEO_TX[G], MG_GEO_TY[G], MG_GEO_NU[G], MG_GEO_T[G]>;

// packed coefficients of a level 1..MG_FUSE_MAX (one thread per cell);
// wide: two words, the couplings E, S, W, N in a byte each and the
// diagonal's float bits
__global__ __launch_bounds__(NT) void k_mg_gcode(const BlockWork *bw,
		const MgLevel *lv, int wide)
{
	const BlockWork b = bw[blockIdx.x];
	const MgLevel L = lv[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)L.w * L.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const int i = (int)(q % L.w), j = (int)(q / L.w);
		const unsigned d = (unsigned)L.D[q];
		const unsigned e = (unsigned)L.wE[q], so = (unsigned)L.wS[q];
		const unsigned we = i > 0 ? (unsigned)L.wE[q - 1] : 0u;
		const unsigned no = j > 0 ? (unsigned)L.wS[q - L.w] : 0u;
		if (wide)
			((uint2 *)L.gcode)[q] = make_uint2(e | so << 8 | we << 16 | no << 24,
					__float_as_uint(L.D[q]));
		else
			L.gcode[q] = d | e << 11 | so << 16 | we << 21 | no << 26;
	}
}

// level 0: byte codes, four to a register per 2x2 block
struct Lev0 {
	static constexpr int CW = 1;   // code words per block
	static constexpr bool NEEDK = false;   // gather ignores the code
	typedef unsigned char Raw;     // one cell's code in memory
	typedef unsigned short Raw2;   // two cells' (first in the low byte)
	typedef unsigned Code;         // one cell's code in registers
	static __device__ __forceinline__ const Raw *codes(const MgLevel &L) { return L.code; }
	static __device__ __forceinline__ Code zero() { return 0u; }
	static __device__ __forceinline__ Code one(Raw r) { return r; }
	static __device__ __forceinline__ void split(Raw2 r, Code &a, Code &b)
	{ a = r & 255u; b = (unsigned)r >> 8; }
	static __device__ __forceinline__ void set(unsigned (&k)[CW], int c, Code x)
	{ k[0] = c ? k[0] | x << (8 * c) : x; }
	static __device__ __forceinline__ Code get(const unsigned (&k)[CW], int c)
	{ return (k[0] >> (8 * c)) & 255u; }
	static __device__ __forceinline__ Code fresh(Code k)
	{
		asm volatile("" : "+v"(k));
		return k;
	}
	static __device__ __forceinline__ float mask(unsigned k) { return (float)(k & 1u); }
	static __device__ __forceinline__ float diag(unsigned k) { return (float)(k >> 5); }
	// 1/diagonal, 0 for an inactive cell (code 0); selects, no branches
	static __device__ __forceinline__ float rdiag(unsigned k)
	{
		const unsigned d = k >> 5;
		float r = d == 2 ? 0.5f : 0.0f;
		r = d == 3 ? (1.0f / 3.0f) : r;
		return d == 4 ? 0.25f : r;
	}
	// s + the neighbours of the cell at z[id] (E, S, W, N); zs = row pitch,
	// de / dw the east / west offsets.  A neighbour that is not an in-image
	// unknown (known pixel, outside the image, the region's frame) holds
	// z = 0 at all times, so adding all four equals adding the unknown ones
	// (x + 0 = x): no masks needed.
	static __device__ __forceinline__ float gather4(unsigned, float s, float e,
			float so, float w, float n)
	{
		s += e;
		s += so;
		s += w;
		s += n;
		return s;
	}
	static __device__ __forceinline__ float gather(unsigned k, float s,
			const float *z, int id, int zs, int de, int dw)
	{
		return gather4(k, s, z[id + de], z[id + zs], z[id + dw], z[id - zs]);
	}
};

// levels 1..MG_FUSE_MAX: packed integer coefficients, one word per cell
struct LevG {
	static constexpr int CW = 4;
	static constexpr bool NEEDK = true;
	typedef unsigned Raw;
	typedef gu2 Raw2;
	typedef unsigned Code;
	static __device__ __forceinline__ const Raw *codes(const MgLevel &L) { return L.gcode; }
	static __device__ __forceinline__ Code zero() { return 0u; }
	static __device__ __forceinline__ Code one(Raw r) { return r; }
	static __device__ __forceinline__ void split(Raw2 r, Code &a, Code &b)
	{ a = r.x; b = r.y; }
	static __device__ __forceinline__ void set(unsigned (&k)[CW], int c, Code x)
	{ k[c] = x; }
	static __device__ __forceinline__ Code get(const unsigned (&k)[CW], int c)
	{ return k[c]; }
	// decode the coefficients again in every half sweep: hoisting the
	// decoded floats out of the sweep loop costs 80 registers per thread
	static __device__ __forceinline__ Code fresh(Code k)
	{
		asm volatile("" : "+v"(k));
		return k;
	}
	static __device__ __forceinline__ float mask(unsigned k)
	{ return (k & 2047u) ? 1.0f : 0.0f; }
	static __device__ __forceinline__ float diag(unsigned k) { return (float)(k & 2047u); }
	// 1/diagonal, 0 for an inactive cell (diagonal 0).  The hardware
	// reciprocal (1 ulp) is the same in the pre- and post-smoother, so the
	// V-cycle stays symmetric.
	static __device__ __forceinline__ float rdiag(unsigned k)
	{
		const unsigned d = k & 2047u;
		return d ? __builtin_amdgcn_rcpf((float)d) : 0.0f;
	}
	static __device__ __forceinline__ float gather4(unsigned k, float s, float e,
			float so, float w, float n)
	{
		s += (float)((k >> 11) & 31u) * e;
		s += (float)((k >> 16) & 31u) * so;
		s += (float)((k >> 21) & 31u) * w;
		s += (float)(k >> 26) * n;
		return s;
	}
	static __device__ __forceinline__ float gather(unsigned k, float s,
			const float *z, int id, int zs, int de, int dw)
	{
		return gather4(k, s, z[id + de], z[id + zs], z[id + dw], z[id - zs]);
	}
};

// levels MG_G32_MAX+1..MG_FUSE_MAX: couplings in a byte each, the diagonal
// as a float (two words per cell)
struct LevG64 {
	static constexpr int CW = 8;
	static constexpr bool NEEDK = true;
	typedef gu2 Raw;
	typedef gu4 Raw2;
	typedef uint2 Code;
	static __device__ __forceinline__ const Raw *codes(const MgLevel &L)
	{ return (const Raw *)L.gcode; }
	static __device__ __forceinline__ Code zero() { return make_uint2(0u, 0u); }
	static __device__ __forceinline__ Code one(Raw r) { return make_uint2(r.x, r.y); }
	static __device__ __forceinline__ void split(Raw2 r, Code &a, Code &b)
	{ a = make_uint2(r.x, r.y); b = make_uint2(r.z, r.w); }
	static __device__ __forceinline__ void set(unsigned (&k)[CW], int c, Code x)
	{ k[2 * c] = x.x; k[2 * c + 1] = x.y; }
	static __device__ __forceinline__ Code get(const unsigned (&k)[CW], int c)
	{ return make_uint2(k[2 * c], k[2 * c + 1]); }
	static __device__ __forceinline__ Code fresh(Code k)
	{
		asm volatile("" : "+v"(k.x), "+v"(k.y));
		return k;
	}
	static __device__ __forceinline__ float diag(Code k) { return __uint_as_float(k.y); }
	static __device__ __forceinline__ float mask(Code k) { return diag(k) != 0.0f ? 1.0f : 0.0f; }
	static __device__ __forceinline__ float rdiag(Code k)
	{
		const float d = diag(k);
		return d != 0.0f ? __builtin_amdgcn_rcpf(d) : 0.0f;
	}
	static __device__ __forceinline__ float gather4(Code k, float s, float e,
			float so, float w, float n)
	{
		s += (float)(k.x & 255u) * e;
		s += (float)((k.x >> 8) & 255u) * so;
		s += (float)((k.x >> 16) & 255u) * w;
		s += (float)(k.x >> 24) * n;
		return s;
	}
	static __device__ __forceinline__ float gather(Code k, float s,
			const float *z, int id, int zs, int de, int dw)
	{
		return gather4(k, s, z[id + de], z[id + zs], z[id + dw], z[id - zs]);
	}
};

// per-thread cells: K 2x2 blocks stacked in a 2 x 2K strip, block kb at
// block row m = m0 K + kb, cell c at (2m + (c>>1), 2pc + (c&1)).  With few
// cells per thread (RD), 1/diagonal and the thread's own z stay in
// registers: a half sweep then reads from LDS only the neighbours that
// other threads own (1.5 per cell instead of 3) and writes its new values
// to both.
template <class P, class G> struct MgCells {
	static constexpr bool RD = G::K <= 4;
	float b[G::K][4];
	float rd[RD ? G::K : 1][4];
	float zr[RD ? G::K : 1][4];
	unsigned k[G::K][P::CW];
	int pc, m0;
};

template <class G>
static __device__ __forceinline__ int brow(int m0, int kb) { return m0 * G::K + kb; }

template <class P, class G>
static __device__ __forceinline__ float cell_rdiag(const MgCells<P, G> &C, int kb,
		int c, typename P::Code k)
{
	if constexpr (MgCells<P, G>::RD) return C.rd[kb][c];
	else return P::rdiag(k);
}

template <class P, class G>
static __device__ __forceinline__ int cell_z(const MgCells<P, G> &C, int kb, int c)
{
	return G::zid(2 * brow<G>(C.m0, kb) + (c >> 1), 2 * C.pc + (c & 1));
}

// Region geometry of one thread: its columns i0, i0+1 and the rows
// j0 + 2*kb + (0, 1); q0 = j0 * w + i0 (may lie outside the image).
struct MgSpan {
	int i0, j0, w, h;
	long q0;
	bool colok[2];
};

template <class G>
static __device__ __forceinline__ MgSpan mg_span(const MgLevel &L, int ox, int oy,
		int pc, int m0)
{
	MgSpan S;
	S.i0 = ox + 2 * pc;
	S.j0 = oy + 2 * brow<G>(m0, 0);
	S.w = L.w;
	S.h = L.h;
	S.q0 = (long)S.j0 * L.w + S.i0;
	S.colok[0] = S.i0 >= 0 && S.i0 < L.w;
	S.colok[1] = S.i0 + 1 >= 0 && S.i0 + 1 < L.w;
	return S;
}

template <class G>
static __device__ __forceinline__ bool row_ok(const MgSpan &S, int kb, int hr)
{
	const int j = S.j0 + 2 * kb + hr;
	return j >= 0 && j < S.h;
}

template <class G>
static __device__ __forceinline__ long row_q(const MgSpan &S, int kb, int hr)
{
	return S.q0 + (long)(2 * kb + hr) * S.w;
}

// zero frame of the region's z: the rows above and below, and the frame
// cell between the even and the odd columns of every row
template <class G>
static __device__ __forceinline__ void z_frame(float *z)
{
	for (int e = threadIdx.x; e < 2 * G::ZS + G::RH; e += G::T) {
		const int at = e < G::ZS ? e : e < 2 * G::ZS ? G::ZOFF + G::RH * G::ZS + (e - G::ZS)
			: G::ZOFF + (e - 2 * G::ZS) * G::ZS + G::PAIRS_X;
		z[at] = 0.0f;
	}
}

// load b and codes (zero outside the image); z of the region from zsrc
// (LOADZ) or 0.  Every load of the thread goes out before the first is
// used: addresses are clamped into the image, and cells outside it take 0
// by a select afterwards.  With an even width the two cells of a block row
// are one aligned pair (region origins are even), loaded together.
template <bool LOADZ, class P, class G>
static __device__ __forceinline__ void mg_load(const MgLevel &L, int ox, int oy,
		MgCells<P, G> &C, float *z, const float *zsrc, const float *ce = nullptr,
		const bool *cin = nullptr, float wc = 0.0f)
{
	C.pc = threadIdx.x % G::PAIRS_X;
	C.m0 = threadIdx.x / G::PAIRS_X;
	const MgSpan S = mg_span<G>(L, ox, oy, C.pc, C.m0);
	const typename P::Raw *code = P::codes(L);
	float bv[G::K][4], zv[G::K][4];
	typename P::Code kv[G::K][4];
	if (!(L.w & 1)) {
		const int ic = clampi(S.i0, 0, L.w - 2);
		gf2 br[G::K][2], zr[G::K][2];
		typename P::Raw2 kr[G::K][2];
#pragma unroll
		for (int kb = 0; kb < G::K; kb++)
#pragma unroll
			for (int hr = 0; hr < 2; hr++) {
				const int j = clampi(S.j0 + 2 * kb + hr, 0, L.h - 1);
				const unsigned q = (unsigned)j * (unsigned)L.w + (unsigned)ic;
				br[kb][hr] = gload_b<gf2>(L.b, 4 * q);
				kr[kb][hr] = gload_b<typename P::Raw2>(code,
						q * (unsigned)sizeof(typename P::Raw));
				if (LOADZ) zr[kb][hr] = gload_b<gf2>(zsrc, 4 * q);
			}
#pragma unroll
		for (int kb = 0; kb < G::K; kb++)
#pragma unroll
			for (int hr = 0; hr < 2; hr++) {
				const bool ok = S.colok[0] && row_ok<G>(S, kb, hr);
				typename P::Code k0, k1;
				P::split(kr[kb][hr], k0, k1);
				bv[kb][2 * hr] = ok ? br[kb][hr].x : 0.0f;
				bv[kb][2 * hr + 1] = ok ? br[kb][hr].y : 0.0f;
				kv[kb][2 * hr] = ok ? k0 : P::zero();
				kv[kb][2 * hr + 1] = ok ? k1 : P::zero();
				zv[kb][2 * hr] = LOADZ && ok ? zr[kb][hr].x : 0.0f;
				zv[kb][2 * hr + 1] = LOADZ && ok ? zr[kb][hr].y : 0.0f;
			}
	} else {
		float br[G::K][4], zr[G::K][4];
		typename P::Raw kr[G::K][4];
#pragma unroll
		for (int kb = 0; kb < G::K; kb++)
#pragma unroll
			for (int c = 0; c < 4; c++) {
				const int j = clampi(S.j0 + 2 * kb + (c >> 1), 0, L.h - 1);
				const int i = clampi(S.i0 + (c & 1), 0, L.w - 1);
				const unsigned q = (unsigned)j * (unsigned)L.w + (unsigned)i;
				br[kb][c] = gload_b<float>(L.b, 4 * q);
				kr[kb][c] = gload_b<typename P::Raw>(code,
						q * (unsigned)sizeof(typename P::Raw));
				if (LOADZ) zr[kb][c] = gload_b<float>(zsrc, 4 * q);
			}
#pragma unroll
		for (int kb = 0; kb < G::K; kb++)
#pragma unroll
			for (int c = 0; c < 4; c++) {
				const bool ok = S.colok[c & 1] && row_ok<G>(S, kb, c >> 1);
				bv[kb][c] = ok ? br[kb][c] : 0.0f;
				kv[kb][c] = ok ? P::one(kr[kb][c]) : P::zero();
				zv[kb][c] = LOADZ && ok ? zr[kb][c] : 0.0f;
			}
	}
#pragma unroll
	for (int kb = 0; kb < G::K; kb++)
#pragma unroll
		for (int c = 0; c < 4; c++) {
			// with ce: a block (one coarse cell) inside the coarse level
			// adds its coarse correction wc * ce at its unknowns
			const float zc = ce && cin[kb] ? zv[kb][c] + wc * ce[kb] * P::mask(kv[kb][c])
				: zv[kb][c];
			C.b[kb][c] = bv[kb][c];
			P::set(C.k[kb], c, kv[kb][c]);
			if constexpr (MgCells<P, G>::RD) {
				C.rd[kb][c] = P::rdiag(kv[kb][c]);
				C.zr[kb][c] = zc;
			}
			z[cell_z(C, kb, c)] = zc;
		}
}

// store z of the thread's tile cells (inside the image) into dst
template <class P, class G>
static __device__ __forceinline__ void mg_store_z(float *dst, const MgLevel &L,
		int ox, int oy, const MgCells<P, G> &C, const float *z, int kb)
{
	const MgSpan S = mg_span<G>(L, ox, oy, C.pc, C.m0);
#pragma unroll
	for (int hr = 0; hr < 2; hr++) {
		if (!row_ok<G>(S, kb, hr)) continue;
		// pixel offsets of the pair's two cells (each stored only inside
		// the level, so each is >= 0 when used)
		const long q = row_q<G>(S, kb, hr);
		const unsigned o0 = 4u * (unsigned)q, o1 = 4u * (unsigned)(q + 1);
		const float z0 = z[cell_z(C, kb, 2 * hr)], z1 = z[cell_z(C, kb, 2 * hr + 1)];
		if (!(L.w & 1)) {
			if (S.colok[0]) gstore_b<gf2>(dst, o0, gf2{z0, z1});
		} else {
			if (S.colok[0]) gstore_b<float>(dst, o0, z0);
			if (S.colok[1]) gstore_b<float>(dst, o1, z1);
		}
	}
}

// red-black half sweep over the thread's cells of colour COLOR: all
// neighbour loads first (cells of one colour do not depend on each other).
// An inactive cell (known pixel, empty aggregate, outside the image) has
// b = 0, no couplings and 1/diagonal = 0, so it keeps z = 0 without a test.
// Neighbours are summed E, S, W, N whether they come from registers or LDS.
template <int COLOR, class P, class G>
static __device__ __forceinline__ void mg_half(MgCells<P, G> &C, float *z, bool first)
{
	constexpr bool RD = MgCells<P, G>::RD;
	float nz[G::K][2];
#pragma unroll
	for (int kb = 0; kb < G::K; kb++)
#pragma unroll
		for (int h = 0; h < 2; h++) {
			// colour of cell c is ((c>>1) + (c&1) + pair origin parity) & 1;
			// origins are even, so cells 0,3 have colour 0 and 1,2 colour 1
			const int c = COLOR == 0 ? (h ? 3 : 0) : (h ? 2 : 1);
			typename P::Code k = P::zero();
			if constexpr (P::NEEDK || !RD) k = P::fresh(P::get(C.k[kb], c));
			float s = C.b[kb][c];
			if (!first) {
				const int id = cell_z(C, kb, c);
				if constexpr (RD) {
					// own cells from registers: the block's other colour, and
					// the blocks above (kb - 1) and below (kb + 1) in the strip
					float e, so, w, n;
					if (c == 0) {
						e = C.zr[kb][1];
						so = C.zr[kb][2];
						w = z[id + G::dw(0)];
						n = kb > 0 ? C.zr[kb > 0 ? kb - 1 : 0][2] : z[id - G::ZS];
					} else if (c == 3) {
						e = z[id + G::de(1)];
						so = kb + 1 < G::K ? C.zr[kb + 1 < G::K ? kb + 1 : kb][1]
							: z[id + G::ZS];
						w = C.zr[kb][2];
						n = C.zr[kb][1];
					} else if (c == 1) {
						e = z[id + G::de(1)];
						so = C.zr[kb][3];
						w = C.zr[kb][0];
						n = kb > 0 ? C.zr[kb > 0 ? kb - 1 : 0][3] : z[id - G::ZS];
					} else {
						e = C.zr[kb][3];
						so = kb + 1 < G::K ? C.zr[kb + 1 < G::K ? kb + 1 : kb][0]
							: z[id + G::ZS];
						w = z[id + G::dw(0)];
						n = C.zr[kb][0];
					}
					s = P::gather4(k, s, e, so, w, n);
				} else {
					s = P::gather(k, s, z, id, G::ZS, G::de(c & 1), G::dw(c & 1));
				}
			}
			nz[kb][h] = s * cell_rdiag(C, kb, c, k);
		}
#pragma unroll
	for (int kb = 0; kb < G::K; kb++)
#pragma unroll
		for (int h = 0; h < 2; h++) {
			const int c = COLOR == 0 ? (h ? 3 : 0) : (h ? 2 : 1);
			if constexpr (RD) C.zr[kb][c] = nz[kb][h];
			z[cell_z(C, kb, c)] = nz[kb][h];
		}
}

template <class P, class G>
static __device__ __forceinline__ bool in_tile(const MgCells<P, G> &C, int kb)
{
	const int m = brow<G>(C.m0, kb);
	return C.pc >= G::HALO / 2 && C.pc < G::HALO / 2 + G::TX / 2
		&& m >= G::HALO / 2 && m < G::HALO / 2 + G::TY / 2;
}

// pre-smoothing from z = 0 (into zd), residual, 2x2 restriction into the
// next level
// (two 1024-thread work-groups per CU need at most 64 registers a thread)
template <class P, class G>
__global__ __launch_bounds__(G::T) __attribute__((amdgpu_waves_per_eu(G::T == 1024 ? 8 : 4)))
void k_mg_down(const int4 *tiles,
		const MgLevel *lv, const MgLevel *coarse, const Cg32Scal *sc)
{
	__shared__ float z[G::ZN];
	const int4 t = tiles[xcd_block()];
	// the descriptors load alongside the activity flag
	const MgLevel L = lv[t.x], Cc = coarse[t.x];
	if (!sc[t.x].active) return;
	const int ox = t.y - G::HALO, oy = t.z - G::HALO;
	MgCells<P, G> C;
	z_frame<G>(z);
	mg_load<false>(L, ox, oy, C, z, nullptr);
	__syncthreads();
	for (int s = 0; s < G::NU; s++) {
		mg_half<0>(C, z, s == 0);
		__syncthreads();
		mg_half<1>(C, z, false);
		__syncthreads();
	}
#pragma unroll
	for (int kb = 0; kb < G::K; kb++) {
		if (!in_tile(C, kb)) continue;
		mg_store_z(L.zd, L, ox, oy, C, z, kb);
		float acc = 0.0f;
#pragma unroll
		for (int c = 0; c < 4; c++) {
			const typename P::Code k = P::get(C.k[kb], c);
			const int id = cell_z(C, kb, c);
			// an inactive cell adds nothing (its unmasked neighbour sum need
			// not be 0)
			const float az = P::diag(k) * z[id] -
				P::gather(k, 0.0f, z, id, G::ZS, G::de(c & 1), G::dw(c & 1));
			acc += P::mask(k) * (C.b[kb][c] - az);
		}
		const int I = (ox >> 1) + C.pc, J = (oy >> 1) + brow<G>(C.m0, kb);
		if (I < Cc.w && J < Cc.h) {
			const size_t cq = (size_t)J * Cc.w + I;
			gstore<float>(Cc.b + cq, acc);
			gstore<float>(Cc.z + cq, 0.0f);
		}
	}
}

// coarse correction, post-smoothing (zd in, z out: a tile reads its halo
// from other tiles' cells, so it must not read the buffer that tiles of the
// same launch write); at level 0 also the partial r.z of
// the tile
// (two 1024-thread work-groups per CU need at most 64 registers a thread)
template <class P, class G, bool DOT>
__global__ __launch_bounds__(G::T) __attribute__((amdgpu_waves_per_eu(G::T == 1024 ? 8 : 4)))
void k_mg_up(const int4 *tiles, const MgLevel *lv,
		const MgLevel *coarse, const Cg32Scal *sc, float wc, double *tpart)
{
	__shared__ float z[G::ZN];
	__shared__ double red[G::T / 64];
	const int tb = xcd_block();
	const int4 t = tiles[tb];
	const MgLevel L = lv[t.x], Cc = coarse[t.x];
	if (!sc[t.x].active) return;
	const int ox = t.y - G::HALO, oy = t.z - G::HALO;
	// every block of the region is one coarse cell (levels are at least
	// 1 x 1): its correction loads with the region
	float ce[G::K];
	bool cin[G::K];
#pragma unroll
	for (int kb = 0; kb < G::K; kb++) {
		const int I = (ox >> 1) + (int)threadIdx.x % G::PAIRS_X;
		const int J = (oy >> 1) + brow<G>((int)threadIdx.x / G::PAIRS_X, kb);
		cin[kb] = I >= 0 && J >= 0 && I < Cc.w && J < Cc.h;
		ce[kb] = gload<float>(Cc.z + (size_t)clampi(J, 0, Cc.h - 1) * Cc.w
				+ clampi(I, 0, Cc.w - 1));
	}
	MgCells<P, G> C;
	z_frame<G>(z);
	mg_load<true>(L, ox, oy, C, z, L.zd, ce, cin, wc);
	__syncthreads();
	for (int s = 0; s < G::NU; s++) {
		mg_half<1>(C, z, false);
		__syncthreads();
		mg_half<0>(C, z, false);
		__syncthreads();
	}
	double acc = 0;
#pragma unroll
	for (int kb = 0; kb < G::K; kb++) {
		if (!in_tile(C, kb)) continue;
		mg_store_z(L.z, L, ox, oy, C, z, kb);
		if (DOT) {
#pragma unroll
			for (int c = 0; c < 4; c++)   // 0 at inactive cells
				acc += (double)C.b[kb][c] * z[cell_z(C, kb, c)];
		}
	}
	if (DOT) {
		const double sum = bsum<G::T>(acc, red);
		if (threadIdx.x == 0) tpart[tb] = sum;
	}
}

// rho_new = r.z folded over the plane's level-0 tiles
__global__ __launch_bounds__(RT) void k_mg_reduce_t(const Cg32Plane *pl,
		Cg32Scal *sc, const double *tpart, int setup)
{
	__shared__ double red[RT / 64];
	if (!sc[blockIdx.x].active) return;
	const Cg32Plane P = pl[blockIdx.x];
	double v = 0;
	for (int k = P.t0 + threadIdx.x; k < P.t0 + P.nt; k += RT) v += tpart[k];
	const double rz = bsum<RT>(v, red);
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		s.beta = setup ? 0.0 : rz / s.rz;
		s.rz = rz;
	}
}

void cg32_setup(const Cg32Batch &B, hipStream_t s);


// ------------------------------------------------------ coarse-level tail
//
// The small levels of a plane (from the first level of at most
// MG_TAIL_CELLS cells down to the coarsest) run inside one work-group: the
// same smoothing, restriction and correction as the per-level kernels, with
// a barrier between phases instead of a launch.  A thread's cells q =
// threadIdx.x + k * blockDim.x are walked with their (i, j) advanced by a
// constant step (no division per cell), and the smoother multiplies by the
// hardware reciprocal of the diagonal (1 ulp, the same in the pre- and
// post-smoother, so the V-cycle stays symmetric).
struct TailIdx { int i, j, di, dj; };

static __device__ __forceinline__ TailIdx tail_idx(int w)
{
	TailIdx t;
	t.j = (int)threadIdx.x / w;
	t.i = (int)threadIdx.x - t.j * w;
	t.dj = (int)blockDim.x / w;
	t.di = (int)blockDim.x - t.dj * w;
	return t;
}

static __device__ __forceinline__ void tail_next(TailIdx &t, int w)
{
	t.i += t.di;
	t.j += t.dj;
	if (t.i >= w) { t.i -= w; t.j++; }
}

static __device__ void tail_smooth(const MgLevel &L, int color, bool first)
{
	const int n = L.w * L.h;
	TailIdx p = tail_idx(L.w);
	for (int q = threadIdx.x; q < n; q += blockDim.x, tail_next(p, L.w)) {
		const int i = p.i, j = p.j;
		if (((i + j) & 1) != color) {
			if (first) L.z[q] = 0.0f;
			continue;
		}
		const float d = L.D[q];
		if (d == 0.0f) continue;
		L.z[q] = (first ? L.b[q] : L.b[q] + offdiag(L, i, j, false)) *
			__builtin_amdgcn_rcpf(d);
	}
	__syncthreads();
}

static __device__ void tail_restrict(const MgLevel &F, const MgLevel &C)
{
	const int n = C.w * C.h;
	TailIdx p = tail_idx(C.w);
	for (int q = threadIdx.x; q < n; q += blockDim.x, tail_next(p, C.w)) {
		const int I = p.i, J = p.j;
		float r = 0;
		for (int c = 0; c < 4; c++) {
			const int i = 2 * I + (c & 1), j = 2 * J + (c >> 1);
			if (i >= F.w || j >= F.h) continue;
			const size_t f = (size_t)j * F.w + i;
			const float d = F.D[f];
			if (d == 0.0f) continue;
			r += F.b[f] - (d * F.z[f] - offdiag(F, i, j, false));
		}
		C.b[q] = r;
		C.z[q] = 0.0f;
	}
	__syncthreads();
}

static __device__ void tail_prolong(const MgLevel &F, const MgLevel &C, float wc)
{
	const int n = F.w * F.h;
	TailIdx p = tail_idx(F.w);
	for (int q = threadIdx.x; q < n; q += blockDim.x, tail_next(p, F.w)) {
		if (F.D[q] == 0.0f) continue;
		F.z[q] += wc * C.z[(p.j >> 1) * C.w + (p.i >> 1)];
	}
	__syncthreads();
}

__global__ __launch_bounds__(MG_TAIL_T) void k_mg_tail(const MgLevel *lvall,
		const int4 *rng, const Cg32Scal *sc, int nu, int csw, float wc)
{
	const int z = blockIdx.x;
	if (!sc[z].active || rng[z].z) return;   // planes of the LDS tail skip
	const int l0 = rng[z].x, d = rng[z].y;
	const MgLevel *lv = lvall + (size_t)z * MG_MAXLEV;
	for (int l = l0; l + 1 < d; l++) {
		for (int it = 0; it < nu; it++) {
			tail_smooth(lv[l], 0, it == 0);
			tail_smooth(lv[l], 1, false);
		}
		tail_restrict(lv[l], lv[l + 1]);
	}
	for (int it = 0; it < csw; it++) {
		tail_smooth(lv[d - 1], 0, it == 0);
		tail_smooth(lv[d - 1], 1, false);
	}
	for (int it = 0; it < csw; it++) {
		tail_smooth(lv[d - 1], 1, false);
		tail_smooth(lv[d - 1], 0, false);
	}
	for (int l = d - 2; l >= l0; l--) {
		tail_prolong(lv[l], lv[l + 1], wc);
		for (int it = 0; it < nu; it++) {
			tail_smooth(lv[l], 1, false);
			tail_smooth(lv[l], 0, false);
		}
	}
}

// The same tail with every level's z, b, D, wE, wS in LDS (planes whose tail
// fits MG_TAIL_LDS_MAX).  Each level is stored with a frame of zero cells
// (width w + 2, height h + 2): the frame holds z = 0, b = 0, D = 0 and no
// couplings, the last column has wE = 0 and the last row wS = 0, so every
// neighbour term is read unconditionally and a cell's loads go out
// together (a frame or border term adds 0).  Inactive cells (D = 0) get
// z = s * 0 = 0 (s is 0 there: no right-hand side, no couplings).
struct TailLv {
	float *z, *b, *D, *wE, *wS;
	int w, h, W;   // W = w + 2
};

// f(i, j) on this thread's cells of a w x h level.  Levels up to 1024 wide
// map threads by shifts: thread t takes column t mod VW and rows t / VW +
// k (T / VW), VW the power of two at or above w; wider ones step (i, j).
template <class F>
static __device__ __forceinline__ void tail_for(int w, int h, F &&f)
{
	if (w <= 1024) {
		const int sh = w > 1 ? 32 - __clz(w - 1) : 0;
		const int i = (int)threadIdx.x & ((1 << sh) - 1);
		if (i >= w) return;
		const int dj = (int)blockDim.x >> sh;
		for (int j = (int)threadIdx.x >> sh; j < h; j += dj) f(i, j);
		return;
	}
	TailIdx t = tail_idx(w);
	for (int q = threadIdx.x; q < w * h; q += blockDim.x, tail_next(t, w)) f(t.i, t.j);
}

static __device__ __forceinline__ int tl_at(const TailLv &L, int i, int j)
{
	return (j + 1) * L.W + i + 1;
}

static __device__ __forceinline__ float tl_off(const TailLv &L, int p)
{
	float s = 0.0f;
	s += L.wE[p] * L.z[p + 1];
	s += L.wE[p - 1] * L.z[p - 1];
	s += L.wS[p] * L.z[p + L.W];
	s += L.wS[p - L.W] * L.z[p - L.W];
	return s;
}

static __device__ __forceinline__ float tl_rdiag(float d)
{
	return d != 0.0f ? __builtin_amdgcn_rcpf(d) : 0.0f;
}

static __device__ void tl_smooth(const TailLv &L, int color, bool first)
{
	tail_for(L.w, L.h, [&](int i, int j) {
		const int p = tl_at(L, i, j);
		if (((i + j) & 1) != color) {
			if (first) L.z[p] = 0.0f;
			return;
		}
		const float s = first ? L.b[p] : L.b[p] + tl_off(L, p);
		L.z[p] = s * tl_rdiag(L.D[p]);
	});
	__syncthreads();
}

static __device__ void tl_restrict(const TailLv &F, const TailLv &C)
{
	tail_for(C.w, C.h, [&](int i, int j) {
		float r = 0;
#pragma unroll
		for (int c = 0; c < 4; c++) {
			// fine cells past an odd edge are frame cells: D = 0
			const int f = tl_at(F, 2 * i + (c & 1), 2 * j + (c >> 1));
			const float d = F.D[f];
			const float v = F.b[f] - (d * F.z[f] - tl_off(F, f));
			r += d != 0.0f ? v : 0.0f;
		}
		const int p = tl_at(C, i, j);
		C.b[p] = r;
		C.z[p] = 0.0f;
	});
	__syncthreads();
}

static __device__ void tl_prolong(const TailLv &F, const TailLv &C, float wc)
{
	tail_for(F.w, F.h, [&](int i, int j) {
		const int p = tl_at(F, i, j);
		const float z = F.z[p], e = wc * C.z[tl_at(C, i >> 1, j >> 1)];
		F.z[p] = F.D[p] != 0.0f ? z + e : z;
	});
	__syncthreads();
}

__global__ __launch_bounds__(MG_TAIL_T) void k_mg_tail_lds(const MgLevel *lvall,
		const int4 *rng, const Cg32Scal *sc, int nu, int csw, float wc)
{
	extern __shared__ float tsm[];
	__shared__ int toff[MG_MAXLEV + 1], tw[MG_MAXLEV], th[MG_MAXLEV];
	const int z = blockIdx.x;
	if (!sc[z].active || !rng[z].z) return;
	const int l0 = rng[z].x, d = rng[z].y;
	const MgLevel *lv = lvall + (size_t)z * MG_MAXLEV;
	if (threadIdx.x == 0) {
		int o = 0;
		for (int l = l0; l < d; l++) {
			toff[l] = o;
			tw[l] = lv[l].w;
			th[l] = lv[l].h;
			o += 5 * (lv[l].w + 2) * (lv[l].h + 2);
		}
		toff[d] = o;
	}
	__syncthreads();
	// level l's framed arrays (z, b, D, wE, wS) from the offsets
	auto T = [&](int l) {
		const int n = (tw[l] + 2) * (th[l] + 2);
		float *p = tsm + toff[l];
		return TailLv{p, p + n, p + 2 * n, p + 3 * n, p + 4 * n, tw[l], th[l], tw[l] + 2};
	};
	// zero everything (frames included), then the interiors
	for (int q = threadIdx.x; q < toff[d]; q += blockDim.x) tsm[q] = 0.0f;
	__syncthreads();
	for (int l = l0; l < d; l++) {
		const TailLv L = T(l);
		const int n = L.w * L.h;
		TailIdx t = tail_idx(L.w);
		for (int q = threadIdx.x; q < n; q += blockDim.x, tail_next(t, L.w)) {
			const int p = tl_at(L, t.i, t.j);
			L.D[p] = lv[l].D[q];
			L.wE[p] = lv[l].wE[q];
			L.wS[p] = lv[l].wS[q];
			if (l == l0) L.b[p] = lv[l].b[q];
		}
	}
	__syncthreads();
	for (int l = l0; l + 1 < d; l++) {
		for (int it = 0; it < nu; it++) {
			tl_smooth(T(l), 0, it == 0);
			tl_smooth(T(l), 1, false);
		}
		tl_restrict(T(l), T(l + 1));
	}
	for (int it = 0; it < csw; it++) {
		tl_smooth(T(d - 1), 0, it == 0);
		tl_smooth(T(d - 1), 1, false);
	}
	for (int it = 0; it < csw; it++) {
		tl_smooth(T(d - 1), 1, false);
		tl_smooth(T(d - 1), 0, false);
	}
	for (int l = d - 2; l >= l0; l--) {
		tl_prolong(T(l), T(l + 1), wc);
		for (int it = 0; it < nu; it++) {
			tl_smooth(T(l), 1, false);
			tl_smooth(T(l), 0, false);
		}
	}
	// the first tail level's correction is read by the next finer level
	const TailLv L0 = T(l0);
	const int n0 = L0.w * L0.h;
	TailIdx t = tail_idx(L0.w);
	for (int q = threadIdx.x; q < n0; q += blockDim.x, tail_next(t, L0.w))
		lv[l0].z[q] = L0.z[tl_at(L0, t.i, t.j)];
}

static const char *const g_down_name[MG_FUSE_MAX + 1] = {
	"mg_down0", "mg_down1", "mg_down2", "mg_down3", "mg_down4", "mg_down5",
	"mg_down6", "mg_down7"};
static const char *const g_up_name[MG_FUSE_MAX + 1] = {
	"mg_up0", "mg_up1", "mg_up2", "mg_up3", "mg_up4", "mg_up5", "mg_up6", "mg_up7"};

// call f with a value of the tile shape type Geo<geo>, geo in [I, N)
template <template <int> class Geo, int I, int N, class F>
static void with_geo_from(int geo, F &f)
{
	if constexpr (I + 1 < N) {
		if (geo != I) { with_geo_from<Geo, I + 1, N>(geo, f); return; }
	}
	f(Geo<I>{});
}

template <template <int> class Geo, int N, class F>
static void with_geo(int geo, F &&f)
{
	with_geo_from<Geo, 0, N>(geo, f);
}

static void vcycle(const Cg32Batch &B, hipStream_t s)
{
	const int L = B.nlev;
	// level 0 (never in the tail): fused pre-smoothing + restriction
	// algorithmic bytes per level-0 pixel: down reads b and the code, writes
	// zd and a quarter coarse b (10 B); up reads zd, b, the code and a
	// quarter coarse z, writes z (14 B)
	if (B.prof) B.prof("mg_down0", 1, 0, s);
	with_geo<GeoF, MG_NFINE>(B.geo0, [&](auto g) {
		using G = decltype(g);
		k_mg_down<Lev0, G><<<B.ntiles0, G::T, 0, s>>>(B.tiles0, B.mg[0].lv,
				B.mg[0].coarse, B.scal);
	});
	if (B.prof) B.prof("mg_down0", 0, 10.0 * B.px0, s);
	// levels >= 1 outside the tail: fused tiles (packed coefficients; per
	// cell down reads b and the code, writes zd and a quarter coarse b and
	// z, 14 B; up reads zd, b, the code and a quarter coarse z, writes z,
	// 17 B; 4 B more per cell for the two-word codes) or per-level kernels
	// (pre-smooth, restrict)
	for (int l = 1; l < L; l++) {
		const MgLevelWork &W = B.mg[l];
		if (W.ntiles) {
			if (B.prof) B.prof(g_down_name[l], 1, 0, s);
			auto down = [&](auto p, auto g) {
				using P = decltype(p);
				using G = decltype(g);
				k_mg_down<P, G><<<W.ntiles, G::T, 0, s>>>(W.tiles, W.lv, W.coarse, B.scal);
			};
			if (l == 1)
				with_geo<GeoF, MG_NFINE>(W.geo, [&](auto g) { down(LevG{}, g); });
			else if (l <= MG_G32_MAX)
				with_geo<GeoC, MG_NGEO>(W.geo, [&](auto g) { down(LevG{}, g); });
			else
				with_geo<GeoC, MG_NGEO>(W.geo, [&](auto g) { down(LevG64{}, g); });
			if (B.prof)
				B.prof(g_down_name[l], 0, (l > MG_G32_MAX ? 18.0 : 14.0) * W.cells, s);
		}
		if (!W.nmid) continue;
		for (int it = 0; it < B.nu; it++) {
			k_mg_smooth<<<W.nmid, NT, 0, s>>>(W.mid, W.lv, B.scal, 0, it == 0, 0);
			k_mg_smooth<<<W.nmid, NT, 0, s>>>(W.mid, W.lv, B.scal, 1, 0, 0);
		}
		k_mg_restrict<<<W.ncells_c, NT, 0, s>>>(W.cells_c, W.lv, W.coarse, B.scal, 0);
	}
	// small levels of every plane in one work-group each
	if (B.prof) B.prof("mg_tail", 1, 0, s);
	if (B.tail_global)
		k_mg_tail<<<B.nplanes, MG_TAIL_T, 0, s>>>(B.tail_lv, B.tail_rng, B.scal, B.nu,
				B.coarse_sweeps, B.wc);
	if (B.tail_lds)
		k_mg_tail_lds<<<B.nplanes, MG_TAIL_T, B.tail_lds, s>>>(B.tail_lv, B.tail_rng,
				B.scal, B.nu, B.coarse_sweeps, B.wc);
	if (B.prof) B.prof("mg_tail", 0, 0.0, s);
	// up: correct, post-smooth in the mirrored order
	for (int l = L - 1; l >= 1; l--) {
		const MgLevelWork &W = B.mg[l];
		if (W.nmid) {
			k_mg_prolong<<<W.nmid, NT, 0, s>>>(W.mid, W.lv, W.coarse, B.scal, B.wc, 0);
			for (int it = 0; it < B.nu; it++) {
				k_mg_smooth<<<W.nmid, NT, 0, s>>>(W.mid, W.lv, B.scal, 1, 0, 0);
				k_mg_smooth<<<W.nmid, NT, 0, s>>>(W.mid, W.lv, B.scal, 0, 0, 0);
			}
		}
		if (W.ntiles) {
			if (B.prof) B.prof(g_up_name[l], 1, 0, s);
			auto up = [&](auto p, auto g) {
				using P = decltype(p);
				using G = decltype(g);
				k_mg_up<P, G, false><<<W.ntiles, G::T, 0, s>>>(W.tiles, W.lv, W.coarse,
						B.scal, B.wc, nullptr);
			};
			if (l == 1)
				with_geo<GeoF, MG_NFINE>(W.geo, [&](auto g) { up(LevG{}, g); });
			else if (l <= MG_G32_MAX)
				with_geo<GeoC, MG_NGEO>(W.geo, [&](auto g) { up(LevG{}, g); });
			else
				with_geo<GeoC, MG_NGEO>(W.geo, [&](auto g) { up(LevG64{}, g); });
			if (B.prof)
				B.prof(g_up_name[l], 0, (l > MG_G32_MAX ? 21.0 : 17.0) * W.cells, s);
		}
	}
	if (B.prof) B.prof("mg_up0", 1, 0, s);
	with_geo<GeoF, MG_NFINE>(B.geo0, [&](auto g) {
		using G = decltype(g);
		k_mg_up<Lev0, G, true><<<B.ntiles0, G::T, 0, s>>>(B.tiles0, B.mg[0].lv,
				B.mg[0].coarse, B.scal, B.wc, B.tpart);
	});
	if (B.prof) B.prof("mg_up0", 0, 14.0 * B.px0, s);
}

// z = M r, rho = r.z (and p = z at setup on the separate-kernel path)
static void precondition(const Cg32Batch &B, int setup, hipStream_t s)
{
	vcycle(B, s);
	k_mg_reduce_t<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.tpart, setup);
	if (setup && B.bilap)   // p = z + 0 * p
		k32_dir<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.nblk);
}

int mg_tail_lds_enable(int bytes)
{
	return hipFuncSetAttribute((const void *)k_mg_tail_lds,
			hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess ? 0 : -1;
}

void mg_build(const Cg32Batch &B, hipStream_t s)
{
	// level 1 with its packed words in one pass from the code bytes
	if (B.nlev > 1 && B.mg[0].nbuild_c)
		k_mg_coarsen0<<<B.mg[0].nbuild_c, NT, 0, s>>>(B.mg[0].build_c, B.mg[0].lv,
				B.mg[0].coarse);
	for (int l = 1; l + 1 < B.nlev; l++) {
		const MgLevelWork &W = B.mg[l];
		if (W.nbuild_c)
			k_mg_coarsen<<<W.nbuild_c, NT, 0, s>>>(W.build_c, W.lv, W.coarse, 0);
	}
	// packed coefficients of the fused levels from 2 (their couplings are final)
	for (int l = 1; l < MG_FUSE_MAX && l + 1 < B.nlev; l++) {
		const MgLevelWork &W = B.mg[l];
		if (W.nbuild_c && B.mg[l + 1].ntiles)
			k_mg_gcode<<<W.nbuild_c, NT, 0, s>>>(W.build_c, W.coarse,
					l + 1 > MG_G32_MAX);
	}
}

void cg32_setup(const Cg32Batch &B, hipStream_t s)
{
	k32_setup<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.part);
	k32_reduce0<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.part, B.nactive);
	if (B.nlev) {
		mg_build(B, s);
		precondition(B, 1, s);
	}
	// p = z, q = A p, p.q into the buffer iteration 0 reads (par 1 -> p)
	if (!B.bilap)
		k32_dirspmv<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part, 1, 1, B.nblk);
}

// Poisson: alpha, r -= alpha q, beta, z = M r, then the fused direction +
// product (x += alpha p_old, p_new, q = A p_new, p.q) reading p_old from the
// buffer of parity par.  Biharmonic: product, alpha, update, beta, z,
// direction (one p buffer).
void cg32_iter(const Cg32Batch &B, hipStream_t s, int par)
{
	const int pc = B.nlev > 0;
	if (B.bilap) {
		k32_spmv<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part, B.nblk);
		k32_reduce_a<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.part, B.nactive);
		k32_update<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part, B.nblk);
		k32_reduce_b<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.part, B.nactive,
				B.trace, pc);
		if (pc) precondition(B, 0, s);
		k32_dir<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.nblk);
		return;
	}
	k32_reduce_a<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.part, B.nactive);
	k32_update<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part, B.nblk);
	k32_reduce_b<<<B.nplanes, RT, 0, s>>>(B.planes, B.scal, B.part, B.nactive,
			B.trace, pc);
	if (pc) precondition(B, 0, s);
	// reads z, the code and p_old (with the four neighbours), writes p_new
	// and q; every other pass also x and the p of two passes before
	if (B.prof) B.prof("k32_dirspmv", 1, 0, s);
	k32_dirspmv<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part, par, 0, B.nblk);
	if (B.prof) B.prof("k32_dirspmv", 0, 0.0, s);
}

// the Poisson iteration writes the solution in the pass that stops a plane
void cg32_finish(const Cg32Batch &B, hipStream_t s)
{
	if (B.bilap) k32_finish<<<B.nblk, NT, 0, s>>>(B.bw, B.planes);
}

// Gradient of the converged biharmonic fill from lambda = A_hh^-T v (the
// solve's output, 0 at known pixels): dL/df = -lambda on the holes, and on
// the known pixels dL/dg = v - (A_hk^T lambda), the transposed operator
// over lambda there.  Work blocks of 1024 pixels, 4 per thread.
__global__ __launch_bounds__(NT) void k_adj_gather_bilap(const AdjPlane *ap,
		const BlockWork *bw)
{
	const BlockWork b = bw[blockIdx.x];
	const AdjPlane A = ap[b.plane];
	const size_t n = (size_t)A.w * A.h;
	for (int k = 0; k < PPT; k++) {
		const size_t q = (size_t)b.start + k * NT + threadIdx.x;
		if (q >= n) break;
		if (isnan(A.g[q])) {
			A.grad_g[q] = 0.0f;
			if (A.grad_f) A.grad_f[q] = -A.lam[q];
			continue;
		}
		const int i = (int)(q % A.w), j = (int)(q / A.w);
		const float t = op32T(i, j, A.w, A.h, [&](size_t z) {
			return isnan(A.g[z]) ? A.lam[z] : 0.0f; });
		A.grad_g[q] = A.v[q] - t;
		if (A.grad_f) A.grad_f[q] = 0.0f;
	}
}

void launch_adj_gather_bilap(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s)
{
	if (nblocks) k_adj_gather_bilap<<<nblocks, NT, 0, s>>>(ap, bw);
}

} // namespace spk
