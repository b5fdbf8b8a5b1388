// Device kernels for the simpois hole-filling path on gfx950 (CDNA4).
//
// All arithmetic that the reference performs in float is performed here in
// float with the same association order; the library is compiled with
// -ffp-contract=off so no multiply-add is fused.  That, plus the hyperplane
// schedule below, makes the Gauss-Seidel sweeps and the pyramid bit-exact
// with simpois.c.
//
// The Gauss-Seidel schedule and its correctness argument are in sp_gs.h.
#include "sp_internal.h"
#include <math.h>
#include <algorithm>
#include <type_traits>

namespace spk {

static __device__ __forceinline__ float coef_of(float g, const float *f, size_t q)
{
	return sp_coef(g, f && isnan(g) ? f[q] : 0.0f);
}

static __device__ __forceinline__ int clampi(int v, int lo, int hi)
{
	return v < lo ? lo : (v > hi ? hi : v);
}

static __device__ __forceinline__ float at(const float *a, int w, int h, int i, int j)
{
	return a[(size_t)clampi(j, 0, h - 1) * w + clampi(i, 0, w - 1)];
}

// block-wide sum of an int, result valid in thread 0
template <int NT>
static __device__ int block_sum_int(int v, int *red)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
	int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) red[wv] = v;
	__syncthreads();
	int r = 0;
	if (threadIdx.x == 0)
		for (int k = 0; k < NT / 64; k++) r += red[k];
	__syncthreads();
	return r;
}

// A double moved across lanes as its two halves: lane i takes lane i + 32
// (lanes 0-31), lane i + 16 (lanes 0-15 and 32-47: the swapped halves of
// permlane32 / permlane16 swaps), or lane i + o within its row of 16 (a
// DPP row shift left by o = 8, 4, 2, 1: CTRL 0x100 + o; lanes whose source
// leaves the row keep their value)
static __device__ __forceinline__ double dbl_of(unsigned lo, unsigned hi)
{
	return __builtin_bit_cast(double, (unsigned long long)hi << 32 | lo);
}
static __device__ __forceinline__ double lanes_up32(double x)
{
	const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
	const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
	const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
			false, false);
	return dbl_of(lo[1], hi[1]);
}
static __device__ __forceinline__ double lanes_up16(double x)
{
	const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
	const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
	const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
			false, false);
	return dbl_of(lo[1], hi[1]);
}
template <int CTRL>
static __device__ __forceinline__ double lanes_up_row(double x)
{
	const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
	const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
	return dbl_of(__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false),
			__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false));
}

// wave sum in lane 0 by the tree v += (v of lane + o), o = 32, 16, ..., 1
// (that of __shfl_down, value for value: lane 0's result reads at step o
// only lanes below 2 o, which the moves above serve, with no trip through
// the LDS crossbar)
static __device__ __forceinline__ double wave_sum_dbl(double v)
{
	v = v + lanes_up32(v);
	v = v + lanes_up16(v);
	v = v + lanes_up_row<0x108>(v);
	v = v + lanes_up_row<0x104>(v);
	v = v + lanes_up_row<0x102>(v);
	v = v + lanes_up_row<0x101>(v);
	return v;
}

// block-wide sum of a double (fixed tree), result valid in thread 0
template <int NT>
static __device__ double block_sum_dbl(double v, double *red)
{
	v = wave_sum_dbl(v);
	int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) red[wv] = v;
	__syncthreads();
	double r = 0;
	if (threadIdx.x == 0)
		for (int k = 0; k < NT / 64; k++) r += red[k];
	__syncthreads();
	return r;
}

constexpr int EW_NT = 256;   // threads per elementwise block
constexpr int EW_PPT = 4;    // pixels per thread

// ------------------------------------------------------------ pyramid ops

// finest level: coefficient plane and hole count
__global__ __launch_bounds__(EW_NT) void k_prepare(const SpPlane *planes,
		const BlockWork *bw, int *bcnt)
{
	__shared__ int red[EW_NT / 64];
	const BlockWork b = bw[blockIdx.x];
	const SpPlane P = planes[b.plane];
	const size_t n = (size_t)P.w * P.h;
	int holes = 0;
	// the thread's loads of g go out together (an index past the plane
	// loads the plane's last pixel and is dropped)
	float gv[EW_PPT];
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		gv[k] = gload<float>(P.g + (q < n ? q : n - 1));
	}
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		const float g = gv[k];
		holes += isnan(g) ? 1 : 0;
		if (P.c) P.c[q] = coef_of(g, P.f, q);
	}
	int s = block_sum_int<EW_NT>(holes, red);
	if (threadIdx.x == 0) bcnt[blockIdx.x] = s;
}

// 2x2 NaN-aware mean of g (and f, times pmsfac), simpois.c:172-196,260-262
__global__ __launch_bounds__(EW_NT) void k_restrict(const SpPlane *fine,
		const SpPlane *coarse, const BlockWork *bw, double pmsfac, int *bcnt)
{
	__shared__ int red[EW_NT / 64];
	const BlockWork b = bw[blockIdx.x];
	const SpPlane F = fine[b.plane], C = coarse[b.plane];
	const size_t n = (size_t)C.w * C.h;
	int holes = 0;
	for (int k = 0; k < EW_PPT; k++) {
		size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		int i = (int)(q % C.w), j = (int)(q / C.w);
		float v[4], m = 0;
		int cnt = 0;
		v[0] = at(F.g, F.w, F.h, 2 * i, 2 * j);
		v[1] = at(F.g, F.w, F.h, 2 * i + 1, 2 * j);
		v[2] = at(F.g, F.w, F.h, 2 * i, 2 * j + 1);
		v[3] = at(F.g, F.w, F.h, 2 * i + 1, 2 * j + 1);
		for (int a = 0; a < 4; a++)
			if (isfinite(v[a])) { m = m + v[a]; cnt++; }
		float g = cnt ? m / (float)cnt : __uint_as_float(SP_QNAN_BITS);
		C.g[q] = g;
		if (C.f) {
			v[0] = at(F.f, F.w, F.h, 2 * i, 2 * j);
			v[1] = at(F.f, F.w, F.h, 2 * i + 1, 2 * j);
			v[2] = at(F.f, F.w, F.h, 2 * i, 2 * j + 1);
			v[3] = at(F.f, F.w, F.h, 2 * i + 1, 2 * j + 1);
			m = 0; cnt = 0;
			for (int a = 0; a < 4; a++)
				if (isfinite(v[a])) { m = m + v[a]; cnt++; }
			float fv = cnt ? m / (float)cnt : __uint_as_float(SP_QNAN_BITS);
			C.f[q] = (float)((double)fv * pmsfac);
		}
		holes += isnan(g) ? 1 : 0;
		if (C.c) C.c[q] = coef_of(g, C.f, q);
	}
	int s = block_sum_int<EW_NT>(holes, red);
	if (threadIdx.x == 0) bcnt[blockIdx.x] = s;
}

// one block per plane: its hole count is the sum of its work blocks' counts
__global__ __launch_bounds__(EW_NT) void k_count(const SpPlane *planes,
		const int *first, const int *bcnt)
{
	__shared__ int red[EW_NT / 64];
	int v = 0;
	for (int k = first[blockIdx.x] + threadIdx.x; k < first[blockIdx.x + 1]; k += EW_NT)
		v += bcnt[k];
	const int s = block_sum_int<EW_NT>(v, red);
	if (threadIdx.x == 0) *planes[blockIdx.x].cnt = s;
}

// The output of a level at q (level_out below): after a level, simpois
// writes the input back at every non-NaN pixel when the level has holes
// (cleant_cgpois.c:251), so the output is g there and u elsewhere.

// u = isfinite(g) ? g : init, init = bilinear x2 of the coarse output
// (simpois.c:159-160, 199-239) or 0 at the coarsest level
__global__ __launch_bounds__(EW_NT) void k_init(const SpPlane *lv,
		const SpPlane *coarse, const BlockWork *bw)
{
	const BlockWork b = bw[blockIdx.x];
	const SpPlane P = lv[b.plane];
	const SpPlane C = coarse[b.plane];   // C.g == nullptr: coarsest level
	const size_t n = (size_t)P.w * P.h;
	// two rounds of loads for the thread's pixels, each issued whole before
	// its values are used: g; then, at holes, the coarse level's output at
	// the four corners (its g and u)
	float gv[EW_PPT];
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		gv[k] = gload<float>(P.g + (q < n ? q : n - 1));
	}
	const bool cany = C.g && *C.cnt > 0;
	float cg[EW_PPT][4], cu[EW_PPT][4], fx[EW_PPT], fy[EW_PPT];
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (!(C.g && q < n && !isfinite(gv[k]))) continue;
		const int i = (int)(q % P.w), j = (int)(q / P.w);
		const float x = (float)((i - 0.5) / 2), y = (float)((j - 0.5) / 2);
		const int ix = (int)floorf(x), iy = (int)floorf(y);
		const int x0 = clampi(ix, 0, C.w - 1), x1 = clampi(ix + 1, 0, C.w - 1);
		const int y0 = clampi(iy, 0, C.h - 1), y1 = clampi(iy + 1, 0, C.h - 1);
		const size_t cq[4] = {(size_t)y0 * C.w + x0, (size_t)y0 * C.w + x1,
			(size_t)y1 * C.w + x0, (size_t)y1 * C.w + x1};
#pragma unroll
		for (int c = 0; c < 4; c++) {
			cg[k][c] = gload<float>(C.g + cq[c]);
			cu[k][c] = gload<float>(C.u + cq[c]);
		}
		fx[k] = x - (float)ix;
		fy[k] = y - (float)iy;
	}
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		const float g = gv[k];
		float init = 0.0f;
		if (C.g && !isfinite(g)) {
			float v[4];
#pragma unroll
			for (int c = 0; c < 4; c++)
				v[c] = cany && !isnan(cg[k][c]) ? cg[k][c] : cu[k][c];
			const float a = v[0], bb = v[1], cc = v[2], d = v[3];
			const float ffx = fx[k], ffy = fy[k];
			float r = 0.0f;
			r = r + a * (1 - ffx) * (1 - ffy);
			r = r + bb * ffx * (1 - ffy);
			r = r + cc * (1 - ffx) * ffy;
			r = r + d * ffx * ffy;
			init = r;
		}
		P.u[q] = isfinite(g) ? g : init;
	}
}

__global__ __launch_bounds__(EW_NT) void k_finalize(const SpPlane *lv,
		float *const *outs, const BlockWork *bw)
{
	const BlockWork b = bw[blockIdx.x];
	const SpPlane P = lv[b.plane];
	float *out = outs[b.plane];
	const size_t n = (size_t)P.w * P.h;
	if (out == P.u) {
		// in place (the finest level solves in the output buffer): u holds g
		// at finite known pixels already (k_init, and the sweeps and the CG
		// write holes only), so only the known +-Inf pixels, which the
		// solve saw as holes' initial values, change
		if (*P.cnt <= 0) return;
		float gv[EW_PPT];
#pragma unroll
		for (int k = 0; k < EW_PPT; k++) {
			const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
			gv[k] = gload<float>(P.g + (q < n ? q : n - 1));
		}
#pragma unroll
		for (int k = 0; k < EW_PPT; k++) {
			const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
			if (q < n && isinf(gv[k])) out[q] = gv[k];
		}
		return;
	}
	// the level's output at the thread's pixels, their loads issued together
	const bool any = *P.cnt > 0;
	float gv[EW_PPT], uv[EW_PPT];
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x, qc = q < n ? q : n - 1;
		gv[k] = gload<float>(P.g + qc);
		uv[k] = gload<float>(P.u + qc);
	}
#pragma unroll
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q < n) out[q] = any && !isnan(gv[k]) ? gv[k] : uv[k];
	}
}

void launch_prepare(const SpPlane *planes, const BlockWork *bw, int nblocks,
		int *bcnt, hipStream_t s)
{
	if (nblocks) k_prepare<<<nblocks, EW_NT, 0, s>>>(planes, bw, bcnt);
}

void launch_restrict(const SpPlane *fine, const SpPlane *coarse,
		const BlockWork *bw, int nblocks, double pmsfac, int *bcnt, hipStream_t s)
{
	if (nblocks) k_restrict<<<nblocks, EW_NT, 0, s>>>(fine, coarse, bw, pmsfac, bcnt);
}

void launch_count(const SpPlane *planes, int nplanes, const int *first,
		const int *bcnt, hipStream_t s)
{
	if (nplanes) k_count<<<nplanes, EW_NT, 0, s>>>(planes, first, bcnt);
}

void launch_init(const SpPlane *lv, const SpPlane *coarse, const BlockWork *bw,
		int nblocks, hipStream_t s)
{
	if (nblocks) k_init<<<nblocks, EW_NT, 0, s>>>(lv, coarse, bw);
}

void launch_finalize(const SpPlane *lv, float *const *outs,
		const BlockWork *bw, int nblocks, hipStream_t s)
{
	if (nblocks) k_finalize<<<nblocks, EW_NT, 0, s>>>(lv, outs, bw);
}

// ------------------------------------------------------------ GS kernels

// one tile: its footprint into LDS, its BT sweeps hyperplane by hyperplane,
// the pixels it updated back to the image
// (FLOW: the form of k_gs_flow, whose coefficients are already in LDS and
// whose values are read and written with agent-scope coherent accesses)
template <class S, bool FLOW = false>
static __device__ __forceinline__ void gs_tile_run(const GsTile &tl, const GsInst &in,
		float *smem)
{
	using G = Geo<S>;
	constexpr bool COH = FLOW;
	const TileCtx<S> c = tile_ctx<S>(tl, in, smem);
	tile_load_dev<S, COH, FLOW ? 2 : 0>(c, in);
	__syncthreads();
	if constexpr (S::R == 1) {
		Gs1Thread t = gs1_thread<S>(c, threadIdx.x);
		for (int sg = 0; sg < G::NSTEP; sg++) {
			gs1_step<S>(c, t, sg);
			__syncthreads();
		}
	} else {
		for (int sg = 0; sg < G::NSTEP; sg++) {
			tile_step<S>(c, threadIdx.x, sg);
			__syncthreads();
		}
	}
	if (COH) tile_store_coh<S>(c, in);
	else tile_store<S>(c, in, threadIdx.x);
}

template <class S>
__global__ __launch_bounds__(Geo<S>::NT) void k_gs_tile(const GsTile *tiles,
		const GsInst *insts)
{
	extern __shared__ float smem[];
	const GsTile tl = tiles[blockIdx.x];
	const GsInst in = insts[tl.inst];
	if (in.active && !*in.active) return;
	gs_tile_run<S>(tl, in, smem);
}

// the tiles of a level in one launch ordered by their dependences
// (tile_flow_loop)
template <class S>
__global__ __launch_bounds__(Geo<S>::NT) void k_gs_flow(const GsTile *tiles, int ntiles,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err)
{
	extern __shared__ float smem[];
	tile_flow_loop<S>(tiles, ntiles, insts, fi, claim, flags, epoch, spin_max, err, smem,
			[&](const GsTile &tl, const GsInst &in) { gs_tile_run<S, true>(tl, in, smem); });
}

int gs_flow_blocks(int R, int var)
{
	int nb = 0;
	sp_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		int per = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gs_flow<S>, Geo<S>::NT,
				Geo<S>::LDS_FLOATS * 4) != hipSuccess)
			per = 1;
		int dev = 0, ncu = 0;
		(void)hipGetDevice(&dev);
		if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)
				!= hipSuccess || ncu <= 0)
			ncu = 256;
		nb = std::max(1, per) * ncu;
	});
	return nb;
}

__global__ void k_gs_flow_reset(unsigned *claim) { *claim = 0; }

void launch_flow_reset(unsigned *claim, hipStream_t s)
{
	k_gs_flow_reset<<<1, 1, 0, s>>>(claim);
}

void launch_gs_flow(int R, int var, const GsTile *tiles, int ntiles, int nblocks,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, hipStream_t s)
{
	if (!ntiles) return;
	launch_flow_reset(claim, s);
	sp_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		k_gs_flow<S><<<nblocks, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(tiles, ntiles,
				insts, fi, claim, flags, epoch, spin_max, err);
	});
}

int tile_lds_bytes(int R, int var)
{
	int b = 0;
	sp_with_shape(R, var, [&](auto sh) { b = Geo<decltype(sh)>::LDS_FLOATS * 4; });
	return b;
}

void launch_gs_tiles(int R, int var, const GsTile *tiles, int ntiles,
		const GsInst *insts, hipStream_t s)
{
	if (!ntiles) return;
	sp_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		k_gs_tile<S><<<ntiles, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(tiles, insts);
	});
}

// whole level in one block (levels that fit in LDS)
template <int R>
__global__ __launch_bounds__(1024) void k_gs_small(const GsInst *insts, const int *ids)
{
	extern __shared__ float smem[];
	const GsInst in = insts[ids[blockIdx.x]];
	if (in.active && !*in.active) return;
	const int W = in.w, H = in.h, T = in.nsweep;
	if (T <= 0) return;
	float *U = smem, *C = smem + small_cells(R, W, H);
	small_load<R>(U, C, in, threadIdx.x, blockDim.x);
	SmallSlots sl;
	small_slots<R>(W, H, threadIdx.x, blockDim.x, sl);
	__syncthreads();
	const int ns = (int)small_nsteps(R, W, H, T);
	for (int s = 0; s < ns; s++) {
		small_step<R>(U, C, W, H, T, in.t, s, sl);
		__syncthreads();
	}
	small_store<R>(U, C, in, threadIdx.x, blockDim.x);
}

int small_lds_bytes(int R, int w, int h) { return 8 * small_cells(R, w, h); }

// exclusive prefix sum of v over the work-group (blockDim.x a multiple of
// 64, at most 1024) and the sum into *tot; wsum: 16 words of LDS
static __device__ int block_exscan(int v, int *wsum, int *tot)
{
	const int lane = (int)threadIdx.x & 63, wid = (int)threadIdx.x >> 6;
	const int nw = (int)blockDim.x >> 6;
	int x = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const int y = __shfl_up(x, o, 64);
		if (lane >= o) x += y;
	}
	if (lane == 63) wsum[wid] = x;
	__syncthreads();
	int pre = 0, sum = 0;
	for (int k = 0; k < nw; k++) {
		const int w = wsum[k];
		if (k < wid) pre += w;
		sum += w;
	}
	*tot = sum;
	return pre + x - v;
}

// whole R = 1 level by its unknown pixels (sp_gs.h gsc_*), one plane per
// work-group
template <int MAXS>
__global__ __launch_bounds__(GSC_MAXNT) void k_gs_cells(const GsInst *insts, const int *ids)
{
	extern __shared__ float smem[];
	const GsInst in = insts[ids[blockIdx.x]];
	if (in.active && !*in.active) return;
	const int W = in.w, H = in.h, T = in.nsweep;
	if (T <= 0) return;
	const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
	float *U = smem;
	int *L = (int *)(smem + (W + 2) * (H + 2));
	int *wsum = L + W * H;
	// the last step of each parity on which an update changed its pixel
	int *mark = wsum + 16;
	if (tid < 2) mark[tid] = -1;
	gsc_load(U, in, tid, nt);
	int tot;
	const int pre = block_exscan(gsc_count(in, tid, nt), wsum, &tot);
	gsc_fill(L, in, tid, nt, pre, tot);
	__syncthreads();
	GscThread<MAXS> th;
	gsc_take<MAXS>(th, U, L, in, tid, nt, tot);
	// even and odd steps in turn, each with its parity known at compile
	// time (a step of either parity behind one branch made the compiler
	// index the values by parity, through scratch memory).  An even step's
	// mark is read during the next (odd) step and an odd step's during the
	// next even one: each word is stable between the barrier after its
	// writes and the barrier after the following step, so every thread
	// reads the same values and leaves the loop at the same step.
	const int ns = (int)small_nsteps(1, W, H, T);
	const bool lead = (tid & 63) == 0;   // one writer per wave
	int last = -1;
	for (int s = 0; s < ns; s += 2) {
		if (__any(gsc_step_par<MAXS, 0>(th, U, W + 2, s, 2 * T, in.t)) && lead) mark[0] = s;
		__syncthreads();
		last = max(last, __builtin_amdgcn_readfirstlane(mark[0]));
		if (gsc_final(last, W, H, s) || s + 1 >= ns) break;
		if (__any(gsc_step_par<MAXS, 1>(th, U, W + 2, s + 1, 2 * T, in.t)) && lead)
			mark[1] = s + 1;
		__syncthreads();
		last = max(last, __builtin_amdgcn_readfirstlane(mark[1]));
		if (gsc_final(last, W, H, s + 1)) break;
	}
	gsc_store<MAXS>(th, in);
}

// slots per thread and parity of the compiled instances
static int gsc_inst(int maxs) { return maxs <= 1 ? 1 : maxs <= 2 ? 2 : maxs <= 4 ? 4 : 8; }

int gs_cells_enable(int bytes)
{
	for (int m : {1, 2, 4, 8}) {
		const void *f = m == 1 ? (const void *)k_gs_cells<1> : m == 2 ? (const void *)k_gs_cells<2>
			: m == 4 ? (const void *)k_gs_cells<4> : (const void *)k_gs_cells<8>;
		if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes)
				!= hipSuccess)
			return -1;
	}
	return 0;
}

void launch_gs_cells(int maxs, const GsInst *insts, const int *ids, int nids,
		int lds_bytes, int threads, hipStream_t s)
{
	if (!nids) return;
	switch (gsc_inst(maxs)) {
	case 1: k_gs_cells<1><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	case 2: k_gs_cells<2><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	case 4: k_gs_cells<4><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	default: k_gs_cells<8><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	}
}

void launch_gs_small(int R, const GsInst *insts, const int *ids, int nids,
		int lds_bytes, int threads, hipStream_t s)
{
	if (!nids) return;
	switch (R) {
	case 1: k_gs_small<1><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	case 2: k_gs_small<2><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	default: k_gs_small<3><<<nids, threads, lds_bytes, s>>>(insts, ids); break;
	}
}

// ------------------------------------------------------------------ CG
//
// Masked conjugate gradient of cleant_cgpois.c:197-264 / minicg.c:24-56.
// Vectors are full-grid doubles, zero at known pixels.  The operator reads
// each unknown rounded to float and evaluates its stencil in float, as the
// reference's float-returning getter does; dot products are double.

constexpr int CG_NT = 256;
// threads of the iteration kernels' blocks (4 unknowns each per pass):
// 512 measured against 256 and 1024 on the 2048^2 biharmonic fill (18.7
// against 19.7 ms per solve: half the partial sums of 256, more blocks
// per CU than 1024)
constexpr int CG_IT = 512;
constexpr int CG_ITU = 4;   // unknowns per thread and pass (2: 19.2, 8: 20.2 ms)

// stencil of A at (i,j) over values val(q) (float), Poisson = -(5-point),
// biharmonic = 13-point, in the reference's term order
template <typename V>
static __device__ __forceinline__ double cg_stencil(int bilap, int i, int j,
		int w, int h, V val)
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
		return -(double)r;
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
	return (double)r;
}

// sum of the nblk partials at part[0..nblk), same fixed tree in every block
template <int NT = CG_NT>
static __device__ double sum_partials(const double *part, int nblk, double *red)
{
	double v = 0;
	// unrolled so that the loads issue together; the adds keep their order
#pragma unroll 4
	for (int k = threadIdx.x; k < nblk; k += NT) v += gload<double>(part + k);
	double s = block_sum_dbl<NT>(v, red);
	__shared__ double bc;
	if (threadIdx.x == 0) bc = s;
	__syncthreads();
	return bc;
}

// The FP64 CG keeps its vectors compact, one entry per unknown in raster
// order (as the reference's mask3 list does, cleant_cgpois.c:6-25): the
// setup lists the unknowns (idx), and the iteration kernels stream over that
// list, so known pixels cost nothing.  The stencil reads p rounded to float
// from a full-grid copy pf (0 at known pixels), at clamped taps.  Indices
// are 32-bit (the solver takes at most 2^30 pixels per plane).
constexpr int CG_SPAN = CG_NT * 4;   // pixels per setup block (4 per thread)

// holes of each CG_SPAN-pixel block into bcnt; pf = 0 at known pixels
__global__ __launch_bounds__(CG_NT) void k_cg_count(const CgDev *dp)
{
	const CgDev d = *dp;
	__shared__ int red[CG_NT / 64];
	const unsigned n = (unsigned)d.w * (unsigned)d.h;
	const unsigned q0 = blockIdx.x * CG_SPAN + threadIdx.x * 4;
	int c = 0;
#pragma unroll
	for (int u = 0; u < 4; u++) {
		const unsigned q = q0 + u;
		if (q >= n) break;
		if (isnan(d.g[q])) c++;
		else d.pf[q] = 0.0f;
	}
	for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0) {
		int t = 0;
		for (int k = 0; k < CG_NT / 64; k++) t += red[k];
		d.bcnt[blockIdx.x] = t;
	}
}

// exclusive prefix sum of bcnt[0..nb) in place, bcnt[nb] = total (one
// work-group: each thread sums a contiguous run, the runs are scanned)
constexpr int CG_SCAN_T = 1024;
__global__ __launch_bounds__(CG_SCAN_T) void k_cg_scan(const CgDev *dp)
{
	const CgDev d = *dp;
	__shared__ int tot[CG_SCAN_T];
	const int nb = d.nb, per = (nb + CG_SCAN_T - 1) / CG_SCAN_T;
	const int a = threadIdx.x * per, e = min(a + per, nb);
	int s = 0;
	for (int k = a; k < e; k++) s += d.bcnt[k];
	tot[threadIdx.x] = s;
	__syncthreads();
	for (int o = 1; o < CG_SCAN_T; o <<= 1) {   // inclusive scan of the runs
		const int v = threadIdx.x >= o ? tot[threadIdx.x - o] : 0;
		__syncthreads();
		tot[threadIdx.x] += v;
		__syncthreads();
	}
	int run = threadIdx.x ? tot[threadIdx.x - 1] : 0;
	for (int k = a; k < e; k++) {
		const int v = d.bcnt[k];
		d.bcnt[k] = run;
		run += v;
	}
	if (threadIdx.x == CG_SCAN_T - 1) d.bcnt[nb] = tot[CG_SCAN_T - 1];
}

// list the unknowns of the block (raster order) and set up each:
// x0 = u; b = -f - A(boundary); r = b - A x0; p = r; partial r.r
__global__ __launch_bounds__(CG_NT) void k_cg_setup(const CgDev *dp)
{
	const CgDev d = *dp;
	__shared__ double red[CG_NT / 64];
	__shared__ int wsum[CG_NT / 64];
	const unsigned n = (unsigned)d.w * (unsigned)d.h;
	const unsigned q0 = blockIdx.x * CG_SPAN + threadIdx.x * 4;
	bool hole[4];
	int c = 0;
#pragma unroll
	for (int u = 0; u < 4; u++) {
		hole[u] = q0 + u < n && isnan(d.g[q0 + u]);
		c += hole[u];
	}
	// rank of the thread's first unknown within the block
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	int incl = c;
	for (int o = 1; o < 64; o <<= 1) {
		const int v = __shfl_up(incl, o, 64);
		if (lane >= o) incl += v;
	}
	if (lane == 63) wsum[wv] = incl;
	__syncthreads();
	int k = d.bcnt[blockIdx.x] + incl - c;
	for (int z = 0; z < wv; z++) k += wsum[z];
	double acc = 0;
	const float *g = d.g, *uu = d.u;
#pragma unroll
	for (int u = 0; u < 4; u++) {
		if (!hole[u]) continue;
		const unsigned q = q0 + u;
		const int j = (int)(q / (unsigned)d.w), i = (int)(q - (unsigned)j * d.w);
		const double a0 = cg_stencil(d.bilap, i, j, d.w, d.h, [&](size_t t) {
			const float v = g[t]; return isnan(v) ? 0.0f : v; });
		const double ax = cg_stencil(d.bilap, i, j, d.w, d.h, [&](size_t t) {
			return isnan(g[t]) ? uu[t] : 0.0f; });
		const float dv = d.f ? -d.f[q] : 0.0f;
		const double b = (double)dv - a0;
		const double r = b - ax;
		d.idx[k] = q;
		d.x[k] = (double)uu[q];
		d.r[k] = r;
		d.p[k] = r;
		d.pf[q] = (float)r;
		acc += r * r;
		k++;
	}
	double s = block_sum_dbl<CG_NT>(acc, red);
	if (threadIdx.x == 0) d.part[blockIdx.x] = s;
}

__global__ __launch_bounds__(CG_NT) void k_cg_rr0(const CgDev *dp)
{
	const CgDev d = *dp;
	__shared__ double red[CG_NT / 64];
	double s = sum_partials(d.part, d.nb, red);
	if (threadIdx.x == 0) {
		d.rr[0] = s;
		d.rrp[0] = s;
		d.done[0] = 0;
		d.done[1] = 0;
		d.done[2] = -1;   // iteration index, advanced by each k_cg_spmv
		d.done[3] = 0;
		d.done[4] = 0;
		d.done[5] = 0;   // k_cg_persist's error word, gate word and barrier words
		d.done[7] = 0;
		for (int k = 0; k < 36; k++) d.bar[64 * k] = 0;
	}
}

// The stop flag of iteration k's kernels is done[3 + (k & 1)]: k_cg_dir of
// the stopping iteration k sets done[3 + ((k + 1) & 1)], and k_cg_spmv of
// iteration k + 1, finding it set, sets done[3 + (k & 1)].  So no kernel
// writes the word it reads (every block of the stopping k_cg_dir still
// updates x, however late it starts), and all threads read one value.
static __device__ __forceinline__ bool cg_stopped(const CgDev &d, int par)
{
	return __hip_atomic_load((__attribute__((address_space(1))) int *)d.done + 3 + par,
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// The iteration kernels walk the list with the same grid stride (entry t,
// t + S, t + 2S, ... with S = nblk * CG_NT) and accumulate in that order, in
// batches of U entries whose loads are issued before the first value is
// used (through the global address space: flat loads under a branch made the
// compiler wait for each batch member before the next).

// taps of A at (i, j) in the reference's term order (cleant_cgpois.c:91-121):
// Poisson (i,j), (i+1,j), (i,j+1), (i-1,j), (i,j-1); the biharmonic adds the
// diagonals and the distance-2 taps
template <int BILAP>
struct CgTaps {
	static constexpr int N = BILAP ? 13 : 5;
};
static __device__ __forceinline__ void cg_tap_offset(int k, int &di, int &dj)
{
	constexpr int DI[13] = {0, 1, 0, -1, 0, 1, 1, -1, -1, 2, 0, -2, 0};
	constexpr int DJ[13] = {0, 0, 1, 0, -1, 1, -1, 1, -1, 0, 2, 0, -2};
	di = DI[k];
	dj = DJ[k];
}

// the stencil over the taps' values (p rounded to float), summed in float
// in the reference's order (the same operations as cg_stencil)
template <int BILAP>
static __device__ __forceinline__ double cg_apply_taps(const float (&t)[CgTaps<BILAP>::N])
{
	if (!BILAP) {
		float r = -4.0f * t[0];
		r = r + t[1];
		r = r + t[2];
		r = r + t[3];
		r = r + t[4];
		return -(double)r;
	}
	float r = 20.0f * t[0];
#pragma unroll
	for (int k = 1; k < 5; k++) r = r - 8.0f * t[k];
#pragma unroll
	for (int k = 5; k < 9; k++) r = r + 2.0f * t[k];
#pragma unroll
	for (int k = 9; k < 13; k++) r = r + t[k];
	return (double)r;
}

// a relaxed atomic load (4 or 8 bytes): an ordered access, which the
// compiler neither moves behind a later branch nor drops
template <typename T>
static __device__ __forceinline__ T cg_aload(const T *p)
{
	static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4 or 8 bytes");
	using B = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;
	const B v = __hip_atomic_load((__attribute__((address_space(1))) B *)p,
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return __builtin_bit_cast(T, v);
}

// q = A p, partial p.q
template <int BILAP, int U>
__global__ __launch_bounds__(CG_IT) void k_cg_spmv(const CgDev *dp, int par)
{
	const CgDev d = *dp;
	__shared__ double red[CG_IT / 64];
	// the launch may have more blocks than this level uses (a graph's grid
	// is fixed); a block's whole work is skipped, so no barrier is split
	if (blockIdx.x >= (unsigned)d.nblk) return;
	const bool stopped = cg_stopped(d, par);
	constexpr int NT = CgTaps<BILAP>::N;
	const int w = d.w, h = d.h;
	const unsigned nc = (unsigned)d.nc, S = (unsigned)d.nblk * CG_IT;
	const float *pf = d.pf;
	unsigned k0 = blockIdx.x * CG_IT + threadIdx.x;
	bool on[U];
	unsigned qq[U];
	double pc[U];
	// unconditional loads at clamped indices, as atomic loads (see
	// cg_part_load: plain ones were moved behind the stop test)
	auto load = [&]() {
#pragma unroll
		for (int u = 0; u < U; u++) {
			const unsigned k = k0 + u * S;
			on[u] = k < nc;
			qq[u] = cg_aload(d.idx + min(k, nc - 1));
			pc[u] = cg_aload(d.p + min(k, nc - 1));
		}
	};
	// the first pass's list and p loads are in flight when the flag is
	// tested
	load();
	if (stopped) {
		// passed on to the next iteration's kernels (see cg_stopped)
		if (blockIdx.x == 0 && threadIdx.x == 0) d.done[4 - par] = 1;
		return;
	}
	double acc = 0;
	for (;;) {
		// every entry's taps, also those past the list's end (their qq is
		// the last entry's, a valid pixel): loads under a per-entry branch
		// made the compiler wait for one entry's taps before the next's
		float t[U][NT];
#pragma unroll
		for (int u = 0; u < U; u++) {
			const int j = (int)(qq[u] / (unsigned)w), i = (int)(qq[u] - (unsigned)j * w);
#pragma unroll
			for (int e = 0; e < NT; e++) {
				int di, dj;
				cg_tap_offset(e, di, dj);
				const int a = clampi(i + di, 0, w - 1), b = clampi(j + dj, 0, h - 1);
				t[u][e] = gload<float>(pf + (unsigned)b * (unsigned)w + (unsigned)a);
			}
		}
#pragma unroll
		for (int u = 0; u < U; u++) {
			if (!on[u]) continue;
			const double y = cg_apply_taps<BILAP>(t[u]);
			gstore<float>(d.q + k0 + u * S, (float)y);   // y is a float widened: exact
			acc += y * pc[u];
		}
		k0 += U * S;
		if (k0 >= nc) break;
		load();
	}
	double s = block_sum_dbl<CG_IT>(acc, red);
	if (threadIdx.x == 0) d.part[blockIdx.x] = s;
	// the first kernel of an iteration advances the index k_cg_dir records
	// the r.r history at
	if (blockIdx.x == 0 && threadIdx.x == 0) d.done[2] += 1;
}

// The partials of the previous kernel, two per thread (nblk <= 2 * CG_IT):
// loaded first, summed later in sum_partials<CG_IT>'s order and tree.  The
// loads are unconditional, at indices clamped into [0, nblk) (a load under a
// branch made the compiler wait for it at the join), and the sum skips the
// clamped ones.  They are relaxed atomic loads, which the compiler keeps
// ahead of the stop test (it moved plain loads behind it).
static __device__ __forceinline__ double cg_part_at(const double *p)
{
	return cg_aload(p);
}
static __device__ __forceinline__ void cg_part_load(const double *part, int nblk,
		double (&pp)[2])
{
	const int k = (int)threadIdx.x;
	pp[0] = cg_part_at(part + min(k, nblk - 1));
	pp[1] = cg_part_at(part + min(k + CG_IT, nblk - 1));
}
static __device__ __forceinline__ double cg_part_sum(const double (&pp)[2], int nblk,
		double *red)
{
	const int k = (int)threadIdx.x;
	double v = 0;
	if (k < nblk) v += pp[0];
	if (k + CG_IT < nblk) v += pp[1];
	return block_sum_dbl<CG_IT>(v, red);   // valid in thread 0
}

// The first pass of k_cg_update and k_cg_dir is written out before their
// loop: the stop flag, r.r, the pass's vectors and the partials are all in
// flight before the flag is tested (inside a loop the compiler tested the
// flag ahead of the vector loads, one memory round trip more).

// alpha = rr_old / p.q ; r -= alpha q ; partial r.r
template <int U>
__global__ __launch_bounds__(CG_IT) void k_cg_update(const CgDev *dp, int par)
{
	const CgDev d = *dp;
	__shared__ double red[CG_IT / 64];
	__shared__ double alpha_s;
	if (blockIdx.x >= (unsigned)d.nblk) return;
	const bool stopped = cg_stopped(d, par);
	const double rr0 = gload<double>(d.rrp + par);
	const unsigned nc = (unsigned)d.nc, S = (unsigned)d.nblk * CG_IT;
	unsigned k0 = blockIdx.x * CG_IT + threadIdx.x;
	double rv[U];
	float qv[U];   // widened at the use: a conversion here would wait
	auto load = [&]() {   // unconditional, clamped (as cg_part_load)
#pragma unroll
		for (int u = 0; u < U; u++) {
			const unsigned k = min(k0 + u * S, nc - 1);
			rv[u] = gload<double>(d.r + k);
			qv[u] = gload<float>(d.q + k);
		}
	};
	load();
	double pp[2];
	cg_part_load(d.part, d.nblk, pp);
	if (stopped) return;   // before any barrier, the whole block
	const double pq = cg_part_sum(pp, d.nblk, red);
	if (threadIdx.x == 0) alpha_s = rr0 / pq;
	__syncthreads();
	const double alpha = alpha_s;
	double acc = 0;
	for (;;) {
#pragma unroll
		for (int u = 0; u < U; u++) {
			const unsigned k = k0 + u * S;
			if (k >= nc) continue;
			const double r = rv[u] - alpha * (double)qv[u];
			gstore<double>(d.r + k, r);
			acc += r * r;
		}
		k0 += U * S;
		if (k0 >= nc) break;
		load();
	}
	double s = block_sum_dbl<CG_IT>(acc, red);
	if (threadIdx.x == 0) d.part[d.nblk + blockIdx.x] = s;
}

// rr_new; x += alpha p (alpha recomputed from the same partials, in the same
// order, as k_cg_update did); absolute stop test; beta; p = r + beta p and
// its float copy.  x is updated here, where p is read anyway, also in the
// iteration that stops.
template <int U>
__global__ __launch_bounds__(CG_IT) void k_cg_dir(const CgDev *dp, int par)
{
	const CgDev d = *dp;
	__shared__ double red[CG_IT / 64];
	__shared__ double alpha_s, rr1_s;
	if (blockIdx.x >= (unsigned)d.nblk) return;
	const bool stopped = cg_stopped(d, par);
	// the history index, for block 0's record at the end
	const int it = blockIdx.x == 0 ? gload<int>(d.done + 2) : 0;
	const unsigned nc = (unsigned)d.nc, S = (unsigned)d.nblk * CG_IT;
	const double rr0 = gload<double>(d.rrp + par);
	unsigned k0 = blockIdx.x * CG_IT + threadIdx.x;
	double rv[U], pv[U], xv[U];
	unsigned qq[U];
	auto load = [&]() {   // unconditional, clamped (as cg_part_load)
#pragma unroll
		for (int u = 0; u < U; u++) {
			const unsigned k = min(k0 + u * S, nc - 1);
			rv[u] = gload<double>(d.r + k);
			pv[u] = gload<double>(d.p + k);
			xv[u] = gload<double>(d.x + k);
			qq[u] = gload<unsigned>(d.idx + k);
		}
	};
	load();
	double pa[2], pb[2];
	cg_part_load(d.part, d.nblk, pa);
	cg_part_load(d.part + d.nblk, d.nblk, pb);
	if (stopped) return;   // before any barrier, the whole block
	// p.q then r.r, each in sum_partials' order
	const double spq = cg_part_sum(pa, d.nblk, red);
	const double srr = cg_part_sum(pb, d.nblk, red);
	if (threadIdx.x == 0) { alpha_s = rr0 / spq; rr1_s = srr; }
	__syncthreads();
	const double alpha = alpha_s, rr1 = rr1_s;
	const bool stop = sqrt(rr1) < d.stop;
	const double beta = rr1 / rr0;
	for (;;) {
#pragma unroll
		for (int u = 0; u < U; u++) {
			const unsigned k = k0 + u * S;
			if (k >= nc) continue;
			gstore<double>(d.x + k, xv[u] + alpha * pv[u]);
			if (stop) continue;
			const double pn = rv[u] + beta * pv[u];
			gstore<double>(d.p + k, pn);
			gstore<float>(d.pf + qq[u], (float)pn);
		}
		k0 += U * S;
		if (k0 >= nc) break;
		load();
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		d.rr[it + 1] = rr1;
		d.rrp[1 - par] = rr1;   // the next iteration's r.r (this one reads rrp[par])
		if (sqrt(rr1) < d.stop) {
			d.done[1] = it + 1;
			d.done[0] = 1;
			d.done[4 - par] = 1;   // the next iteration's stop flag
		}
	}
}

// word 3 of a raw buffer resource on gfx9 (untyped 32-bit data, no
// swizzle) and the cache policy bit sc1: an access coherent at agent scope,
// the bit the compiler sets for cg_aload
constexpr int CG_RSRC_WORD3 = 0x00020000;
constexpr int CG_SC1 = 16;

// a relaxed atomic store (4 or 8 bytes) at agent scope: coherent across the
// XCDs' L2 caches, for values other work-groups of the launch read
template <typename T>
static __device__ __forceinline__ void cg_astore(T *p, T v)
{
	static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4 or 8 bytes");
	using B = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;
	__hip_atomic_store((__attribute__((address_space(1))) B *)p, __builtin_bit_cast(B, v),
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier of k_cg_persist, e-th (from 1): work-group b counts in at
// group counter bar[64 * (b % G)] (G = 1 << lg groups; with G a multiple of
// 8 a group's blocks share an XCD, block i running on XCD i mod 8), the
// last of a group at the global counter bar[64 * 32], and the last of all
// writes e to the release word bar[64 * 33], which the work-groups poll
// (the words are 256 bytes apart: arrivals and polls do not share a line).
// Each work-group's waves have completed their stores before it counts in.
// A wait past spin_max polls, or one that finds *err set, sets *err and
// returns false (the work-group leaves; so does every other one at its next
// wait).
static __device__ __forceinline__ bool cg_grid_sync(unsigned *bar, unsigned e, int nblk,
		int lg, int *err, int spin_max)
{
	__shared__ int s_ok;
	__builtin_amdgcn_s_waitcnt(0);   // this wave's stores have completed
	__syncthreads();
	if (threadIdx.x == 0) {
		const int G = 1 << lg, g = (int)blockIdx.x & (G - 1), ng = nblk < G ? nblk : G;
		const unsigned ing = (unsigned)((nblk >> lg) + (g < (nblk & (G - 1))));   // in group g
		typedef __attribute__((address_space(1))) unsigned gu;
		const unsigned o = __hip_atomic_fetch_add((gu *)(bar + 64 * g), 1u,
				__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (o == e * ing - 1) {
			const unsigned o2 = __hip_atomic_fetch_add((gu *)(bar + 64 * 32), 1u,
					__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (o2 == e * (unsigned)ng - 1) cg_astore(bar + 64 * 33, e);
		}
		int ok = 1, n = 0;
		while ((int)(cg_aload(bar + 64 * 33) - e) < 0) {
			if (++n > spin_max || ((n & 255) == 0 && cg_aload(err))) {
				ok = 0;
				cg_astore(err, 1);
				break;
			}
			__builtin_amdgcn_s_sleep(1);
		}
		s_ok = ok;
	}
	__syncthreads();
	return s_ok != 0;
}

// Start gate of k_cg_persist: every work-group counts in at bar[64 * 35]
// and waits for the decision word bar[64 * 34]: 1 (go), set by the last of
// the nblk to count in, or 2 (leave), set by a work-group that waited more
// than spin_max polls; one compare-and-swap from 0 decides, and every
// work-group acts on the one value.  Go means all work-groups were running
// at once, and none leaves before the end, so all stay resident for the
// barriers; with leave no work-group has done any work (a launch whose
// work-groups cannot all be resident, as when other work holds CUs), and
// the host runs the iterations as separate launches instead.
static __device__ __forceinline__ bool cg_start_gate(unsigned *bar, int nblk, int *done,
		int spin_max)
{
	__shared__ int s_go;
	if (threadIdx.x == 0) {
		typedef __attribute__((address_space(1))) unsigned gu;
		gu *dec = (gu *)(bar + 64 * 34);
		const unsigned o = __hip_atomic_fetch_add((gu *)(bar + 64 * 35), 1u,
				__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (o == (unsigned)nblk - 1) {
			unsigned z = 0;
			__hip_atomic_compare_exchange_strong(dec, &z, 1u, __ATOMIC_RELAXED,
					__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		unsigned f;
		int n = 0;
		while ((f = cg_aload(bar + 64 * 34)) == 0) {
			if (++n > spin_max) {
				unsigned z = 0;
				__hip_atomic_compare_exchange_strong(dec, &z, 2u, __ATOMIC_RELAXED,
						__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			}
			__builtin_amdgcn_s_sleep(1);
		}
		s_go = f == 1;
		if (f != 1) done[7] = 1;   // for the host: the launch did nothing
	}
	__syncthreads();
	return s_go != 0;
}

// The FP64 CG's iterations in one launch whose work-groups are all resident
// (checked by cg_start_gate), for a level whose unknowns take one pass of the
// iteration kernels (nc <= CG_ITU * nblk * CG_IT).  Each thread keeps its
// entries' x, r and p in registers from the first iteration to the last;
// only p's float copy (read by the neighbours' stencils) and the per-block
// partial sums go through memory, with agent-scope coherent stores and
// loads, and three grid barriers per iteration stand for the boundaries of
// k_cg_spmv, k_cg_update and k_cg_dir.  Same entries per thread, same
// arithmetic and the same partial sums in the same trees as those kernels:
// the same iterates, r.r history and stop.  Writes x, the history and
// done[0..1] as they do.  bar holds the gate and barrier words, done[5] is
// the error word and done[7] the gate's give-up mark, all cleared by
// k_cg_rr0.
// (at most 128 registers: two work-groups resident per CU)
template <int BILAP>
__global__ __launch_bounds__(CG_IT) __attribute__((amdgpu_waves_per_eu(4)))
void k_cg_persist(const CgDev *dp, int maxit, int spin_max, int gate_spin, int lg)
{
	constexpr int U = CG_ITU, NT = CgTaps<BILAP>::N;
	const CgDev d = *dp;
	__shared__ double red[CG_IT / 64];
	__shared__ double s_alpha, s_rr1;
	const int nblk = d.nblk, w = d.w, h = d.h;
	const unsigned nc = (unsigned)d.nc, S = (unsigned)nblk * CG_IT;
	const unsigned k0 = blockIdx.x * CG_IT + threadIdx.x;
	unsigned *bar = d.bar;
	int *err = d.done + 5;
	if (!cg_start_gate(bar, nblk, d.done, gate_spin)) return;
	float *pf = d.pf;
	// the stencils read pf through a buffer resource: a 32-bit byte offset
	// per tap (one register per address instead of two; the plane has at
	// most 2^30 pixels) and the agent-coherent cache policy of cg_aload
	const unsigned long long pfb = (unsigned long long)w * (unsigned)h * 4u;
	const __amdgpu_buffer_rsrc_t pfr = __builtin_amdgcn_make_buffer_rsrc(pf, (short)0,
			(int)(unsigned)(pfb < 0xfffffffcull ? pfb : 0xfffffffcull), CG_RSRC_WORD3);
	bool on[U];
	unsigned qq[U];
	double x[U], r[U], p[U];
#pragma unroll
	for (int u = 0; u < U; u++) {
		const unsigned k = k0 + u * S;
		on[u] = k < nc;
		const unsigned kc = min(k, nc - 1);   // past the end: the last entry, unused
		qq[u] = gload<unsigned>(d.idx + kc);
		x[u] = gload<double>(d.x + kc);
		r[u] = gload<double>(d.r + kc);
		p[u] = gload<double>(d.p + kc);
	}
	double rr0 = gload<double>(d.rr);
	unsigned epoch = 0;
	int it = 0, stopped = 0;
	while (it < maxit) {
		// the 13 U tap offsets are recomputed here in every iteration: held
		// from one to the next (the compiler's choice, as they do not
		// change) they took 52 registers and made the kernel spill
#pragma unroll
		for (int u = 0; u < U; u++) asm volatile("" : "+v"(qq[u]));
		// q = A p and the partial p.q (k_cg_spmv)
		float t[U][NT];
#pragma unroll
		for (int u = 0; u < U; u++) {
			const int j = (int)(qq[u] / (unsigned)w), i = (int)(qq[u] - (unsigned)j * w);
#pragma unroll
			for (int e = 0; e < NT; e++) {
				int di, dj;
				cg_tap_offset(e, di, dj);
				const int a = clampi(i + di, 0, w - 1), b = clampi(j + dj, 0, h - 1);
				const unsigned off = ((unsigned)b * (unsigned)w + (unsigned)a) << 2;
				t[u][e] = __builtin_bit_cast(float,
						__builtin_amdgcn_raw_buffer_load_b32(pfr, off, 0, CG_SC1));
			}
		}
		double q[U], acc = 0;
#pragma unroll
		for (int u = 0; u < U; u++) {
			q[u] = cg_apply_taps<BILAP>(t[u]);
			if (on[u]) acc += q[u] * p[u];
		}
		double sb = block_sum_dbl<CG_IT>(acc, red);
		if (threadIdx.x == 0) cg_astore(d.part + blockIdx.x, sb);
		if (!cg_grid_sync(bar, ++epoch, nblk, lg, err, spin_max)) return;
		// alpha = rr0 / p.q, r -= alpha q and the partial r.r (k_cg_update)
		double pa[2];
		cg_part_load(d.part, nblk, pa);
		const double pq = cg_part_sum(pa, nblk, red);
		if (threadIdx.x == 0) s_alpha = rr0 / pq;
		__syncthreads();
		const double alpha = s_alpha;
		acc = 0;
#pragma unroll
		for (int u = 0; u < U; u++) {
			if (!on[u]) continue;
			r[u] = r[u] - alpha * q[u];
			acc += r[u] * r[u];
		}
		sb = block_sum_dbl<CG_IT>(acc, red);
		if (threadIdx.x == 0) cg_astore(d.part + nblk + blockIdx.x, sb);
		if (!cg_grid_sync(bar, ++epoch, nblk, lg, err, spin_max)) return;
		// r.r, x += alpha p, the stop test, p = r + beta p (k_cg_dir)
		double pb[2];
		cg_part_load(d.part + nblk, nblk, pb);
		const double rs = cg_part_sum(pb, nblk, red);
		if (threadIdx.x == 0) s_rr1 = rs;
		__syncthreads();
		const double rr1 = s_rr1;
		const bool stop = sqrt(rr1) < d.stop;
		const double beta = rr1 / rr0;
#pragma unroll
		for (int u = 0; u < U; u++) {
			if (!on[u]) continue;
			x[u] = x[u] + alpha * p[u];
			if (stop) continue;
			p[u] = r[u] + beta * p[u];
			cg_astore(pf + qq[u], (float)p[u]);
		}
		if (blockIdx.x == 0 && threadIdx.x == 0) d.rr[it + 1] = rr1;
		rr0 = rr1;
		it++;
		if (stop) { stopped = 1; break; }
		if (it < maxit) {
			// p's float copy complete before the next stencils read it
			if (!cg_grid_sync(bar, ++epoch, nblk, lg, err, spin_max)) return;
		}
	}
#pragma unroll
	for (int u = 0; u < U; u++)
		if (on[u]) d.x[k0 + u * S] = x[u];
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		d.done[0] = stopped;
		d.done[1] = it;
	}
}

int cg_persist_capacity(int bilap)
{
	const void *f = bilap ? (const void *)k_cg_persist<1> : (const void *)k_cg_persist<0>;
	int per = 0, dev = 0, ncu = 0;
	if (hipGetDevice(&dev) != hipSuccess
			|| hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, CG_IT, 0) != hipSuccess
			|| hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)
			!= hipSuccess)
		return 0;
	return per * ncu;
}

void launch_cg_persist(const CgDev *dp, int bilap, int nblk, int maxit, int spin_max,
		int gate_spin, int lg, hipStream_t s)
{
	if (bilap) k_cg_persist<1><<<nblk, CG_IT, 0, s>>>(dp, maxit, spin_max, gate_spin, lg);
	else k_cg_persist<0><<<nblk, CG_IT, 0, s>>>(dp, maxit, spin_max, gate_spin, lg);
}

// u[holes] = (float) x
__global__ __launch_bounds__(CG_NT) void k_cg_finish(const CgDev *dp)
{
	const CgDev d = *dp;
	const unsigned nc = (unsigned)d.nc;
	for (unsigned k = blockIdx.x * CG_NT + threadIdx.x; k < nc; k += (unsigned)d.nblk * CG_NT)
		d.u[d.idx[k]] = (float)d.x[k];
}

int cg_setup_blocks(int w, int h)
{
	return (int)(((size_t)w * h + CG_SPAN - 1) / CG_SPAN);
}

void launch_cg_setup(const CgDev *dp, int nb, hipStream_t s)
{
	k_cg_count<<<nb, CG_NT, 0, s>>>(dp);
	k_cg_scan<<<1, CG_SCAN_T, 0, s>>>(dp);
	k_cg_setup<<<nb, CG_NT, 0, s>>>(dp);
	k_cg_rr0<<<1, CG_NT, 0, s>>>(dp);
}

int cg_iter_span() { return CG_ITU * CG_IT; }

void launch_cg_iter(const CgDev *dp, int bilap, int grid, int par, hipStream_t s)
{
	if (bilap) k_cg_spmv<1, CG_ITU><<<grid, CG_IT, 0, s>>>(dp, par);
	else k_cg_spmv<0, CG_ITU><<<grid, CG_IT, 0, s>>>(dp, par);
	k_cg_update<CG_ITU><<<grid, CG_IT, 0, s>>>(dp, par);
	k_cg_dir<CG_ITU><<<grid, CG_IT, 0, s>>>(dp, par);
}

void launch_cg_finish(const CgDev *dp, int nblk, hipStream_t s)
{
	k_cg_finish<<<nblk, CG_NT, 0, s>>>(dp);
}

// ------------------------------------------------------------- residual

// per block: sum over holes of (b - A u)^2 and b^2, in double from float
// data, for the masked system of cleant_cgpois.c:225-236
__global__ __launch_bounds__(EW_NT) void k_residual(const SpPlane *lv,
		const BlockWork *bw, int bilap, double *part)
{
	__shared__ double red[EW_NT / 64];
	const BlockWork b = bw[blockIdx.x];
	const SpPlane P = lv[b.plane];
	const size_t n = (size_t)P.w * P.h;
	double rr = 0, bb = 0;
	for (int k = 0; k < EW_PPT; k++) {
		size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		if (!isnan(P.g[q])) continue;
		const int i = (int)(q % P.w), j = (int)(q / P.w);
		double au = 0, a0 = 0;
		auto term = [&](int di, int dj, double cf) {
			size_t k2 = (size_t)clampi(j + dj, 0, P.h - 1) * P.w
				+ clampi(i + di, 0, P.w - 1);
			au += cf * (double)P.u[k2];
			float gv = P.g[k2];
			if (!isnan(gv)) a0 += cf * (double)gv;
		};
		if (!bilap) {
			term(0, 0, -4); term(1, 0, 1); term(0, 1, 1); term(-1, 0, 1);
			term(0, -1, 1);
			au = -au; a0 = -a0;
		} else {
			term(0, 0, 20); term(1, 0, -8); term(0, 1, -8); term(-1, 0, -8);
			term(0, -1, -8); term(1, 1, 2); term(1, -1, 2); term(-1, 1, 2);
			term(-1, -1, 2); term(2, 0, 1); term(0, 2, 1); term(-2, 0, 1);
			term(0, -2, 1);
		}
		double bv = (P.f ? -(double)P.f[q] : 0.0) - a0;
		double res = bv - (au - a0);
		rr += res * res;
		bb += bv * bv;
	}
	double s1 = block_sum_dbl<EW_NT>(rr, red);
	double s2 = block_sum_dbl<EW_NT>(bb, red);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = s1;
		part[2 * blockIdx.x + 1] = s2;
	}
}

// out[3p..3p+2] = (rr, bb, rel) for plane p; blocks of plane p are
// plane_first[p] .. plane_first[p+1]-1
__global__ __launch_bounds__(EW_NT) void k_residual_reduce(const double *part,
		const int *plane_first, int nplanes, double *out)
{
	__shared__ double red[EW_NT / 64];
	const int p = blockIdx.x;
	double rr = 0, bb = 0;
	for (int k = plane_first[p] + threadIdx.x; k < plane_first[p + 1]; k += EW_NT) {
		rr += part[2 * k];
		bb += part[2 * k + 1];
	}
	rr = block_sum_dbl<EW_NT>(rr, red);
	bb = block_sum_dbl<EW_NT>(bb, red);
	if (threadIdx.x == 0) {
		out[3 * p] = rr;
		out[3 * p + 1] = bb;
		out[3 * p + 2] = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
	}
}

void launch_residual(const SpPlane *lv, const BlockWork *bw, int nblocks,
		int bilap, double *part, hipStream_t s)
{
	if (nblocks) k_residual<<<nblocks, EW_NT, 0, s>>>(lv, bw, bilap, part);
}

// the descriptors, work blocks (per of EW_NT * EW_PPT pixels for each of pd
// planar w x h planes) and first blocks of a residual, written on the device
__global__ __launch_bounds__(EW_NT) void k_residual_plan(SpPlane *lv, BlockWork *bw,
		int *first, const float *u, const float *g, const float *f, int w, int h,
		int pd, int per)
{
	const int t = blockIdx.x * EW_NT + threadIdx.x;
	const size_t n = (size_t)w * h;
	if (t < pd * per) bw[t] = BlockWork{t / per, (t % per) * (EW_NT * EW_PPT)};
	if (t < pd)
		lv[t] = SpPlane{const_cast<float *>(g) + n * t,
			f ? const_cast<float *>(f) + n * t : nullptr,
			const_cast<float *>(u) + n * t, nullptr, nullptr, w, h};
	if (t <= pd) first[t] = t * per;
}

void launch_residual_plan(SpPlane *lv, BlockWork *bw, int *first, const float *u,
		const float *g, const float *f, int w, int h, int pd, int per, hipStream_t s)
{
	const int n = std::max(pd * per, pd + 1);
	k_residual_plan<<<(n + EW_NT - 1) / EW_NT, EW_NT, 0, s>>>(lv, bw, first, u, g, f,
			w, h, pd, per);
}

// plain streaming copy: each thread moves 4 x 16 bytes (all loads issued
// before the stores), one pass over the buffer; the measured bandwidth
// ceiling of the device and the calibration of the HBM counters
constexpr int SC_T = 256, SC_U = 4;
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(SC_T) void k_stream_copy(v4f *__restrict__ dst,
		const v4f *__restrict__ src, size_t n4)
{
	const size_t q0 = (size_t)blockIdx.x * (SC_T * SC_U) + threadIdx.x;
	v4f v[SC_U];
#pragma unroll
	for (int u = 0; u < SC_U; u++) {
		const size_t q = q0 + (size_t)u * SC_T;
		if (q < n4) v[u] = __builtin_nontemporal_load(src + q);
	}
#pragma unroll
	for (int u = 0; u < SC_U; u++) {
		const size_t q = q0 + (size_t)u * SC_T;
		if (q < n4) __builtin_nontemporal_store(v[u], dst + q);
	}
}

void launch_stream_copy(float *dst, const float *src, size_t n4, hipStream_t s)
{
	const size_t blocks = (n4 + SC_T * SC_U - 1) / (SC_T * SC_U);
	if (blocks)
		k_stream_copy<<<(unsigned)blocks, SC_T, 0, s>>>((v4f *)dst,
				(const v4f *)src, n4);
}

void launch_residual_reduce(const double *part, const int *plane_first,
		int nplanes, double *out, hipStream_t s)
{
	if (nplanes)
		k_residual_reduce<<<nplanes, EW_NT, 0, s>>>(part, plane_first, nplanes, out);
}

} // namespace spk

// ------------------------------------------------------------- adjoint
//
// Converged solve of the masked Poisson system A u_h = -f_h - A_hk g_k
// (A = -Laplacian, clamped, symmetric).  For a loss with gradient v = dL/du:
//   lambda = A^-1 v_h                         (the adjoint: a masked solve)
//   dL/df_h = -lambda_h,   dL/df_k = 0
//   dL/dg_k = v_k + sum of lambda over the hole 4-neighbours of k
//   dL/dg_h = 0 at holes (NaN inputs)
namespace spk {

// adjoint system of one plane: boundary 0 at known pixels, data -v at holes
__global__ __launch_bounds__(EW_NT) void k_adj_prepare(const AdjPlane *ap,
		const BlockWork *bw)
{
	const BlockWork b = bw[blockIdx.x];
	const AdjPlane A = ap[b.plane];
	const size_t n = (size_t)A.w * A.h;
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		const bool hole = isnan(A.g[q]);
		A.g2[q] = hole ? __uint_as_float(SP_QNAN_BITS) : 0.0f;
		A.f2[q] = hole ? -A.v[q] : 0.0f;
	}
}

__global__ __launch_bounds__(EW_NT) void k_adj_gather(const AdjPlane *ap,
		const BlockWork *bw)
{
	const BlockWork b = bw[blockIdx.x];
	const AdjPlane A = ap[b.plane];
	const size_t n = (size_t)A.w * A.h;
	for (int k = 0; k < EW_PPT; k++) {
		const size_t q = (size_t)b.start + k * EW_NT + threadIdx.x;
		if (q >= n) break;
		const int i = (int)(q % A.w), j = (int)(q / A.w);
		if (isnan(A.g[q])) {
			A.grad_g[q] = 0.0f;
			if (A.grad_f) A.grad_f[q] = -A.lam[q];
			continue;
		}
		float s = A.v[q];
		if (i + 1 < A.w && isnan(A.g[q + 1])) s += A.lam[q + 1];
		if (j + 1 < A.h && isnan(A.g[q + A.w])) s += A.lam[q + A.w];
		if (i > 0 && isnan(A.g[q - 1])) s += A.lam[q - 1];
		if (j > 0 && isnan(A.g[q - A.w])) s += A.lam[q - A.w];
		A.grad_g[q] = s;
		if (A.grad_f) A.grad_f[q] = 0.0f;
	}
}

void launch_adj_prepare(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s)
{
	if (nblocks) k_adj_prepare<<<nblocks, EW_NT, 0, s>>>(ap, bw);
}

void launch_adj_gather(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s)
{
	if (nblocks) k_adj_gather<<<nblocks, EW_NT, 0, s>>>(ap, bw);
}

} // namespace spk

// ------------------------------------------- adjoint of the finite sweeps
//
// Gradient of simpois's own output (simpois.c:247-285 with -c 0: niter
// Gauss-Seidel sweeps at every pyramid level); the host driver (sp_host.hip,
// adjoint_sweeps) explains the algebra.  Data are finite or NaN (holes).
// Each launch covers the planes of one image at one level: plane
// blockIdx.y, its arrays at offset blockIdx.y * (pixels of the level); one
// pixel per thread, sums in a fixed order.
namespace spk {

constexpr int AS_NT = 256;

static __device__ __forceinline__ bool as_hole(const unsigned char *hole, int w, int h,
		int i, int j)
{
	return i >= 0 && j >= 0 && i < w && j < h && hole[(size_t)j * w + i];
}

// pixel index of this thread in its plane, or -1; *pl = plane offset
static __device__ __forceinline__ long as_pix(size_t n, size_t *pl)
{
	*pl = (size_t)blockIdx.y * n;
	const size_t q = (size_t)blockIdx.x * AS_NT + threadIdx.x;
	return q < n ? (long)q : -1;
}

// holes of the finest level from g
__global__ __launch_bounds__(AS_NT) void k_as_mask0(const float *g, size_t n,
		unsigned char *hole)
{
	size_t o;
	const long q = as_pix(n, &o);
	if (q >= 0) hole[o + q] = isnan(g[o + q]);
}

// holes of the next coarser level: a coarse pixel is NaN when none of its
// four clamped samples is finite (simpois.c:172-196)
__global__ __launch_bounds__(AS_NT) void k_as_mask(const unsigned char *hf, int wf,
		int hf_, unsigned char *hc, int wc, int hc_)
{
	const size_t nc = (size_t)wc * hc_, nf = (size_t)wf * hf_;
	size_t oc;
	const long q = as_pix(nc, &oc);
	if (q < 0) return;
	const unsigned char *F = hf + (size_t)blockIdx.y * nf;
	const int I = (int)(q % wc), J = (int)(q / wc);
	const int i0 = min(2 * I, wf - 1), i1 = min(2 * I + 1, wf - 1);
	const int j0 = min(2 * J, hf_ - 1), j1 = min(2 * J + 1, hf_ - 1);
	hc[oc + q] = F[(size_t)j0 * wf + i0] && F[(size_t)j0 * wf + i1] &&
		F[(size_t)j1 * wf + i0] && F[(size_t)j1 * wf + i1];
}

// Level start.  The level's output is g at its known pixels and u at its
// holes (cleant_cgpois.c:251-256), so dL/dg gets v at the known pixels and
// the sweeps' adjoint starts from v at the holes.  The sweep problem is set
// up on the image turned by 180 degrees: boundary 0 at the known pixels,
// NaN at the holes, data -v at the holes.
__global__ __launch_bounds__(AS_NT) void k_as_setup(const unsigned char *hole,
		const float *v, float *gg, float *g_rot, float *f_rot, int w, int h)
{
	const size_t n = (size_t)w * h;
	size_t o;
	const long q = as_pix(n, &o);
	if (q < 0) return;
	const size_t r = o + (n - 1 - q);   // the pixel turned by 180 degrees
	if (hole[o + q]) {
		gg[o + q] = 0.0f;
		g_rot[r] = __uint_as_float(SP_QNAN_BITS);
		f_rot[r] = -v[o + q];
	} else {
		gg[o + q] = v[o + q];
		g_rot[r] = 0.0f;
		f_rot[r] = 0.0f;
	}
}

__global__ __launch_bounds__(AS_NT) void k_as_fill(float *a, size_t n, float v)
{
	size_t o;
	const long q = as_pix(n, &o);
	if (q >= 0) a[o + q] = v;
}

// acc = u (first) or acc + u
__global__ __launch_bounds__(AS_NT) void k_as_acc(float *acc, const float *u, size_t n,
		int first)
{
	size_t o;
	const long q = as_pix(n, &o);
	if (q >= 0) acc[o + q] = first ? u[o + q] : acc[o + q] + u[o + q];
}

// Level end, from the turned buffers nu0 (the sweeps' result) and acc (the
// sum of the intermediate ones): at a hole, dL/d(init) = (I/t + U) nu0 (U:
// the hole couplings to the east and south, later in raster order) and
// dL/df = -acc; at a known pixel, dL/dg += acc over its hole neighbours
// (the sweeps read it as a neighbour value).
__global__ __launch_bounds__(AS_NT) void k_as_finish(const unsigned char *hole,
		const float *nu0, const float *acc, int w, int h, float t, float *gi,
		float *gf, float *gg)
{
	const size_t n = (size_t)w * h;
	size_t o;
	const long q = as_pix(n, &o);
	if (q < 0) return;
	const unsigned char *H = hole + o;
	const float *N0 = nu0 + o, *AC = acc + o;
	const int i = (int)(q % w), j = (int)(q / w);
	auto rot = [&](int a, int b) { return n - 1 - ((size_t)b * w + a); };
	if (H[q]) {
		float s = N0[rot(i, j)] / t;
		if (as_hole(H, w, h, i + 1, j)) s -= N0[rot(i + 1, j)];
		if (as_hole(H, w, h, i, j + 1)) s -= N0[rot(i, j + 1)];
		gi[o + q] = s;
		gf[o + q] = -AC[rot(i, j)];
	} else {
		float s = gg[o + q];
		if (as_hole(H, w, h, i + 1, j)) s += AC[rot(i + 1, j)];
		if (as_hole(H, w, h, i, j + 1)) s += AC[rot(i, j + 1)];
		if (as_hole(H, w, h, i - 1, j)) s += AC[rot(i - 1, j)];
		if (as_hole(H, w, h, i, j - 1)) s += AC[rot(i, j - 1)];
		gg[o + q] = s;
		gi[o + q] = 0.0f;
		gf[o + q] = 0.0f;
	}
}

// a level without sweeps: the output at the holes is the initial value
__global__ __launch_bounds__(AS_NT) void k_as_nosweep(const unsigned char *hole,
		const float *v, float *gg, float *gi, float *gf, int w, int h)
{
	size_t o;
	const long q = as_pix((size_t)w * h, &o);
	if (q < 0) return;
	const bool hl = hole[o + q];
	gg[o + q] = hl ? 0.0f : v[o + q];
	gi[o + q] = hl ? v[o + q] : 0.0f;
	gf[o + q] = 0.0f;
}

// bilinear x2 of simpois.c:199-239 at fine coordinate i: corners floor(x),
// floor(x) + 1 (clamped to n) with weights 1 - (x - floor(x)) and
// x - floor(x), x = (i - 0.5) / 2
static __device__ __forceinline__ void as_bilin(int i, int n, int &a, int &b, float &wa,
		float &wb)
{
	const float x = (float)((i - 0.5) / 2);
	const int ip = (int)floorf(x);
	const float px = x - (float)ip;
	a = clampi(ip, 0, n - 1);
	b = clampi(ip + 1, 0, n - 1);
	wa = 1 - px;
	wb = px;
}

// v of the coarse level = the transpose of the prolongation applied to
// dL/d(init) at the fine holes (the initial value is read at holes only)
__global__ __launch_bounds__(AS_NT) void k_as_zoom_in_t(const float *gi,
		const unsigned char *hole, int wf, int hf, float *vc, int wc, int hc)
{
	size_t oc;
	const long q = as_pix((size_t)wc * hc, &oc);
	if (q < 0) return;
	const size_t of = (size_t)blockIdx.y * wf * hf;
	const int I = (int)(q % wc), J = (int)(q / wc);
	float s = 0.0f;
	for (int j = max(2 * J - 1, 0); j <= min(2 * J + 2, hf - 1); j++) {
		int ja, jb;
		float wja, wjb;
		as_bilin(j, hc, ja, jb, wja, wjb);
		const float wy = (ja == J ? wja : 0.0f) + (jb == J ? wjb : 0.0f);
		if (wy == 0.0f) continue;
		for (int i = max(2 * I - 1, 0); i <= min(2 * I + 2, wf - 1); i++) {
			const size_t p = of + (size_t)j * wf + i;
			if (!hole[p]) continue;
			int ia, ib;
			float wia, wib;
			as_bilin(i, wc, ia, ib, wia, wib);
			const float wx = (ia == I ? wia : 0.0f) + (ib == I ? wib : 0.0f);
			s += gi[p] * wx * wy;
		}
	}
	vc[oc + q] = s;
}

// fine += scale * (transpose of the NaN-aware 2x2 mean) applied to the
// coarse gradient gc: a fine pixel is sampled only by its coarse pixel
// (twice per axis on the clamped last row / column of an odd size), which
// divides by its number of finite samples.  masked: g (only known pixels
// are finite samples); otherwise f (every sample finite).
__global__ __launch_bounds__(AS_NT) void k_as_zoom_out_t(const float *gc, int wc, int hc,
		const unsigned char *hole, int masked, int wf, int hf, float scale, float *fine)
{
	size_t of;
	const long q = as_pix((size_t)wf * hf, &of);
	if (q < 0) return;
	const unsigned char *H = hole + of;
	if (masked && H[q]) return;
	const int i = (int)(q % wf), j = (int)(q / wf), I = i >> 1, J = j >> 1;
	const int i0 = min(2 * I, wf - 1), i1 = min(2 * I + 1, wf - 1);
	const int j0 = min(2 * J, hf - 1), j1 = min(2 * J + 1, hf - 1);
	int cnt = 4;
	if (masked)
		cnt = !H[(size_t)j0 * wf + i0] + !H[(size_t)j0 * wf + i1] +
			!H[(size_t)j1 * wf + i0] + !H[(size_t)j1 * wf + i1];
	const int mult = ((i0 == i) + (i1 == i)) * ((j0 == j) + (j1 == j));
	fine[of + q] += scale * gc[(size_t)blockIdx.y * wc * hc + (size_t)J * wc + I] *
		(float)mult / (float)cnt;
}

static dim3 as_grid(size_t n, int np) { return dim3((unsigned)((n + AS_NT - 1) / AS_NT), np); }

void as_mask0(const float *g, size_t n, int np, unsigned char *hole, hipStream_t s)
{
	k_as_mask0<<<as_grid(n, np), AS_NT, 0, s>>>(g, n, hole);
}
void as_mask(const unsigned char *hf, int wf, int hf_, unsigned char *hc, int wc, int hc_,
		int np, hipStream_t s)
{
	k_as_mask<<<as_grid((size_t)wc * hc_, np), AS_NT, 0, s>>>(hf, wf, hf_, hc, wc, hc_);
}
void as_setup(const unsigned char *hole, const float *v, float *gg, float *g_rot,
		float *f_rot, int w, int h, int np, hipStream_t s)
{
	k_as_setup<<<as_grid((size_t)w * h, np), AS_NT, 0, s>>>(hole, v, gg, g_rot, f_rot, w, h);
}
void as_fill(float *a, size_t n, int np, float v, hipStream_t s)
{
	k_as_fill<<<as_grid(n, np), AS_NT, 0, s>>>(a, n, v);
}
void as_acc(float *acc, const float *u, size_t n, int np, int first, hipStream_t s)
{
	k_as_acc<<<as_grid(n, np), AS_NT, 0, s>>>(acc, u, n, first);
}
void as_finish(const unsigned char *hole, const float *nu0, const float *acc, int w, int h,
		int np, float t, float *gi, float *gf, float *gg, hipStream_t s)
{
	k_as_finish<<<as_grid((size_t)w * h, np), AS_NT, 0, s>>>(hole, nu0, acc, w, h, t, gi,
			gf, gg);
}
void as_nosweep(const unsigned char *hole, const float *v, float *gg, float *gi, float *gf,
		int w, int h, int np, hipStream_t s)
{
	k_as_nosweep<<<as_grid((size_t)w * h, np), AS_NT, 0, s>>>(hole, v, gg, gi, gf, w, h);
}
void as_zoom_in_t(const float *gi, const unsigned char *hole, int wf, int hf, float *vc,
		int wc, int hc, int np, hipStream_t s)
{
	k_as_zoom_in_t<<<as_grid((size_t)wc * hc, np), AS_NT, 0, s>>>(gi, hole, wf, hf, vc, wc,
			hc);
}
void as_zoom_out_t(const float *gc, int wc, int hc, const unsigned char *hole, int masked,
		int wf, int hf, int np, float scale, float *fine, hipStream_t s)
{
	k_as_zoom_out_t<<<as_grid((size_t)wf * hf, np), AS_NT, 0, s>>>(gc, wc, hc, hole,
			masked, wf, hf, scale, fine);
}

} // namespace spk
