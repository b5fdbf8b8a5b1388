// Device kernels of the AMLE hole filling (amle.c): initial bounds, the
// tiled raster-order sweeps (sp_amle.h) and the diagnostics of the trace.
#include "sp_internal.h"
#include "sp_amle.h"

#include <climits>

namespace spk {

constexpr int AM_T = 256;

// first strict minimum / maximum in raster order: the smaller value wins,
// equal values (only +0 / -0 differ in bits) keep the smaller index
struct AmleMM { float mn; int imn; float mx; int imx; };

__device__ __forceinline__ void mm_min(float &m, int &im, float v, int iv)
{
	if (v < m || (v == m && iv < im)) { m = v; im = iv; }
}
__device__ __forceinline__ void mm_max(float &m, int &im, float v, int iv)
{
	if (v > m || (v == m && iv < im)) { m = v; im = iv; }
}

__device__ void mm_block(AmleMM &r)
{
	__shared__ AmleMM sm[AM_T];
	sm[threadIdx.x] = r;
	__syncthreads();
	for (int s = AM_T / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s) {
			AmleMM o = sm[threadIdx.x + s], m = sm[threadIdx.x];
			mm_min(m.mn, m.imn, o.mn, o.imn);
			mm_max(m.mx, m.imx, o.mx, o.imx);
			sm[threadIdx.x] = m;
		}
		__syncthreads();
	}
	r = sm[0];
}

// amle.c:42-55 (get_minmax over the whole plane; NaN never compares):
// block partials over chunk blockIdx.x of plane blockIdx.y
__global__ __launch_bounds__(AM_T) void k_amle_minmax(const float *in, int n,
		int chunk, AmleMM *part)
{
	const float *x = in + (size_t)n * blockIdx.y;
	const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
	AmleMM r{INFINITY, INT_MAX, -INFINITY, INT_MAX};
	for (int q = lo + threadIdx.x; q < hi; q += AM_T) {
		const float v = x[q];
		mm_min(r.mn, r.imn, v, q);
		mm_max(r.mx, r.imx, v, q);
	}
	mm_block(r);
	if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
}

__global__ __launch_bounds__(AM_T) void k_amle_minmax_final(const AmleMM *part,
		int nb, float2 *mm)
{
	AmleMM r{INFINITY, INT_MAX, -INFINITY, INT_MAX};
	for (int b = threadIdx.x; b < nb; b += AM_T) {
		const AmleMM o = part[(size_t)blockIdx.x * nb + b];
		mm_min(r.mn, r.imn, o.mn, o.imn);
		mm_max(r.mx, r.imx, o.mx, o.imx);
	}
	mm_block(r);
	if (threadIdx.x == 0) mm[blockIdx.x] = make_float2(r.mn, r.mx);
}

// amle.c:90-110: tsup (u) starts at the plane's maximum on the holes, or at
// the initial image there; tinf at the minimum; c marks the known pixels
__global__ __launch_bounds__(AM_T) void k_amle_prep(const float *in,
		const float *init, const float2 *mm, float *u, float *c, float *tinf,
		int n)
{
	const int l = blockIdx.y;
	const size_t o = (size_t)n * l;
	const float2 b = mm[l];
	for (int q = blockIdx.x * AM_T + threadIdx.x; q < n; q += gridDim.x * AM_T) {
		const float x = in[o + q];
		const bool hole = isnan(x);
		u[o + q] = hole ? (init ? init[o + q] : b.y) : x;
		c[o + q] = hole ? 0.0f : sp_float(SP_KNOWN_BITS);
		if (tinf) tinf[o + q] = hole ? b.x : x;
	}
}

template <class S>
__global__ __launch_bounds__(Geo<S>::NT) void k_amle_tile(const GsTile *tiles,
		const GsInst *insts, int nn)
{
	using G = Geo<S>;
	extern __shared__ float smem[];
	const GsTile tl = tiles[blockIdx.x];
	const GsInst in = insts[tl.inst];
	const TileCtx<S> c = tile_ctx<S>(tl, in, smem);
	tile_load_dev<S>(c, in);
	__syncthreads();
	for (int sg = 0; sg < G::NSTEP; sg++) {
		amle_tile_step<S>(c, threadIdx.x, sg, nn);
		__syncthreads();
	}
	tile_store<S>(c, in, threadIdx.x);
}

// nn = 4 on planes larger than 1 x 1: per-thread state across the steps
template <class S>
__global__ __launch_bounds__(Geo<S>::NT) void k_amle4_tile(const GsTile *tiles,
		const GsInst *insts)
{
	using G = Geo<S>;
	extern __shared__ float smem[];
	const GsTile tl = tiles[blockIdx.x];
	const GsInst in = insts[tl.inst];
	const TileCtx<S> c = tile_ctx<S>(tl, in, smem);
	tile_load_dev<S>(c, in);
	Amle4Thread t = amle4_thread<S>(c, threadIdx.x);
	__syncthreads();
	for (int sg = 0; sg < G::NSTEP; sg++) {
		amle4_step<S>(c, t, sg);
		__syncthreads();
	}
	tile_store<S>(c, in, threadIdx.x);
}

// the tiles of a pass in one launch ordered by their dependences
// (tile_flow_loop), with the update of k_amle4_tile (FAST4) or k_amle_tile
template <class S, bool FAST4>
__global__ __launch_bounds__(Geo<S>::NT) void k_amle_flow(const GsTile *tiles, int ntiles,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, int nn)
{
	using G = Geo<S>;
	extern __shared__ float smem[];
	tile_flow_loop<S>(tiles, ntiles, insts, fi, claim, flags, epoch, spin_max, err, smem,
		[&](const GsTile &tl, const GsInst &in) {
			const TileCtx<S> c = tile_ctx<S>(tl, in, smem);
			tile_load_dev<S, true, 2>(c, in);
			if constexpr (FAST4 && S::R == 1) {
				Amle4Thread t = amle4_thread<S>(c, threadIdx.x);
				__syncthreads();
				for (int sg = 0; sg < G::NSTEP; sg++) {
					amle4_step<S>(c, t, sg);
					__syncthreads();
				}
			} else {
				__syncthreads();
				for (int sg = 0; sg < G::NSTEP; sg++) {
					amle_tile_step<S>(c, threadIdx.x, sg, nn);
					__syncthreads();
				}
			}
			tile_store_coh<S>(c, in);
		});
}

// trace diagnostics over the holes of plane blockIdx.y, block partials:
// holes, max |tinf - tsup| (NaN never wins, amle.c:120-121), sum of it,
// sums of |old - new| of tinf and of tsup over the last sweep
__global__ __launch_bounds__(AM_T) void k_amle_diag(const float *c,
		const float *tinf, const float *tsup, const float *tinf0,
		const float *tsup0, int n, int chunk, double *part)
{
	const size_t o = (size_t)n * blockIdx.y;
	const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
	double cnt = 0, e = 0, ea = 0, ai = 0, as = 0;
	for (int q = lo + threadIdx.x; q < hi; q += AM_T) {
		if (sp_known(c[o + q])) continue;
		const float t = fabsf(tinf[o + q] - tsup[o + q]);
		cnt += 1;
		if (t > e) e = t;
		ea += t;
		ai += fabsf(tinf0[o + q] - tinf[o + q]);
		as += fabsf(tsup0[o + q] - tsup[o + q]);
	}
	__shared__ double sm[5][AM_T];
	sm[0][threadIdx.x] = cnt;
	sm[1][threadIdx.x] = e;
	sm[2][threadIdx.x] = ea;
	sm[3][threadIdx.x] = ai;
	sm[4][threadIdx.x] = as;
	__syncthreads();
	for (int s = AM_T / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s) {
			sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
			const double m = sm[1][threadIdx.x + s];
			if (m > sm[1][threadIdx.x]) sm[1][threadIdx.x] = m;
			sm[2][threadIdx.x] += sm[2][threadIdx.x + s];
			sm[3][threadIdx.x] += sm[3][threadIdx.x + s];
			sm[4][threadIdx.x] += sm[4][threadIdx.x + s];
		}
		__syncthreads();
	}
	if (threadIdx.x < 5)
		part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 5 + threadIdx.x] =
			sm[threadIdx.x][0];
}

static int am_chunks(int n, int *chunk)
{
	// blocks of 16K pixels (64 per thread), at most 2048 per plane
	int nb = (n + 16383) / 16384;
	nb = nb < 1 ? 1 : nb > 2048 ? 2048 : nb;
	*chunk = (n + nb - 1) / nb;
	return (n + *chunk - 1) / *chunk;
}

int amle_part_blocks(int n)
{
	int chunk;
	return am_chunks(n, &chunk);
}

void launch_amle_prep(const float *in, const float *init, float *u, float *c,
		float *tinf, void *part, float2 *mm, int n, int pd, hipStream_t s)
{
	int chunk;
	const int nb = am_chunks(n, &chunk);
	k_amle_minmax<<<dim3(nb, pd), AM_T, 0, s>>>(in, n, chunk, (AmleMM *)part);
	k_amle_minmax_final<<<pd, AM_T, 0, s>>>((const AmleMM *)part, nb, mm);
	int gx = (n + AM_T * 4 - 1) / (AM_T * 4);
	gx = gx > 4096 ? 4096 : gx;
	k_amle_prep<<<dim3(gx, pd), AM_T, 0, s>>>(in, init, mm, u, c, tinf, n);
}

size_t amle_part_bytes(int n, int pd)
{
	const size_t nb = (size_t)amle_part_blocks(n);
	const size_t a = nb * pd * sizeof(AmleMM), b = nb * pd * 5 * sizeof(double);
	return a > b ? a : b;
}

void launch_amle_diag(const float *c, const float *tinf, const float *tsup,
		const float *tinf0, const float *tsup0, int n, int pd, double *part,
		hipStream_t s)
{
	int chunk;
	const int nb = am_chunks(n, &chunk);
	k_amle_diag<<<dim3(nb, pd), AM_T, 0, s>>>(c, tinf, tsup, tinf0, tsup0, n,
			chunk, part);
}

int amle_tile_lds_bytes(int R, int var)
{
	int b = 0;
	amle_with_shape(R, var, [&](auto sh) { b = Geo<decltype(sh)>::LDS_FLOATS * 4; });
	return b;
}

void launch_amle_tiles(int R, int var, const GsTile *tiles, int ntiles,
		const GsInst *insts, int nn, bool fast4, hipStream_t s)
{
	if (!ntiles) return;
	amle_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		if constexpr (S::R == 1) {
			if (fast4) {
				k_amle4_tile<S><<<ntiles, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(
						tiles, insts);
				return;
			}
		}
		k_amle_tile<S><<<ntiles, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(tiles,
				insts, nn);
	});
}

int amle_flow_blocks(int R, int var, bool fast4)
{
	int nb = 0;
	amle_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		int per = 0;
		const hipError_t e = fast4
			? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_amle_flow<S, true>,
					Geo<S>::NT, Geo<S>::LDS_FLOATS * 4)
			: hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_amle_flow<S, false>,
					Geo<S>::NT, Geo<S>::LDS_FLOATS * 4);
		if (e != hipSuccess || per < 1) per = 1;
		int dev = 0, ncu = 0;
		(void)hipGetDevice(&dev);
		if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)
				!= hipSuccess || ncu <= 0)
			ncu = 256;
		nb = per * ncu;
	});
	return nb;
}

void launch_amle_flow(int R, int var, const GsTile *tiles, int ntiles, int nblocks,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, int nn, bool fast4, hipStream_t s)
{
	if (!ntiles) return;
	launch_flow_reset(claim, s);
	amle_with_shape(R, var, [&](auto sh) {
		using S = decltype(sh);
		if (fast4)
			k_amle_flow<S, true><<<nblocks, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(
					tiles, ntiles, insts, fi, claim, flags, epoch, spin_max, err, nn);
		else
			k_amle_flow<S, false><<<nblocks, Geo<S>::NT, Geo<S>::LDS_FLOATS * 4, s>>>(
					tiles, ntiles, insts, fi, claim, flags, epoch, spin_max, err, nn);
	});
}

} // namespace spk
