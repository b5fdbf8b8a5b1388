// Internal interface between the host driver (sp_host.hip) and the device
// kernels (sp_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sp_gs.h"

// One plane at one pyramid level.
struct SpPlane {
	float *g;        // boundary data, NaN = hole
	float *f;        // data term or nullptr
	float *u;        // solution (in place)
	float *c;        // coefficient: data term at holes, SP_KNOWN_BITS elsewhere
	int *cnt;        // number of holes (device counter)
	int w, h;
};

// Work block for elementwise kernels over many planes of uneven size.
struct BlockWork { int plane; int start; };

// One plane of an adjoint (gradient) computation.
struct AdjPlane {
	const float *g;    // forward boundary data, NaN = hole
	const float *v;    // dL/du
	float *g2, *f2;    // adjoint system: 0 / NaN boundary, data -v
	float *lam;        // adjoint solution
	float *grad_g, *grad_f;   // outputs (grad_f may be null)
	int w, h;
};

namespace spk {

// Loads and stores through the global address space.  A pointer read from
// a descriptor in memory is generic, and a flat access also counts against
// the LDS counter: every wait for an LDS read then waits for it too.  The
// type must be a scalar or an ext_vector type (a struct loaded through a
// reference falls back to a flat access).
typedef float gf2 __attribute__((ext_vector_type(2)));
typedef unsigned gu2 __attribute__((ext_vector_type(2)));
typedef unsigned gu4 __attribute__((ext_vector_type(4)));
template <class T>
static __device__ __forceinline__ T gload(const void *p)
{
	return *(const __attribute__((address_space(1))) T *)p;
}
template <class T>
static __device__ __forceinline__ void gstore(void *p, T v)
{
	*(__attribute__((address_space(1))) T *)p = v;
}
// the same at a 32-bit byte offset from a base held in scalar registers:
// the loads' own address form, without 64-bit address arithmetic per lane
// (a plane holds at most 2^30 pixels, so its byte offsets fit)
template <class T>
static __device__ __forceinline__ T gload_b(const void *base, unsigned off)
{
	return *(const __attribute__((address_space(1))) T *)(
			(const __attribute__((address_space(1))) char *)base + off);
}
template <class T>
static __device__ __forceinline__ void gstore_b(void *base, unsigned off, T v)
{
	*(__attribute__((address_space(1))) T *)((__attribute__((address_space(1))) char *)base
			+ off) = v;
}

// tile_load's footprint copy with a thread's loads issued in batches of up
// to 8 cells before the batch's first LDS write: one at a time, each LDS
// write waited for its own load.  Every load is of an address clamped into
// the image (no branch around it: a value merged after a branch made the
// compiler wait there), and a cell outside the image then takes, for R = 1,
// the NaN value and the known marker by a select.  The same cells get the
// same values as in tile_load.
//
// COH: the values (not the coefficients) are read with agent-scope
// coherent loads, which see stores other XCDs made with agent-scope
// coherent stores without any cache invalidation.  PART: 0 values and
// coefficients, 1 the coefficients only, 2 the values only.
template <class S, bool COH = false, int PART = 0>
static __device__ __forceinline__ void tile_load_dev(const TileCtx<S> &c, const GsInst &in)
{
	using G = Geo<S>;
	constexpr int R = S::R;
	constexpr int N = G::LW * G::LH;
	constexpr int IT = (N + G::NT - 1) / G::NT;
	constexpr int B = IT < 8 ? IT : 8;
	const int tid = (int)threadIdx.x;
	for (int e0 = 0; e0 < IT; e0 += B) {
		float uv[B], cv[B];
		bool ok[B];
#pragma unroll
		for (int u = 0; u < B; u++) {
			const int e = tid + (e0 + u) * G::NT;
			const int row = e / G::LW, col = e - row * G::LW;
			const int j = row + c.jlo, i = col + c.K - G::SH * j;
			ok[u] = e < N && i >= 0 && i < c.W && j >= 0 && j < c.H;
			const size_t q = (size_t)sp_clamp(j, 0, c.H - 1) * c.W + sp_clamp(i, 0, c.W - 1);
			if (PART != 1) {
				if (COH)
					uv[u] = __hip_atomic_load((__attribute__((address_space(1))) float *)
							(in.u + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				else
					uv[u] = gload<float>(in.u + q);
			}
			if (PART != 2) cv[u] = gload<float>(in.c + q);
		}
#pragma unroll
		for (int u = 0; u < B; u++) {
			const int e = tid + (e0 + u) * G::NT;
			if (e >= N) continue;
			if (ok[u]) {
				if (PART != 1) c.U[e] = uv[u];
				if (PART != 2) c.C[e] = cv[u];
			} else if (R == 1) {
				if (PART != 1) c.U[e] = sp_float(SP_QNAN_BITS);
				if (PART != 2) c.C[e] = sp_float(SP_KNOWN_BITS);
			}
		}
	}
}

// tile_store with agent-scope coherent stores (the values a later tile of
// the same launch reads with coherent loads)
template <class S>
static __device__ __forceinline__ void tile_store_coh(const TileCtx<S> &c, const GsInst &in)
{
	using G = Geo<S>;
	for (int e = threadIdx.x; e < G::LW * G::LH; e += G::NT) {
		const int row = e / G::LW, col = e - row * G::LW;
		const int j = row + c.jlo, i = col + c.K - G::SH * j;
		if (i < 0 || i >= c.W || j < 0 || j >= c.H || sp_known(c.C[e])) continue;
		if (tile_owns<S>(c, i, j))
			__hip_atomic_store((__attribute__((address_space(1))) float *)(in.u
					+ (size_t)j * c.W + i), c.U[e], __ATOMIC_RELAXED,
					__HIP_MEMORY_SCOPE_AGENT);
	}
}

// The skewed tiles of a level (or of a pass of AMLE sweeps) in one launch,
// ordered by their dependences instead of by one launch per wavefront.
// Work-groups claim tiles in wavefront order from *claim; a tile starts once
// each existing predecessor (I'-a, J'-b, T'-c), a, b, c in {0, 1} not all 0,
// has published its stores.  Every dependence between updates (true, anti
// and output) has a skewed distance in {0, 1}^3, so these are the only
// tiles whose order matters.  The values of the image move between tiles
// only through agent-scope coherent loads and stores (coherent across the
// XCDs' L2 caches without writing back or invalidating them); a tile
// publishes (flag = epoch, a coherent store) once its stores have completed,
// and a waiting tile polls the flags with coherent loads.  The predecessors
// of a claimed tile were claimed before it by running work-groups, so every
// wait ends; a wait past spin_max polls sets *err and lets the tile run (the
// host reports the error).
//
// Before the wait the tile's coefficients go into LDS (they do not change
// during the launch); body(tile, instance) then loads the values with
// coherent loads, runs the sweeps and stores with tile_store_coh.
template <class S, class Body>
static __device__ __forceinline__ void tile_flow_loop(const GsTile *tiles, int ntiles,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, float *smem, Body body)
{
	__shared__ int s_k;
	for (;;) {
		__syncthreads();   // the previous tile's LDS and s_k are free
		if (threadIdx.x == 0) s_k = (int)min(atomicAdd(claim, 1u), (unsigned)ntiles);
		__syncthreads();
		const int k = s_k;
		if (k >= ntiles) return;
		const GsTile tl = tiles[k];
		const GsFlowInst f = fi[tl.inst];
		const int d = f.base + (tl.tt * f.nJ + tl.tj) * f.nI + tl.ti;
		const GsInst in = insts[tl.inst];
		const bool run = !in.active || *in.active;
		if (run) tile_load_dev<S, false, 1>(tile_ctx<S>(tl, in, smem), in);
		if (threadIdx.x < 7) {
			const int a = (threadIdx.x + 1) & 1, b = (threadIdx.x + 1) >> 1 & 1;
			const int cc = (threadIdx.x + 1) >> 2;
			if (tl.ti >= a && tl.tj >= b && tl.tt >= cc) {
				int *p = flags + d - a - b * f.nI - cc * f.nI * f.nJ;
				int n = 0;
				while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
						< epoch) {
					if (++n > spin_max) {
						atomicOr(err, 1);
						break;
					}
					__builtin_amdgcn_s_sleep(2);
				}
			}
		}
		__syncthreads();
		if (run) body(tl, in);
		__builtin_amdgcn_s_waitcnt(0);   // this wave's stores have completed
		__syncthreads();
		if (threadIdx.x == 0)
			__hip_atomic_store(flags + d, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
}

// elementwise / pyramid
// prepare / restrict write the holes of each work block to bcnt;
// launch_count folds them into each member's counter
void launch_prepare(const SpPlane *planes, const BlockWork *bw, int nblocks,
		int *bcnt, hipStream_t s);
void launch_restrict(const SpPlane *fine, const SpPlane *coarse,
		const BlockWork *bw, int nblocks, double pmsfac, int *bcnt, hipStream_t s);
void launch_count(const SpPlane *planes, int nplanes, const int *first,
		const int *bcnt, hipStream_t s);
void launch_init(const SpPlane *lv, const SpPlane *coarse, const BlockWork *bw,
		int nblocks, hipStream_t s);
void launch_finalize(const SpPlane *lv, float *const *outs,
		const BlockWork *bw, int nblocks, hipStream_t s);

// Gauss-Seidel
// plambda "x,l" (variant 0) / "x,lr" (1) of pd planar w x h planes under
// boundary rule mode (0 zero, 1 nearest, 2 reflection, 3 periodic, 4 NaN)
void launch_plambda_lap(float *out, const float *in, int w, int h, int pd,
		int variant, int mode, hipStream_t s);

// AMLE (sp_amle.h): bounds of each plane into mm and the initial u / c
// (and tinf when non-null); part is scratch of amle_part_bytes
size_t amle_part_bytes(int n, int pd);
int amle_part_blocks(int n);
void launch_amle_prep(const float *in, const float *init, float *u, float *c,
		float *tinf, void *part, float2 *mm, int n, int pd, hipStream_t s);
// one launch of the AMLE tiles of one wavefront (shape variant var of radius R)
int amle_tile_lds_bytes(int R, int var);
// fast4: nn = 4 on planes larger than 1 x 1 (k_amle4_tile)
void launch_amle_tiles(int R, int var, const GsTile *tiles, int ntiles,
		const GsInst *insts, int nn, bool fast4, hipStream_t s);
// the tiles of a pass in one launch ordered by dependences (as
// launch_gs_flow; the claim counter is reset first)
int amle_flow_blocks(int R, int var, bool fast4);
void launch_amle_flow(int R, int var, const GsTile *tiles, int ntiles, int nblocks,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, int nn, bool fast4, hipStream_t s);
// per-block partials (holes, max, sums) of the trace diagnostics: 5 doubles
// per block, amle_part_blocks(n) blocks per plane
void launch_amle_diag(const float *c, const float *tinf, const float *tsup,
		const float *tinf0, const float *tsup0, int n, int pd, double *part,
		hipStream_t s);

// one explicit diffusion step of heatd (A: w*h*3 interleaved field or null)
void launch_heatd_step(float *out, const float *in, const float *A, int w, int h,
		int pd, float t, hipStream_t s);
// T isotropic steps (1..heatd_multi_max()) in one launch on register tiles
int heatd_multi_max();
// planes large enough for T steps on tiles of `rows` rows (48 isotropic)
bool heatd_multi_fits(int w, int h, int T, int rows);
// (false, nothing launched, when the planes are too small for T)
bool launch_heatd_multi(float *out, const float *in, int w, int h, int pd, int T,
		float t, hipStream_t s);
// T steps with the field A (w*h*3 interleaved) on tiles of rows = 16, 24,
// 32 or 48 rows (at least 3T - 1); false, nothing launched, for other rows
// or planes too small
bool launch_heatd_multi_field(float *out, const float *in, const float *A, int w, int h,
		int pd, int T, int rows, float t, hipStream_t s);

// dst[0..4*n4) = src[0..4*n4), 16-byte aligned
void launch_stream_copy(float *dst, const float *src, size_t n4, hipStream_t s);

// tiles of shape variant `var` of radius R (sp_gs.h); one launch per wavefront
int tile_lds_bytes(int R, int var);
int small_lds_bytes(int R, int w, int h);
// whole R = 1 levels by unknown pixel (k_gs_cells): allow `bytes` of LDS
// (0 on success), launch one work-group of `threads` per listed member
int gs_cells_enable(int bytes);
void launch_gs_cells(int maxs, const GsInst *insts, const int *ids, int nids,
		int lds_bytes, int threads, hipStream_t s);
void launch_gs_tiles(int R, int var, const GsTile *tiles, int ntiles,
		const GsInst *insts, hipStream_t s);
void launch_gs_small(int R, const GsInst *insts, const int *ids, int nids,
		int lds_bytes, int threads, hipStream_t s);
// the tiles of a level in one launch ordered by dependences (k_gs_flow):
// nblocks work-groups (the resident count, gs_flow_blocks) claim the tiles
// from *claim, which the launch first sets to 0; flags: one per cell of
// each member's tile grid (GsFlowInst), INT_MAX for grid cells without a
// tile, below epoch for a tile not yet run in this launch
int gs_flow_blocks(int R, int var);
void launch_flow_reset(unsigned *claim, hipStream_t s);
void launch_gs_flow(int R, int var, const GsTile *tiles, int ntiles, int nblocks,
		const GsInst *insts, const GsFlowInst *fi, unsigned *claim, int *flags,
		int epoch, int spin_max, int *err, hipStream_t s);

// CG (FP64 vectors).  Vectors are compact: entry k belongs to the k-th
// unknown in raster order, at pixel idx[k].
struct CgDev {
	const float *g, *f;   // plane data
	float *u;             // solution plane (read for x0, written at the end)
	double *x, *r, *p;    // nc entries each
	float *q;             // A p: the stencil sums in float, so q is a float
	                      // (the reference widens it to double)
	unsigned *idx;        // pixel of each unknown, written by the setup
	int *bcnt;            // unknowns per setup block, then their prefix sums (nb + 1)
	float *pf;            // full grid: p rounded to float at the unknowns (what
	                      // the stencil reads), 0 at known pixels
	double *part;         // per-block partial sums (max(2 * nblk, nb))
	double *rr;           // rr history, rr[0] = r0.r0, rr[k+1] = after iter k
	double *rrp;          // [k & 1] = r.r before iteration k (2 entries)
	int *done;            // [0] = stop flag (for the host), [1] = iterations
	                      // run, [2] = index of the iteration under way,
	                      // [3 + (k & 1)] = stop flag iteration k reads,
	                      // [5] = k_cg_persist's error word, [7] = set when
	                      // it gave the launch up at its start (8 entries)
	unsigned *bar;        // k_cg_persist's gate and barrier words (2304 entries)
	int w, h, bilap;
	int nblk;             // blocks of the iteration kernels
	int nb;               // setup blocks: cg_setup_blocks(w, h)
	int nc;               // unknowns (NaN pixels of g)
	double stop;          // absolute stop threshold on sqrt(rr)
};
// The kernels read the descriptor from device memory (dp), so captured
// iterations serve every level and plane: the host rewrites *dp between
// them.  An iteration launch's grid may exceed d.nblk (extra blocks return).
// Iteration k is launched with par = k & 1 (chunks of even length keep the
// parity of their nodes).
int cg_setup_blocks(int w, int h);
void launch_cg_setup(const CgDev *dp, int nb, hipStream_t s);
int cg_iter_span();   // unknowns per block of the iteration kernels
void launch_cg_iter(const CgDev *dp, int bilap, int grid, int par, hipStream_t s);
// all iterations in one launch of nblk work-groups (done[7] set: it found
// its work-groups not all resident and did nothing); capacity = work-groups
// that can be resident at once on the device
int cg_persist_capacity(int bilap);
// spin_max bounds each barrier wait, gate_spin the start gate's wait; lg:
// log2 of the barriers' arrival groups (0..5)
void launch_cg_persist(const CgDev *dp, int bilap, int nblk, int maxit, int spin_max,
		int gate_spin, int lg, hipStream_t s);
void launch_cg_finish(const CgDev *dp, int nblk, hipStream_t s);

// adjoint of the converged Poisson fill
void launch_adj_prepare(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s);
void launch_adj_gather(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s);
// the same for the biharmonic fill (sp_cg.hip): the transposed 13-point
// operator over lambda at the known pixels
void launch_adj_gather_bilap(const AdjPlane *ap, const BlockWork *bw, int nblocks,
		hipStream_t s);

// adjoint of the finite sweeps (sp_kernels.hip): np planes of one level,
// plane l's arrays at offset l * (pixels of the level)
void as_mask0(const float *g, size_t n, int np, unsigned char *hole, hipStream_t s);
void as_mask(const unsigned char *hf, int wf, int hf_, unsigned char *hc, int wc, int hc_,
		int np, hipStream_t s);
void as_setup(const unsigned char *hole, const float *v, float *gg, float *g_rot,
		float *f_rot, int w, int h, int np, hipStream_t s);
void as_fill(float *a, size_t n, int np, float v, hipStream_t s);
void as_acc(float *acc, const float *u, size_t n, int np, int first, hipStream_t s);
void as_finish(const unsigned char *hole, const float *nu0, const float *acc, int w, int h,
		int np, float t, float *gi, float *gf, float *gg, hipStream_t s);
void as_nosweep(const unsigned char *hole, const float *v, float *gg, float *gi, float *gf,
		int w, int h, int np, hipStream_t s);
void as_zoom_in_t(const float *gi, const unsigned char *hole, int wf, int hf, float *vc,
		int wc, int hc, int np, hipStream_t s);
void as_zoom_out_t(const float *gc, int wc, int hc, const unsigned char *hole, int masked,
		int wf, int hf, int np, float scale, float *fine, hipStream_t s);

// FP64 relative residual of the Poisson (bilap=0) or biharmonic system
void launch_residual(const SpPlane *lv, const BlockWork *bw, int nblocks,
		int bilap, double *part, hipStream_t s);
void launch_residual_reduce(const double *part, const int *plane_first,
		int nplanes, double *out, hipStream_t s);
// lv[pd], bw[pd * per] (per blocks of 1024 pixels per plane) and first[pd + 1]
// of the residual of pd planar w x h planes, written by a kernel
void launch_residual_plan(SpPlane *lv, BlockWork *bw, int *first, const float *u,
		const float *g, const float *f, int w, int h, int pd, int per, hipStream_t s);

} // namespace spk
