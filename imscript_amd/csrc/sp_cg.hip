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

namespace spk {

constexpr int NT = 256;
constexpr int PPT = 4;

static __device__ __forceinline__ int clampi(int v, int lo, int hi)
{
	return v < lo ? lo : (v > hi ? hi : v);
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

// pixel q of thread k-th visit inside work block b, or -1 past the plane
static __device__ __forceinline__ long pix(const BlockWork &b, int k, size_t n)
{
	const size_t q = (size_t)b.start + (size_t)k * NT + threadIdx.x;
	return q < n ? (long)q : -1;
}

// b, r0 = b - A x0 (x0 = u at holes), p = r, x = x0; partial r.r and b.b
__global__ __launch_bounds__(NT) void k32_setup(const BlockWork *bw,
		const Cg32Plane *pl, double *part)
{
	__shared__ double red[NT / 64];
	const BlockWork b = bw[blockIdx.x];
	const Cg32Plane P = pl[b.plane];
	double rr = 0, bb = 0;
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)P.w * P.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		const float gq = P.g[q];
		if (!isnan(gq)) {
			P.x[q] = 0; P.r[q] = 0; P.p[q] = 0; P.q[q] = 0;
			continue;
		}
		const int i = (int)(q % P.w), j = (int)(q / P.w);
		const float *g = P.g, *u = P.u;
		const float a0 = op32(P.bilap, i, j, P.w, P.h, [&](size_t k) {
			float v = g[k]; return isnan(v) ? 0.0f : v; });
		const float ax = op32(P.bilap, i, j, P.w, P.h, [&](size_t k) {
			return isnan(g[k]) ? u[k] : 0.0f; });
		const float bv = (P.f ? -P.f[q] : 0.0f) - a0;
		const float r = bv - ax;
		P.x[q] = (double)u[q];
		P.r[q] = r;
		P.p[q] = r;
		P.q[q] = 0;
		rr += (double)r * r;
		bb += (double)bv * bv;
	}
	const double s1 = bsum<NT>(rr, red), s2 = bsum<NT>(bb, red);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = s1;
		part[2 * blockIdx.x + 1] = s2;
	}
}

// one block per plane: fold the plane's slots in a fixed order
static __device__ void fold(const double *part, const Cg32Plane &P, int stride,
		int off, double *red, double *out)
{
	double v = 0;
	for (int k = P.blk0 + threadIdx.x; k < P.blk0 + P.nblk; k += NT)
		v += part[stride * k + off];
	const double s = bsum<NT>(v, red);
	if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(NT) void k32_reduce0(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part, int *nactive)
{
	__shared__ double red[NT / 64];
	__shared__ double rr, bb;
	const Cg32Plane P = pl[blockIdx.x];
	fold(part, P, 2, 0, red, &rr);
	fold(part, P, 2, 1, red, &bb);
	__syncthreads();
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		s.rr = rr;
		s.bnorm = sqrt(bb);
		s.stop = s.rel ? s.eps * s.bnorm : s.eps;
		s.iters = 0;
		const int n = *P.cnt;
		if (s.maxit_from_count) s.maxit = (int)(float)n;
		s.active = (s.maxit > 0 && n > 0) ? 1 : 0;
		if (s.active) atomicAdd(nactive, 1);
	}
}

// q = A p, partial p.q
__global__ __launch_bounds__(NT) void k32_spmv(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, double *part)
{
	__shared__ double red[NT / 64];
	const BlockWork b = bw[blockIdx.x];
	if (!sc[b.plane].active) return;
	const Cg32Plane P = pl[b.plane];
	const float *p = P.p;
	double acc = 0;
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)P.w * P.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		if (!isnan(P.g[q])) continue;
		const int i = (int)(q % P.w), j = (int)(q / P.w);
		const float y = op32(P.bilap, i, j, P.w, P.h, [&](size_t k) { return p[k]; });
		P.q[q] = y;
		acc += (double)y * p[q];
	}
	const double s = bsum<NT>(acc, red);
	if (threadIdx.x == 0) part[2 * blockIdx.x] = s;
}

__global__ __launch_bounds__(NT) void k32_reduce_a(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part)
{
	__shared__ double red[NT / 64];
	__shared__ double pq;
	if (!sc[blockIdx.x].active) return;
	fold(part, pl[blockIdx.x], 2, 0, red, &pq);
	if (threadIdx.x == 0) sc[blockIdx.x].alpha = sc[blockIdx.x].rr / pq;
}

// x += alpha p ; r -= alpha q ; partial r.r
__global__ __launch_bounds__(NT) void k32_update(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc, double *part)
{
	__shared__ double red[NT / 64];
	const BlockWork b = bw[blockIdx.x];
	const Cg32Scal &s = sc[b.plane];
	if (!s.active) return;
	const Cg32Plane P = pl[b.plane];
	const double ad = s.alpha;
	const float a = (float)ad;
	double acc = 0;
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)P.w * P.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		if (!isnan(P.g[q])) continue;
		P.x[q] = P.x[q] + ad * (double)P.p[q];
		const float r = P.r[q] - a * P.q[q];
		P.r[q] = r;
		acc += (double)r * r;
	}
	const double t = bsum<NT>(acc, red);
	if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = t;
}

// rr_new, absolute stop test on sqrt(rr_new), beta
__global__ __launch_bounds__(NT) void k32_reduce_b(const Cg32Plane *pl, Cg32Scal *sc,
		const double *part, int *nactive, double *trace)
{
	__shared__ double red[NT / 64];
	__shared__ double rr1;
	if (!sc[blockIdx.x].active) return;
	fold(part, pl[blockIdx.x], 2, 1, red, &rr1);
	if (threadIdx.x == 0) {
		Cg32Scal &s = sc[blockIdx.x];
		if (trace && s.iters < s.trace_cap) trace[s.trace0 + s.iters] = rr1;
		s.iters++;
		s.beta = rr1 / s.rr;
		s.rr = rr1;
		if (sqrt(rr1) < s.stop || s.iters >= s.maxit) {
			s.active = 0;
			atomicSub(nactive, 1);
		}
	}
}

// p = r + beta p
__global__ __launch_bounds__(NT) void k32_dir(const BlockWork *bw,
		const Cg32Plane *pl, const Cg32Scal *sc)
{
	const BlockWork b = bw[blockIdx.x];
	const Cg32Scal &s = sc[b.plane];
	if (!s.active) return;
	const Cg32Plane P = pl[b.plane];
	const float be = (float)s.beta;
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)P.w * P.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		if (!isnan(P.g[q])) continue;
		P.p[q] = P.r[q] + be * P.p[q];
	}
}

__global__ __launch_bounds__(NT) void k32_finish(const BlockWork *bw,
		const Cg32Plane *pl)
{
	const BlockWork b = bw[blockIdx.x];
	const Cg32Plane P = pl[b.plane];
	for (int k_ = 0; k_ < PPT; k_++) {
		const long qq = pix(b, k_, (size_t)P.w * P.h);
		if (qq < 0) break;
		const size_t q = (size_t)qq;
		if (isnan(P.g[q])) P.u[q] = (float)P.x[q];
	}
}

void cg32_setup(const Cg32Batch &B, hipStream_t s)
{
	k32_setup<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.part);
	k32_reduce0<<<B.nplanes, NT, 0, s>>>(B.planes, B.scal, B.part, B.nactive);
}

void cg32_iter(const Cg32Batch &B, hipStream_t s)
{
	k32_spmv<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part);
	k32_reduce_a<<<B.nplanes, NT, 0, s>>>(B.planes, B.scal, B.part);
	k32_update<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal, B.part);
	k32_reduce_b<<<B.nplanes, NT, 0, s>>>(B.planes, B.scal, B.part, B.nactive, B.trace);
	k32_dir<<<B.nblk, NT, 0, s>>>(B.bw, B.planes, B.scal);
}

void cg32_finish(const Cg32Batch &B, hipStream_t s)
{
	k32_finish<<<B.nblk, NT, 0, s>>>(B.bw, B.planes);
}

} // namespace spk
