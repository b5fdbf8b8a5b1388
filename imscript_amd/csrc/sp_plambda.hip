// plambda's 3x3 Laplacian image operators, the producer of simpois's
// guidance field (plambda.c:2589-2605 stencils, :2680-2689 application,
// getpixel.c boundary rules).  Planar float planes, bit-identical to the
// reference: every output is r = 0; r += s[k] * P(i-1+k%3, j-1+k/3) for
// k = 0..8 in that order, zero weights included (0 * NaN = NaN, 0 * -x =
// -0 as there), float multiply and add without contraction.
#include <stdint.h>
#include "sp_internal.h"

namespace spk {

constexpr int LP_T = 256, LP_PX = 4;   // threads per block, pixels per thread

// stencils "l" and "lr" (plambda.c:2589-2605)
template <int V> struct LapStencil;
template <> struct LapStencil<0> {
	static constexpr float s[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
};
template <> struct LapStencil<1> {
	static constexpr float s[9] = {0.5f, 0, 0.5f, 0, -2, 0, 0.5f, 0, 0.5f};
};

static __device__ __forceinline__ int lp_gmod(int x, int m)
{
	const int r = x % m;
	return r < 0 ? r + m : r;
}

static __device__ __forceinline__ int lp_reflex(int n, int p)
{   // getpixel.c:74-83
	int r = lp_gmod(n, 2 * p);
	if (r == p) r -= 1;
	if (r > p) r = 2 * p - r;
	return r;
}

// one sample under boundary rule `mode` (0 zero, 1 nearest, 2 reflection,
// 3 periodic, 4 NaN; getpixel.c:10-99)
static __device__ __forceinline__ float lp_sample(const float *x, int w, int h,
		int i, int j, int mode)
{
	const bool in = i >= 0 && i < w && j >= 0 && j < h;
	if (in) return x[(size_t)j * w + i];
	switch (mode) {
	case 0: return 0.0f;
	case 2: i = lp_reflex(i, w); j = lp_reflex(j, h); break;
	case 3: i = lp_gmod(i, w); j = lp_gmod(j, h); break;
	case 4: return __builtin_nanf("");
	default:
		i = i < 0 ? 0 : (i >= w ? w - 1 : i);
		j = j < 0 ? 0 : (j >= h ? h - 1 : j);
	}
	return x[(size_t)j * w + i];
}

// Grid: x = strips of LP_RS rows of all planes, y = column chunks of
// LP_T*LP_PX.  A thread walks its strip down with a rolling window of rows:
// each step loads the next LP_U rows (a 16-byte load and a scalar on each
// side when its columns are inside the plane and the planes are 16-byte
// aligned; the boundary rule otherwise), so LP_U rows of loads are in
// flight at once, then writes LP_U output rows.
constexpr int LP_RS = 32, LP_U = 2;

template <bool VEC>
static __device__ __forceinline__ void lp_row(float (&v)[LP_PX + 2], const float *x,
		int w, int h, int i0, int jj, bool colin, int mode)
{
	if (VEC && colin && jj >= 0 && jj < h) {
		const float *r = x + (size_t)jj * w + i0;
		const float4 c = *(const float4 *)r;
		v[0] = r[-1];
		v[1] = c.x; v[2] = c.y; v[3] = c.z; v[4] = c.w;
		v[5] = r[LP_PX];
		return;
	}
#pragma unroll
	for (int dx = 0; dx < LP_PX + 2; dx++) v[dx] = lp_sample(x, w, h, i0 - 1 + dx, jj, mode);
}

template <int V, bool VEC>
__global__ __launch_bounds__(LP_T) void k_plambda_lap(float *__restrict__ out,
		const float *__restrict__ in, int w, int h, int mode)
{
	const int strips = (h + LP_RS - 1) / LP_RS;
	const int l = blockIdx.x / strips, j0 = (blockIdx.x - l * strips) * LP_RS;
	const int i0 = (blockIdx.y * LP_T + threadIdx.x) * LP_PX;
	if (i0 >= w) return;
	const int j1 = j0 + LP_RS < h ? j0 + LP_RS : h;
	const float *x = in + (size_t)l * w * h;
	float *y = out + (size_t)l * w * h;
	constexpr const float *s = LapStencil<V>::s;
	const bool colin = i0 >= 1 && i0 + LP_PX < w;
	// rows j-1 .. j+LP_U of the current step (rows past the plane take the
	// boundary rule, rows past the strip are read and not written)
	float v[LP_U + 2][LP_PX + 2];
	lp_row<VEC>(v[0], x, w, h, i0, j0 - 1, colin, mode);
	lp_row<VEC>(v[1], x, w, h, i0, j0, colin, mode);
	for (int j = j0; j < j1; j += LP_U) {
#pragma unroll
		for (int t = 0; t < LP_U; t++)
			lp_row<VEC>(v[2 + t], x, w, h, i0, j + 1 + t, colin, mode);
#pragma unroll
		for (int t = 0; t < LP_U; t++) {
			if (j + t >= j1) break;
			float o[LP_PX];
#pragma unroll
			for (int p = 0; p < LP_PX; p++) {
				float r = 0.0f;
#pragma unroll
				for (int k = 0; k < 9; k++) r = r + s[k] * v[t + k / 3][p + k % 3];
				o[p] = r;
			}
			float *yr = y + (size_t)(j + t) * w;
			if (VEC && i0 + LP_PX <= w) {
				*(float4 *)(yr + i0) = make_float4(o[0], o[1], o[2], o[3]);
			} else {
#pragma unroll
				for (int p = 0; p < LP_PX; p++)
					if (i0 + p < w) yr[i0 + p] = o[p];
			}
		}
#pragma unroll
		for (int dx = 0; dx < LP_PX + 2; dx++) {
			v[0][dx] = v[LP_U][dx];
			v[1][dx] = v[LP_U + 1][dx];
		}
	}
}

void launch_plambda_lap(float *out, const float *in, int w, int h, int pd,
		int variant, int mode, hipStream_t s)
{
	const dim3 grid((h + LP_RS - 1) / LP_RS * pd, (w + LP_T * LP_PX - 1) / (LP_T * LP_PX));
	const bool vec = !(w & 3) && !((uintptr_t)out & 15) && !((uintptr_t)in & 15);
	if (variant) {
		if (vec) k_plambda_lap<1, true><<<grid, LP_T, 0, s>>>(out, in, w, h, mode);
		else k_plambda_lap<1, false><<<grid, LP_T, 0, s>>>(out, in, w, h, mode);
	} else {
		if (vec) k_plambda_lap<0, true><<<grid, LP_T, 0, s>>>(out, in, w, h, mode);
		else k_plambda_lap<0, false><<<grid, LP_T, 0, s>>>(out, in, w, h, mode);
	}
}

} // namespace spk
