// heatd's explicit linear diffusion (heatd.c:10-103): n steps of
// y = x + t * L(x), L the 9-point scheme of the diffusion matrix (a, b, c)
// of each pixel (the identity (1, 0, 1) without a field) divided by 4,
// clamped borders.  Planar float planes, bit-identical to the reference:
// the nine products s[k] * v[k] are summed for k = 0..8 in order, zero
// weights included (0 * Inf is NaN there too), without contraction.
#include <stdint.h>
#include "sp_internal.h"

namespace spk {

constexpr int HD_T = 256, HD_PX = 4;   // threads per block, pixels per thread

// grid: x = rows of all planes, y = column chunks of HD_T*HD_PX.  A: the
// field, w*h*3 interleaved (a, b, c) shared by every plane, or unused.
template <bool ANISO, bool VEC>
__global__ __launch_bounds__(HD_T) void k_heatd_step(float *__restrict__ out,
		const float *__restrict__ in, const float *__restrict__ A, int w, int h,
		float t)
{
	const int row = blockIdx.x, l = row / h, j = row - l * h;
	const int i0 = (blockIdx.y * HD_T + threadIdx.x) * HD_PX;
	if (i0 >= w) return;
	const float *x = in + (size_t)l * w * h;
	float *y = out + (size_t)l * w * h + (size_t)j * w;
	// the 3 x (HD_PX + 2) window, rows and columns clamped (heatd.c:1-8)
	float v[3][HD_PX + 2];
	if (VEC) {
		// w is a multiple of 4, so the thread's 4 pixels are in the row:
		// one 16-byte load per clamped row, the two side columns clamped
		const bool lt = i0 > 0, rt = i0 + HD_PX < w;
#pragma unroll
		for (int dy = 0; dy < 3; dy++) {
			int jj = j - 1 + dy;
			jj = jj < 0 ? 0 : jj >= h ? h - 1 : jj;
			const float *r = x + (size_t)jj * w + i0;
			const float4 c = *(const float4 *)r;
			v[dy][0] = lt ? r[-1] : c.x;
			v[dy][1] = c.x; v[dy][2] = c.y; v[dy][3] = c.z; v[dy][4] = c.w;
			v[dy][5] = rt ? r[HD_PX] : c.w;
		}
	} else {
#pragma unroll
		for (int dy = 0; dy < 3; dy++) {
			int jj = j - 1 + dy;
			jj = jj < 0 ? 0 : jj >= h ? h - 1 : jj;
#pragma unroll
			for (int dx = 0; dx < HD_PX + 2; dx++) {
				int i = i0 - 1 + dx;
				i = i < 0 ? 0 : i >= w ? w - 1 : i;
				v[dy][dx] = x[(size_t)jj * w + i];
			}
		}
	}
	float fa[HD_PX], fb[HD_PX], fc[HD_PX];
	if (ANISO) {
		const float *ap = A + 3 * ((size_t)j * w + i0);
		if (VEC && i0 + HD_PX <= w) {
			const float4 q0 = *(const float4 *)ap, q1 = *(const float4 *)(ap + 4),
				q2 = *(const float4 *)(ap + 8);
			fa[0] = q0.x; fb[0] = q0.y; fc[0] = q0.z;
			fa[1] = q0.w; fb[1] = q1.x; fc[1] = q1.y;
			fa[2] = q1.z; fb[2] = q1.w; fc[2] = q2.x;
			fa[3] = q2.y; fb[3] = q2.z; fc[3] = q2.w;
		} else {
#pragma unroll
			for (int p = 0; p < HD_PX; p++) {
				const bool ok = i0 + p < w;
				fa[p] = ok ? ap[3 * p] : 1.0f;
				fb[p] = ok ? ap[3 * p + 1] : 0.0f;
				fc[p] = ok ? ap[3 * p + 2] : 1.0f;
			}
		}
	} else {
#pragma unroll
		for (int p = 0; p < HD_PX; p++) { fa[p] = 1.0f; fb[p] = 0.0f; fc[p] = 1.0f; }
	}
	float o[HD_PX];
#pragma unroll
	for (int p = 0; p < HD_PX; p++) {
		const float a = fa[p], b = fb[p], c = fc[p];
		const float hb = b / 2, nhb = -b / 2, m = -2 * a - 2 * c;
		const float s[9] = {hb, c, nhb, a, m, a, nhb, c, hb};
		float r = 0.0f;
#pragma unroll
		for (int k = 0; k < 9; k++) r = r + s[k] * v[k / 3][p + k % 3];
		o[p] = v[1][p + 1] + t * (r / 4);
	}
	if (VEC && i0 + HD_PX <= w) {
		*(float4 *)(y + i0) = make_float4(o[0], o[1], o[2], o[3]);
	} else {
#pragma unroll
		for (int p = 0; p < HD_PX; p++)
			if (i0 + p < w) y[i0 + p] = o[p];
	}
}

void launch_heatd_step(float *out, const float *in, const float *A, int w, int h,
		int pd, float t, hipStream_t s)
{
	const dim3 grid(h * pd, (w + HD_T * HD_PX - 1) / (HD_T * HD_PX));
	const bool vec = !(w & 3) && !((uintptr_t)out & 15) && !((uintptr_t)in & 15) &&
		!((uintptr_t)A & 15);
	if (A) {
		if (vec) k_heatd_step<true, true><<<grid, HD_T, 0, s>>>(out, in, A, w, h, t);
		else k_heatd_step<true, false><<<grid, HD_T, 0, s>>>(out, in, A, w, h, t);
	} else {
		if (vec) k_heatd_step<false, true><<<grid, HD_T, 0, s>>>(out, in, A, w, h, t);
		else k_heatd_step<false, false><<<grid, HD_T, 0, s>>>(out, in, A, w, h, t);
	}
}

} // namespace spk
