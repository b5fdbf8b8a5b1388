// heatd's explicit linear diffusion (heatd.c:10-103): n steps of
// y = x + t * L(x), L the 9-point scheme of the diffusion matrix (a, b, c)
// of each pixel (the identity (1, 0, 1) without a field) divided by 4,
// clamped borders.  Planar float planes, bit-identical to the reference:
// the nine products s[k] * v[k] are summed for k = 0..8 in order, zero
// weights included (0 * Inf is NaN there too), without contraction.
#include <stdint.h>
#include <type_traits>
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

// ------------------------------------------------ several steps per launch
//
// Isotropic diffusion only.  Each wave owns a tile of 64 columns (one per
// lane) by HT_R rows held in registers, loaded with a halo of T pixels, and
// runs T steps on it: after step k the values are exact everywhere but
// within k pixels of the tile's edge, so the central OW x OH = (64 - 2T) x
// (HT_R - 2T) pixels are written back.  Horizontal neighbours come from
// the adjacent lanes, vertical ones from the lane's own registers; no LDS,
// no barrier.  The last tile of a row or column is moved back to end at the
// image's edge (its overlap with the previous tile is computed twice, with
// the same result), so that the image's first and last rows sit at tile
// rows T and HT_R - 1 - T, fixed when compiling; there the clamped borders
// of heatd.c:1-8 are applied, and per lane at the first and last columns.
// The planes must be at least OW x OH (the host's choice).
//
// Update: with the isotropic weights {0, 1, -0, 1, -4, 1, -0, 1, 0} the
// reference's ordered sum r = 0 + 0*NW + N + (-0)*NE + W + (-4)*C + E +
// (-0)*SW + S + 0*SE equals ((((0 + N) + W) + (-4)*C) + E) + S whenever the
// four corners are finite: r starts at +0 and a sum is -0 only if both terms
// are, so r is never -0 and adding a zero product leaves it unchanged.  A
// non-finite corner makes its product, and r, NaN.

constexpr int HT_R = 48;          // rows per wave tile
constexpr int HT_WAVES = 4;       // wave tiles per block, stacked vertically

// grid: x = tile columns of all planes (tiles_x per plane), y = groups of
// HT_WAVES tile rows
template <int T>
__global__ __launch_bounds__(64 * HT_WAVES) void k_heatd_multi(float *__restrict__ out,
		const float *__restrict__ in, int w, int h, int tiles_x, int tiles_y, float t)
{
	constexpr int OW = 64 - 2 * T, OH = HT_R - 2 * T;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int ty = blockIdx.y * HT_WAVES + wave;
	if (ty >= tiles_y) return;
	const int l = blockIdx.x / tiles_x, tx = blockIdx.x - l * tiles_x;
	const float *x = in + (size_t)l * w * h;
	float *y = out + (size_t)l * w * h;
	const int x0 = min(tx * OW, w - OW), y0 = min(ty * OH, h - OH);
	const int gi = x0 - T + lane, gj0 = y0 - T;
	const int ci = gi < 0 ? 0 : gi >= w ? w - 1 : gi;
	float v[HT_R];
#pragma unroll
	for (int r = 0; r < HT_R; r++) {
		int gj = gj0 + r;
		gj = gj < 0 ? 0 : gj >= h ? h - 1 : gj;
		v[r] = x[(size_t)gj * w + ci];
	}
	// per-lane clamps wherever the tile's columns reach an image border
	const bool atl = gi == 0, atr = gi == w - 1;
	const bool edge_cols = x0 - T <= 0 || x0 + OW + T - 1 >= w - 1;
	// clamped rows only in the first and the last tile of a column (image
	// rows 0 and h-1 at tile rows T and HT_R-1-T); another tile whose rows
	// reach past the image does not clamp, so rows within T-1 of such a
	// border are wrong there and are left to the first / last tile
	const bool top = y0 == 0, bot = y0 == h - OH;
	for (int k = 0; k < T; k++) {
		// window of old values: rows r-1 (u), r (c), r+1 (d), each with its
		// west / east neighbours
		// the corners of row r are the west / east neighbours of rows r-1
		// and r+1: whether a row's pair is finite is kept with the row
		float c = v[0];
		float cw = __shfl_up(c, 1, 64), ce = __shfl_down(c, 1, 64);
		if (edge_cols) { cw = atl ? c : cw; ce = atr ? c : ce; }
		bool cf = isfinite(cw) && isfinite(ce);
		float u = c;
		bool uf = cf;
#pragma unroll
		for (int r = 0; r < HT_R; r++) {
			const float d = r + 1 < HT_R ? v[r + 1] : c;
			float dw = __shfl_up(d, 1, 64), de = __shfl_down(d, 1, 64);
			if (edge_cols) { dw = atl ? d : dw; de = atr ? d : de; }
			const bool df = isfinite(dw) && isfinite(de);
			float n = u, so = d;
			bool nf = uf, sf = df;
			if (r == T && top) { n = c; nf = cf; }
			if (r == HT_R - 1 - T && bot) { so = c; sf = cf; }
			float q = 0.0f;
			q = q + n;
			q = q + cw;
			q = q + -4.0f * c;
			q = q + ce;
			q = q + so;
			q = nf && sf ? q : __builtin_nanf("");
			const float nv = c + t * (q * 0.25f);
			u = c; uf = cf;
			c = d; cw = dw; ce = de; cf = df;
			v[r] = nv;
		}
	}
	if (lane < T || lane >= 64 - T) return;
	const int jlo = top ? 0 : T - 1, jhi = bot ? h : h - T + 1;
#pragma unroll
	for (int r = T; r < HT_R - T; r++) {
		const int gj = gj0 + r;
		if (gj >= jlo && gj < jhi) y[(size_t)gj * w + gi] = v[r];
	}
}

// With a diffusion-matrix field: the same register tiles, HR rows per wave,
// each lane also holding the (a, b, c) of its column's pixels, and the
// reference's nine ordered products per pixel (heatd.c:10-45 with
// s = {b/2, c, -b/2, a, -2a-2c, a, -b/2, c, b/2}, r / 4), the corners being
// the west / east neighbours of the rows above and below.  Borders as in
// k_heatd_multi: a clamped north (south) row brings its corners with it, and
// the per-row west / east clamps make the corners of the border columns.
// The field is read once per launch instead of once per step.
template <int T, int HR>
__global__ __launch_bounds__(64 * HT_WAVES) void k_heatd_multi_field(float *__restrict__ out,
		const float *__restrict__ in, const float *__restrict__ A, int w, int h,
		int tiles_x, int tiles_y, float t)
{
	constexpr int OW = 64 - 2 * T, OH = HR - 2 * T;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int ty = blockIdx.y * HT_WAVES + wave;
	if (ty >= tiles_y) return;
	const int l = blockIdx.x / tiles_x, tx = blockIdx.x - l * tiles_x;
	const float *x = in + (size_t)l * w * h;
	float *y = out + (size_t)l * w * h;
	const int x0 = min(tx * OW, w - OW), y0 = min(ty * OH, h - OH);
	const int gi = x0 - T + lane, gj0 = y0 - T;
	const int ci = gi < 0 ? 0 : gi >= w ? w - 1 : gi;
	float v[HR], fa[HR], fb[HR], fc[HR];
#pragma unroll
	for (int r = 0; r < HR; r++) {
		int gj = gj0 + r;
		gj = gj < 0 ? 0 : gj >= h ? h - 1 : gj;
		const size_t q = (size_t)gj * w + ci;
		v[r] = x[q];
		fa[r] = A[3 * q];
		fb[r] = A[3 * q + 1];
		fc[r] = A[3 * q + 2];
	}
	const bool atl = gi == 0, atr = gi == w - 1;
	const bool edge_cols = x0 - T <= 0 || x0 + OW + T - 1 >= w - 1;
	const bool top = y0 == 0, bot = y0 == h - OH;
	for (int k = 0; k < T; k++) {
		float c = v[0];
		float cw = __shfl_up(c, 1, 64), ce = __shfl_down(c, 1, 64);
		if (edge_cols) { cw = atl ? c : cw; ce = atr ? c : ce; }
		float u = c, uw = cw, ue = ce;
#pragma unroll
		for (int r = 0; r < HR; r++) {
			const float d = r + 1 < HR ? v[r + 1] : c;
			float dw = __shfl_up(d, 1, 64), de = __shfl_down(d, 1, 64);
			if (edge_cols) { dw = atl ? d : dw; de = atr ? d : de; }
			float n = u, nw = uw, ne = ue, so = d, sw = dw, se = de;
			if (r == T && top) { n = c; nw = cw; ne = ce; }
			if (r == HR - 1 - T && bot) { so = c; sw = cw; se = ce; }
			const float a = fa[r], b = fb[r], cc = fc[r];
			const float hb = b / 2, nhb = -b / 2, m = -2 * a - 2 * cc;
			float q = 0.0f;
			q = q + hb * nw;
			q = q + cc * n;
			q = q + nhb * ne;
			q = q + a * cw;
			q = q + m * c;
			q = q + a * ce;
			q = q + nhb * sw;
			q = q + cc * so;
			q = q + hb * se;
			const float nv = c + t * (q / 4);
			u = c; uw = cw; ue = ce;
			c = d; cw = dw; ce = de;
			v[r] = nv;
		}
	}
	if (lane < T || lane >= 64 - T) return;
	const int jlo = top ? 0 : T - 1, jhi = bot ? h : h - T + 1;
#pragma unroll
	for (int r = T; r < HR - T; r++) {
		const int gj = gj0 + r;
		if (gj >= jlo && gj < jhi) y[(size_t)gj * w + gi] = v[r];
	}
}

int heatd_multi_max() { return 8; }

bool heatd_multi_fits(int w, int h, int T, int rows)
{
	return w >= 64 - 2 * T && h >= rows - 2 * T;
}

bool launch_heatd_multi_field(float *out, const float *in, const float *A, int w, int h,
		int pd, int T, int rows, float t, hipStream_t s)
{
	// rows - 2T >= T - 1: the first and last tiles of a column write every
	// row that the others leave (those within T - 1 of the image's edge)
	if (T < 1 || T > heatd_multi_max() ||
			(rows != 16 && rows != 24 && rows != 32 && rows != 48) ||
			!heatd_multi_fits(w, h, T, rows) || rows - 2 * T < T - 1)
		return false;
	auto go = [&](auto tc, auto rc) {
		constexpr int TT = decltype(tc)::value, RR = decltype(rc)::value;
		constexpr int OW = 64 - 2 * TT, OH = RR - 2 * TT;
		const int tiles_x = (w + OW - 1) / OW, tiles_y = (h + OH - 1) / OH;
		const dim3 grid(tiles_x * pd, (tiles_y + HT_WAVES - 1) / HT_WAVES);
		k_heatd_multi_field<TT, RR><<<grid, 64 * HT_WAVES, 0, s>>>(out, in, A, w, h,
				tiles_x, tiles_y, t);
	};
	auto by_rows = [&](auto tc) {
		constexpr int TT = decltype(tc)::value;
		if constexpr (16 - 2 * TT >= TT - 1)
			if (rows == 16) return go(tc, std::integral_constant<int, 16>{});
		if constexpr (24 - 2 * TT >= TT - 1)
			if (rows == 24) return go(tc, std::integral_constant<int, 24>{});
		if (rows == 32) return go(tc, std::integral_constant<int, 32>{});
		go(tc, std::integral_constant<int, 48>{});
	};
	switch (T) {
	case 1: by_rows(std::integral_constant<int, 1>{}); break;
	case 2: by_rows(std::integral_constant<int, 2>{}); break;
	case 3: by_rows(std::integral_constant<int, 3>{}); break;
	case 4: by_rows(std::integral_constant<int, 4>{}); break;
	case 5: by_rows(std::integral_constant<int, 5>{}); break;
	case 6: by_rows(std::integral_constant<int, 6>{}); break;
	case 7: by_rows(std::integral_constant<int, 7>{}); break;
	default: by_rows(std::integral_constant<int, 8>{}); break;
	}
	return true;
}

bool launch_heatd_multi(float *out, const float *in, int w, int h, int pd, int T,
		float t, hipStream_t s)
{
	if (T < 1 || T > heatd_multi_max() || !heatd_multi_fits(w, h, T, HT_R)) return false;
	auto go = [&](auto tc) {
		constexpr int TT = decltype(tc)::value;
		constexpr int OW = 64 - 2 * TT, OH = HT_R - 2 * TT;
		const int tiles_x = (w + OW - 1) / OW, tiles_y = (h + OH - 1) / OH;
		const dim3 grid(tiles_x * pd, (tiles_y + HT_WAVES - 1) / HT_WAVES);
		k_heatd_multi<TT><<<grid, 64 * HT_WAVES, 0, s>>>(out, in, w, h, tiles_x,
				tiles_y, t);
	};
	switch (T) {
	case 1: go(std::integral_constant<int, 1>{}); break;
	case 2: go(std::integral_constant<int, 2>{}); break;
	case 3: go(std::integral_constant<int, 3>{}); break;
	case 4: go(std::integral_constant<int, 4>{}); break;
	case 5: go(std::integral_constant<int, 5>{}); break;
	case 6: go(std::integral_constant<int, 6>{}); break;
	case 7: go(std::integral_constant<int, 7>{}); break;
	default: go(std::integral_constant<int, 8>{}); break;
	}
	return true;
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
