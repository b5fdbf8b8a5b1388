/*
 * plambda_lap_oracle.c -- CPU restatement of plambda's 3x3 Laplacian image
 * operators "x,l" and "x,lr", written from scratch as TEST INFRASTRUCTURE.
 *
 * The test-suite and the CPU-baseline leg of the benchmark load it to check
 * the HIP producer of simpois's guidance field; the product never calls it.
 * Built with the rest of the oracle (-O3 -ffp-contract=off).
 *
 * Pinned by golden vectors written by the reference plambda binary
 * (tests/golden/make_golden_plambda.py, compiled from the reference sources
 * by oracle/Makefile) and by closed forms (Laplacian of quadratics, the
 * boundary rules on tiny images).
 */
#include <math.h>
#include <stddef.h>

/* gmod and positive_reflex of getpixel.c:64-83 */
static int orc_gmod(int x, int m)
{
	int r = x % m;
	return r < 0 ? r + m : r;
}

static int orc_reflex(int n, int p)
{
	int r = orc_gmod(n, 2 * p);
	if (r == p) r -= 1;
	if (r > p) r = 2 * p - r;
	return r;
}

/* one sample of plane x (w x h) at (i, j) under plambda's boundary rule
 * (plambda.c:2573-2588 selecting getpixel.c): 0 zero (:10-16), 1 nearest
 * (:18-29, the default), 2 reflection (:85-91), 3 periodic (:93-99),
 * 4 NaN (:31-38) */
static float orc_sample(const float *x, int w, int h, int i, int j, int mode)
{
	const int in = i >= 0 && i < w && j >= 0 && j < h;
	switch (mode) {
	case 0:
		return in ? x[(size_t)j * w + i] : 0.0f;
	case 2:
		i = orc_reflex(i, w);
		j = orc_reflex(j, h);
		break;
	case 3:
		i = orc_gmod(i, w);
		j = orc_gmod(j, h);
		break;
	case 4:
		return in ? x[(size_t)j * w + i] : NAN;
	default:
		if (i < 0) i = 0;
		if (j < 0) j = 0;
		if (i >= w) i = w - 1;
		if (j >= h) j = h - 1;
	}
	return x[(size_t)j * w + i];
}

/* the stencils of plambda.c:2589-2605: "l" (5-point) and "lr" (the
 * diagonal Laplacian of the Roberts scheme) */
static const float orc_st_l[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
static const float orc_st_lr[9] = {0.5f, 0, 0.5f, 0, -2, 0, 0.5f, 0, 0.5f};

/* out = apply_3x3_stencil (plambda.c:2680-2689) at every pixel of every
 * plane of the planar w x h x pd image in: r = 0; r += s[k] * P(i-1+k%3,
 * j-1+k/3) for k = 0..8, zero weights included (0 * NaN is NaN there too).
 * variant 0 = "x,l", 1 = "x,lr". */
void orc_plambda_lap(float *out, const float *in, int w, int h, int pd,
		int variant, int mode)
{
	const float *s = variant ? orc_st_lr : orc_st_l;
	for (int l = 0; l < pd; l++) {
		const float *x = in + (size_t)l * w * h;
		float *y = out + (size_t)l * w * h;
		for (int j = 0; j < h; j++)
		for (int i = 0; i < w; i++) {
			float r = 0;
			for (int k = 0; k < 9; k++)
				r += s[k] * orc_sample(x, w, h, i - 1 + k % 3, j - 1 + k / 3, mode);
			y[(size_t)j * w + i] = r;
		}
	}
}
