/*
 * heatd_oracle.c -- CPU restatement of imscript's heatd (explicit linear
 * diffusion, isotropic or with a per-pixel diffusion matrix).  TEST
 * INFRASTRUCTURE: only the tests, the smoke check and the benchmark's
 * CPU-baseline leg use it, as the checker of the GPU path.
 *
 * Follows src/heatd.c:
 *   clamped pixel access                      heatd.c:1-8    (getpixel)
 *   9-point anisotropic Laplacian, / 4        heatd.c:10-56  (laplacian)
 *   one explicit step y = x + t * L(x)         heatd.c:58-70  (linear_diffusion_step)
 *   n steps, x := y after each                 heatd.c:72-88  (linear_diffusion)
 *   planes one after the other                 heatd.c:90-103 (linear_diffusion_separable)
 *
 * All arithmetic is float, without contraction, in the reference's order:
 * the nine products are summed for k = 0..8 including the zero-weight
 * corners of the isotropic scheme (0 * Inf is NaN there too).
 *
 * Parity: pinned by golden vectors of the reference heatd built from its
 * sources (oracle/_ref/heatd) and by closed forms (constants and linear
 * ramps are fixed points, mass conservation with clamped borders, the
 * anisotropic scheme with (a, b, c) = (1, 0, 1) equals the isotropic one).
 */
#include <stddef.h>
#include <string.h>

static float orc_hd_pix(const float *x, int w, int h, int i, int j)
{
	if (i < 0) i = 0;
	if (j < 0) j = 0;
	if (i >= w) i = w - 1;
	if (j >= h) j = h - 1;
	return x[j * w + i];
}

/* heatd.c:11-56; A: w*h*3 interleaved (a, b, c) or NULL */
static float orc_hd_lap(const float *x, const float *A, int w, int h, int i, int j)
{
	float a = 1, b = 0, c = 1;
	if (A) {
		a = A[3 * (j * w + i) + 0];
		b = A[3 * (j * w + i) + 1];
		c = A[3 * (j * w + i) + 2];
	}
	float m2a = -2 * a;
	float s[9] = {
		b / 2, c, -b / 2,
		a, m2a - 2 * c, a,
		-b / 2, c, b / 2,
	};
	float v[9];
	int k = 0;
	for (int dj = -1; dj <= 1; dj++)
		for (int di = -1; di <= 1; di++)
			v[k++] = orc_hd_pix(x, w, h, i + di, j + dj);
	float r = 0;
	for (k = 0; k < 9; k++) {
		float p = s[k] * v[k];
		r = r + p;
	}
	return r / 4;
}

/* heatd.c:58-70 */
void orc_heatd_step(float *y, const float *x, int w, int h, const float *A, float t)
{
	for (int j = 0; j < h; j++)
		for (int i = 0; i < w; i++) {
			float l = orc_hd_lap(x, A, w, h, i, j);
			float d = t * l;
			y[j * w + i] = x[j * w + i] + d;
		}
}

/* heatd.c:90-103 with :72-88: n steps on each of pd planes; x ends equal
 * to y when n > 0, and y is not written when n <= 0 */
void orc_heatd(float *y, float *x, int w, int h, int pd, const float *A, int n,
		float t)
{
	size_t np = (size_t)w * h;
	for (int l = 0; l < pd; l++)
		for (int s = 0; s < n; s++) {
			orc_heatd_step(y + np * l, x + np * l, w, h, A, t);
			memcpy(x + np * l, y + np * l, np * sizeof(float));
		}
}
