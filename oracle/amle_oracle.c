/*
 * amle_oracle.c -- CPU restatement of imscript's amle (absolutely minimizing
 * Lipschitz extension) hole filling.  TEST INFRASTRUCTURE: only the tests,
 * the smoke check and the benchmark's CPU-baseline leg use it, as the
 * checker of the GPU path; it is not part of the product.
 *
 * Follows src/amle.c:
 *   neighbour table and in-image gather    amle.c:15-43   (get_nvals)
 *   first-index strict min/max scan         amle.c:57-65   (get_minmax_idx)
 *   one raster-order in-place sweep         amle.c:67-88   (amle_iteration)
 *   initial bounds (global min / max)       amle.c:90-110  (amle_init)
 *   max and mean |tinf - tsup| over holes   amle.c:112-139
 *   driver: holes in raster order, NITER sweeps of tinf then tsup,
 *   output tsup                             amle.c:203-246 (amle)
 *
 * The output depends on tsup only (amle.c:244); tinf feeds the diagnostic
 * line printed every tenth sweep (amle.c:230-233), so it is computed only
 * when diagnostics are asked for.  All arithmetic is float, without
 * contraction, in the reference's order.
 *
 * Parity: pinned by golden vectors of the reference amle built from its
 * sources (oracle/_ref/amle) and by closed forms (linear interpolation on a
 * strip, constant images, the max principle).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* amle.c:16-29: {dx, dy, dx*dx + dy*dy} in the reference's order */
static const int orc_amle_nbr[32][3] = {
	{+1, 0, 1}, {0, +1, 1}, {-1, 0, 1}, {0, -1, 1},
	{+1, +1, 2}, {-1, -1, 2},
	{-1, +1, 2}, {+1, -1, 2},
	{+2, +1, 5}, {+1, +2, 5}, {+2, -1, 5}, {+1, -2, 5},
	{-2, -1, 5}, {-1, -2, 5}, {-2, +1, 5}, {-1, +2, 5},
	{+3, +1, 10}, {+1, +3, 10}, {+3, -1, 10}, {+1, -3, 10},
	{-3, -1, 10}, {-1, -3, 10}, {-3, +1, 10}, {-1, +3, 10},
	{+3, +2, 13}, {+2, +3, 13}, {+3, -2, 13}, {+2, -3, 13},
	{-3, -2, 13}, {-2, -3, 13}, {-3, +2, 13}, {-2, +3, 13},
};

/* new value of hole (i, j): amle.c:74-81 with get_nvals / get_minmax_idx */
static float orc_amle_update(const float *x, int w, int h, int i, int j, int nn)
{
	float value[32] = {0}, weight[32] = {0};
	int nv = 0;
	for (int p = 0; p < nn; p++) {
		int ii = i + orc_amle_nbr[p][0], jj = j + orc_amle_nbr[p][1];
		if (ii >= 0 && jj >= 0 && ii < w && jj < h) {
			value[nv] = x[w * jj + ii];
			weight[nv] = (float)orc_amle_nbr[p][2];
			nv++;
		}
	}
	int mn = 0, mx = 0;
	for (int k = 1; k < nv; k++) {
		if (value[k] < value[mn]) mn = k;
		if (value[k] > value[mx]) mx = k;
	}
	float a = weight[mx], b = weight[mn];
	float p1 = a * value[mn];
	float p2 = b * value[mx];
	float num = p1 + p2;
	float den = a + b;
	return num / den;
}

/* one sweep over the holes in raster order; returns the float running sum
 * of |old - new| (amle.c:67-88) */
static float orc_amle_sweep(float *x, int w, int h, const int *holes, int nh,
		int nn)
{
	float actus = 0;
	for (int p = 0; p < nh; p++) {
		int idx = holes[p], i = idx % w, j = idx / w;
		float nx = orc_amle_update(x, w, h, i, j, nn);
		actus += fabsf(x[idx] - nx);
		x[idx] = nx;
	}
	return actus;
}

/* Sweeps of amle_iteration on an image already initialised: x holds the
 * current values, `hole` marks the pixels to update (nonzero = hole). */
void orc_amle_sweeps(float *x, const unsigned char *hole, int w, int h,
		int nn, int nsweep)
{
	int *holes = malloc(sizeof(int) * (size_t)(w * h > 0 ? w * h : 1)), nh = 0;
	for (int q = 0; q < w * h; q++)
		if (hole[q]) holes[nh++] = q;
	for (int s = 0; s < nsweep; s++)
		orc_amle_sweep(x, w, h, holes, nh, nn);
	free(holes);
}

/*
 * amle(y, x, w, h) of one plane.  nn: AMLE_NN (1..32); niter: AMLE_NITER;
 * init: the AMLE_INIT image (w*h, tsup at holes starts from it) or NULL.
 * diag: NULL, or room for 4*niter floats receiving, per sweep, e, ea,
 * actus_inf, actus_sup of amle.c:226-229 (tinf is then computed too).
 * Returns the number of holes.
 */
int orc_amle(float *y, const float *x, int w, int h, int nn, int niter,
		const float *init, float *diag)
{
	int n = w * h;
	int *holes = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), nh = 0;
	for (int j = 0; j < h; j++)
		for (int i = 0; i < w; i++)
			if (isnan(x[j * w + i])) holes[nh++] = j * w + i;
	/* amle.c:42-55 get_minmax over the whole image (NaN never compares) */
	float mn = INFINITY, mx = -INFINITY;
	for (int q = 0; q < n; q++) {
		if (x[q] < mn) mn = x[q];
		if (x[q] > mx) mx = x[q];
	}
	float *tinf = diag ? malloc(sizeof(float) * (size_t)(n > 0 ? n : 1)) : NULL;
	for (int q = 0; q < n; q++) {
		int hole = isnan(x[q]);
		y[q] = hole ? mx : x[q];
		if (hole && init) y[q] = init[q];
		if (tinf) tinf[q] = hole ? mn : x[q];
	}
	for (int it = 0; it < niter; it++) {
		float ai = 0;
		if (tinf) ai = orc_amle_sweep(tinf, w, h, holes, nh, nn);
		float as = orc_amle_sweep(y, w, h, holes, nh, nn);
		if (diag) {
			float e = 0, ea = 0;
			for (int p = 0; p < nh; p++) {
				float t = fabsf(tinf[holes[p]] - y[holes[p]]);
				if (t > e) e = t;
				ea += t;
			}
			diag[4 * it + 0] = e;
			diag[4 * it + 1] = ea / nh;
			diag[4 * it + 2] = ai;
			diag[4 * it + 3] = as;
		}
	}
	free(tinf);
	free(holes);
	return nh;
}
