/*
 * simpois_oracle.c -- CPU restatement of imscript's simpois hole-filling
 * path, written from scratch as TEST INFRASTRUCTURE.
 *
 * This file is the oracle: only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it, and only to check or time the
 * HIP implementation in imscript_amd/.  The product never calls it.
 *
 * Build: -O3 -ffp-contract=off (no FMA contraction, no -march), which
 * reproduces the reference's own `cc -O3` arithmetic bit for bit.
 *
 * Every function cites the reference lines it restates.  Arithmetic is
 * kept in the reference's types and association order:
 *   - GS stencils are float expressions (simpois.c:35-90);
 *   - the CG operator reads float-rounded values and evaluates its stencil
 *     in float, while the CG vectors and dot products are double
 *     (cleant_cgpois.c:48-121, minicg.c:12-56);
 *   - environment knobs (PMSFAC, PONLIT, CG_EPS) are passed explicitly in
 *     struct orc_knobs instead of being read with getenv once per process
 *     (smapa.h:44-60), so one process can test several settings.
 *
 * Pinned by: the golden fixtures in tests/golden (written by the reference
 * binary itself), the compiled reference in oracle/_ref when present, and
 * closed-form tests (harmonic ramps, cubic annihilation, maximum principle,
 * restriction/prolongation weights, dense FP64 solves of the masked CG
 * system).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_knobs {
	double pmsfac;   /* PMSFAC, default 3    (simpois.c:244) */
	double ponlit;   /* PONLIT, default 0    (simpois.c:245) */
	double cg_eps;   /* CG_EPS, default -1   (cleant_cgpois.c:125) */
} orc_knobs;

/* optional record of the CG convergence trace (minicg.c:46):
 * one (n, iter, rr_new) triple per CG iteration, all levels, in order */
typedef struct orc_trace {
	int cap, len;
	int *n, *iter;
	double *rr;
} orc_trace;

static void trace_push(orc_trace *t, int n, int iter, double rr)
{
	if (!t || t->len >= t->cap) { if (t) t->len++; return; }
	t->n[t->len] = n;
	t->iter[t->len] = iter;
	t->rr[t->len] = rr;
	t->len++;
}

static void *orc_alloc(size_t n)
{
	void *p = malloc(n ? n : 1);
	if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
	return p;
}

/* ---------------------------------------------------------------- stencils */

/* nearest-value (Neumann) extension, simpois.c:25-32 */
static inline float at(const float *u, int w, int h, int i, int j)
{
	i = i < 0 ? 0 : (i >= w ? w - 1 : i);
	j = j < 0 ? 0 : (j >= h ? h - 1 : j);
	return u[(size_t)j * w + i];
}

/* 5-point Laplacian, simpois.c:35-46: ((((-4c + E) + S) + W) + N) in float */
static float lap5(const float *u, int w, int h, int i, int j)
{
	float c = at(u, w, h, i, j);
	float r = -4 * c;
	r = r + at(u, w, h, i + 1, j);
	r = r + at(u, w, h, i, j + 1);
	r = r + at(u, w, h, i - 1, j);
	r = r + at(u, w, h, i, j - 1);
	return r;
}

/* nested Laplacian of the Laplacian, simpois.c:48-57 */
static float lap5_2(const float *u, int w, int h, int i, int j)
{
	float r = -4 * lap5(u, w, h, i, j);
	r = r + lap5(u, w, h, i + 1, j);
	r = r + lap5(u, w, h, i, j + 1);
	r = r + lap5(u, w, h, i - 1, j);
	r = r + lap5(u, w, h, i, j - 1);
	return r;
}

/* triple nesting, simpois.c:59-68 */
static float lap5_3(const float *u, int w, int h, int i, int j)
{
	float r = -4 * lap5_2(u, w, h, i, j);
	r = r + lap5_2(u, w, h, i + 1, j);
	r = r + lap5_2(u, w, h, i, j + 1);
	r = r + lap5_2(u, w, h, i - 1, j);
	r = r + lap5_2(u, w, h, i, j - 1);
	return r;
}

/* Laplacian that ignores non-finite neighbours, simpois.c:72-90.
 * Neighbour order E, S, W, N; each accepted term adds (nbr - centre). */
static float lap_neumann(const float *u, int w, int h, int i, int j)
{
	float x = at(u, w, h, i, j);
	float nb[4] = {
		at(u, w, h, i + 1, j), at(u, w, h, i, j + 1),
		at(u, w, h, i - 1, j), at(u, w, h, i, j - 1) };
	float r = 0;
	for (int k = 0; k < 4; k++)
		if (isfinite(nb[k]))
			r += nb[k] - x;
	return r;
}

/* operator choice from the sign of the step, simpois.c:96-106 */
enum { OP_NEUMANN = 0, OP_BILAP = 1, OP_TRILAP = 2 };

static int op_of_step(float tstep, float *eff)
{
	if (!(tstep < 0)) { *eff = tstep; return OP_NEUMANN; }
	if (tstep < -1000) {
		float t = tstep;
		t += 1000;
		t *= -1;
		*eff = t;
		return OP_TRILAP;
	}
	*eff = tstep;
	return OP_BILAP;
}

static float apply_op(int op, const float *u, int w, int h, int i, int j)
{
	if (op == OP_NEUMANN) return lap_neumann(u, w, h, i, j);
	if (op == OP_BILAP) return lap5_2(u, w, h, i, j);
	return lap5_3(u, w, h, i, j);
}

/* ---------------------------------------------------------- Gauss-Seidel */

/* niter in-place raster-order sweeps over the NaN pixels of g,
 * simpois.c:93-119 and :163-164.  f may be NULL (data term 0). */
void orc_gs_sweeps(float *u, const float *g, const float *f, int w, int h,
		float tstep, int niter)
{
	float t;
	int op = op_of_step(tstep, &t);
	for (int it = 0; it < niter; it++)
		for (int j = 0; j < h; j++)
		for (int i = 0; i < w; i++) {
			size_t ij = (size_t)j * w + i;
			if (!isnan(g[ij])) continue;
			float l = apply_op(op, u, w, h, i, j);
			float d = f ? f[ij] : 0;
			u[ij] = u[ij] + t * (l - d);
		}
}

/* The same sweeps executed hyperplane by hyperplane: the update of pixel
 * (i,j) at sweep k belongs to hyperplane s = i + R*j + (R*R+1)*k, with
 * R = 1, 2, 3 for the 5, 13 and 25-point operators.  Inside a hyperplane
 * the updates are visited in an order scrambled by `seed`.  This is the
 * schedule the GPU kernels use; equality with orc_gs_sweeps is the claim
 * that it reproduces raster-order Gauss-Seidel bit for bit. */
void orc_gs_sweeps_hyperplane(float *u, const float *g, const float *f,
		int w, int h, float tstep, int niter, unsigned seed)
{
	float t;
	int op = op_of_step(tstep, &t);
	int R = op + 1, B = R, C = R * R + 1;
	long smax = (long)(w - 1) + (long)B * (h - 1) + (long)C * (niter - 1);
	/* at most one point per (sweep, row) on a hyperplane */
	int cap = h * (niter > 0 ? niter : 1) + 1;
	int *pi = orc_alloc(sizeof(int) * cap), *pj = orc_alloc(sizeof(int) * cap);
	unsigned st = seed * 2654435761u + 1;
	for (long s = 0; s <= smax && niter > 0; s++) {
		int n = 0;
		for (int k = 0; k < niter; k++)
		for (int j = 0; j < h; j++) {
			long i = s - (long)B * j - (long)C * k;
			if (i < 0 || i >= w) continue;
			pi[n] = (int)i + w * k; /* packs (i, k) */
			pj[n] = j;
			n++;
		}
		for (int a = n - 1; a > 0; a--) { /* Fisher-Yates */
			st = st * 1103515245u + 12345u;
			int b = (int)((st >> 8) % (unsigned)(a + 1));
			int x = pi[a]; pi[a] = pi[b]; pi[b] = x;
			x = pj[a]; pj[a] = pj[b]; pj[b] = x;
		}
		for (int q = 0; q < n; q++) {
			int i = pi[q] % w, j = pj[q];
			size_t ij = (size_t)j * w + i;
			if (!isnan(g[ij])) continue;
			float l = apply_op(op, u, w, h, i, j);
			float d = f ? f[ij] : 0;
			u[ij] = u[ij] + t * (l - d);
		}
	}
	free(pi);
	free(pj);
}

/* ---------------------------------------------------------------- pyramid */

/* 2x2 NaN-aware block mean, simpois.c:172-196 */
void orc_restrict(float *out, int ow, int oh, const float *in, int iw, int ih)
{
	for (int j = 0; j < oh; j++)
	for (int i = 0; i < ow; i++) {
		float v[4] = {
			at(in, iw, ih, 2 * i, 2 * j), at(in, iw, ih, 2 * i + 1, 2 * j),
			at(in, iw, ih, 2 * i, 2 * j + 1),
			at(in, iw, ih, 2 * i + 1, 2 * j + 1) };
		float m = 0;
		int cnt = 0;
		for (int k = 0; k < 4; k++)
			if (isfinite(v[k])) { m += v[k]; cnt++; }
		out[(size_t)ow * j + i] = cnt ? m / cnt : NAN;
	}
}

/* cell-centred bilinear x2 upsampling, simpois.c:199-239.  The sample
 * position (i - 0.5)/2 is formed in double and rounded to float; the four
 * weights are accumulated left to right as in evaluate_bilinear_cell. */
void orc_prolong(float *out, int ow, int oh, const float *in, int iw, int ih)
{
	for (int j = 0; j < oh; j++)
	for (int i = 0; i < ow; i++) {
		float x = (float)((i - 0.5) / 2);
		float y = (float)((j - 0.5) / 2);
		int ix = (int)floor(x), iy = (int)floor(y);
		float a = at(in, iw, ih, ix, iy), b = at(in, iw, ih, ix + 1, iy);
		float c = at(in, iw, ih, ix, iy + 1);
		float d = at(in, iw, ih, ix + 1, iy + 1);
		float fx = x - ix, fy = y - iy;
		float r = 0;
		r += a * (1 - fx) * (1 - fy);
		r += b * fx * (1 - fy);
		r += c * (1 - fx) * fy;
		r += d * fx * fy;
		out[(size_t)ow * j + i] = r;
	}
}

/* ---------------------------------------------------- masked CG operators */

typedef struct cg_sys {
	int w, h, n;
	int *unk;          /* raster index of each unknown (cleant_cgpois.c:6-25) */
	int *slot;         /* raster index -> unknown number or -1 (:27-38) */
	float *bnd;        /* known values, later zeroed (:218-236) */
	int bilap;
} cg_sys;

/* value of pixel (i,j) seen by the operator: the known data, or the
 * unknown rounded to float (the getter returns float), cleant_cgpois.c:48-65 */
static inline float cg_at(const cg_sys *s, const double *x, int i, int j)
{
	i = i < 0 ? 0 : (i >= s->w ? s->w - 1 : i);
	j = j < 0 ? 0 : (j >= s->h ? s->h - 1 : j);
	int q = s->slot[(size_t)j * s->w + i];
	return q < 0 ? s->bnd[(size_t)j * s->w + i] : (float)x[q];
}

/* y = A x.  Poisson: y = -(float stencil), cleant_cgpois.c:67-89.
 * Biharmonic: 13-point float stencil, cleant_cgpois.c:91-121. */
static void cg_apply(const cg_sys *s, double *y, const double *x)
{
	for (int p = 0; p < s->n; p++) {
		int i = s->unk[p] % s->w, j = s->unk[p] / s->w;
		float r;
		if (!s->bilap) {
			r = -4 * cg_at(s, x, i, j);
			r = r + cg_at(s, x, i + 1, j);
			r = r + cg_at(s, x, i, j + 1);
			r = r + cg_at(s, x, i - 1, j);
			r = r + cg_at(s, x, i, j - 1);
			y[p] = -(double)r;
		} else {
			r = 20 * cg_at(s, x, i, j);
			r = r - 8 * cg_at(s, x, i + 1, j);
			r = r - 8 * cg_at(s, x, i, j + 1);
			r = r - 8 * cg_at(s, x, i - 1, j);
			r = r - 8 * cg_at(s, x, i, j - 1);
			r = r + 2 * cg_at(s, x, i + 1, j + 1);
			r = r + 2 * cg_at(s, x, i + 1, j - 1);
			r = r + 2 * cg_at(s, x, i - 1, j + 1);
			r = r + 2 * cg_at(s, x, i - 1, j - 1);
			r = r + cg_at(s, x, i + 2, j);
			r = r + cg_at(s, x, i, j + 2);
			r = r + cg_at(s, x, i - 2, j);
			r = r + cg_at(s, x, i, j - 2);
			y[p] = (double)r;
		}
	}
}

/* sequential double dot product, minicg.c:12-18 */
static double dot(const double *a, const double *b, int n)
{
	double r = 0;
	for (int i = 0; i < n; i++)
		r += a[i] * b[i];
	return r;
}

/* conjugate gradient with absolute stopping rule, minicg.c:24-56.
 * Returns the number of iterations run. */
static int cg_run(const cg_sys *s, double *x, const double *b, const double *x0,
		int maxit, double stop, orc_trace *tr)
{
	int n = s->n, it = 0;
	double *r = orc_alloc(sizeof(double) * n);
	double *p = orc_alloc(sizeof(double) * n);
	double *q = orc_alloc(sizeof(double) * n);
	cg_apply(s, q, x0);
	for (int k = 0; k < n; k++) {
		x[k] = x0[k];
		r[k] = b[k] - q[k];
		p[k] = r[k];
	}
	for (it = 0; it < maxit; it++) {
		cg_apply(s, q, p);
		double pq = dot(q, p, n);
		double rr0 = dot(r, r, n);
		double alpha = rr0 / pq;
		for (int k = 0; k < n; k++) x[k] = x[k] + alpha * p[k];
		for (int k = 0; k < n; k++) r[k] = r[k] - alpha * q[k];
		double rr1 = dot(r, r, n);
		trace_push(tr, n, it, rr1);
		if (sqrt(rr1) < stop) { it++; break; }
		double beta = rr1 / rr0;
		for (int k = 0; k < n; k++) p[k] = r[k] + beta * p[k];
	}
	free(r); free(p); free(q);
	return it;
}

/* masked CG refinement of one plane, cleant_cgpois.c:197-264.
 * out may alias init (as poisson_rec passes u for both). */
int orc_cg_extend(float *out, const float *in, const float *dat, int w, int h,
		const float *init, int maxit, float eps, int bilap,
		const orc_knobs *kn, orc_trace *tr)
{
	size_t wh = (size_t)w * h;
	int n = 0;
	for (size_t k = 0; k < wh; k++) n += isnan(in[k]) ? 1 : 0;
	if (!n) {
		for (size_t k = 0; k < wh; k++)
			out[k] = isfinite(in[k]) ? in[k] : init[k];
		return 0;
	}
	cg_sys s = { w, h, n, orc_alloc(sizeof(int) * n),
		orc_alloc(sizeof(int) * wh), orc_alloc(sizeof(float) * wh), bilap };
	n = 0;
	for (size_t k = 0; k < wh; k++) {
		s.slot[k] = -1;
		if (isnan(in[k])) { s.slot[k] = n; s.unk[n++] = (int)k; }
		s.bnd[k] = in[k];
	}
	/* right-hand side: data term minus the known-pixel contribution */
	double *b = orc_alloc(sizeof(double) * n);
	double *z = calloc(n ? n : 1, sizeof(double));
	if (!z) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
	cg_apply(&s, b, z);
	for (int p = 0; p < n; p++) {
		float d = dat ? -dat[s.unk[p]] : 0;
		b[p] = d - b[p];
	}
	for (size_t k = 0; k < wh; k++)
		if (s.slot[k] < 0) s.bnd[k] = 0;
	double *x = orc_alloc(sizeof(double) * n);
	for (int p = 0; p < n; p++) z[p] = init[s.unk[p]];
	float e = kn && kn->cg_eps >= 0 ? (float)kn->cg_eps : eps;
	float mi = maxit == -1 ? (float)n : (float)maxit;
	int its = cg_run(&s, x, b, z, (int)mi, e, tr);
	for (size_t k = 0; k < wh; k++) out[k] = in[k];
	for (int p = 0; p < n; p++) out[s.unk[p]] = (float)x[p];
	free(b); free(z); free(x);
	free(s.unk); free(s.slot); free(s.bnd);
	return its;
}

/* ------------------------------------------------------- multiscale driver */

/* coarse-to-fine recursion on one plane, simpois.c:247-285 */
void orc_poisson_rec(float *u, const float *g, const float *f, int w, int h,
		float tstep, int niter, int scale, int cgit,
		const orc_knobs *kn, orc_trace *tr)
{
	size_t wh = (size_t)w * h;
	float *init = orc_alloc(sizeof(float) * wh);
	if (scale > 1 && (w > 1 || h > 1)) {
		int ws = (int)ceil(w / 2.0), hs = (int)ceil(h / 2.0);
		size_t n2 = (size_t)ws * hs;
		float *gs = orc_alloc(sizeof(float) * n2);
		float *fs = f ? orc_alloc(sizeof(float) * n2) : NULL;
		float *us = orc_alloc(sizeof(float) * n2);
		orc_restrict(gs, ws, hs, g, w, h);
		if (f) {
			orc_restrict(fs, ws, hs, f, w, h);
			double fac = kn ? kn->pmsfac : 3;
			for (size_t k = 0; k < n2; k++) fs[k] = fs[k] * fac;
		}
		orc_poisson_rec(us, gs, fs, ws, hs, tstep, niter, scale - 1, cgit,
				kn, tr);
		orc_prolong(init, w, h, us, ws, hs);
		free(gs); free(fs); free(us);
	} else {
		for (size_t k = 0; k < wh; k++) init[k] = 0;
	}
	double ponlit = kn ? kn->ponlit : 0;
	if (ponlit && ponlit != w)
		niter = cgit = 0;

	/* simpois.c:144-168 */
	for (size_t k = 0; k < wh; k++)
		u[k] = isfinite(g[k]) ? g[k] : init[k];
	orc_gs_sweeps(u, g, f, w, h, tstep, niter);
	free(init);

	/* simpois.c:278-284 */
	if (cgit && tstep > 0)
		orc_cg_extend(u, g, f, w, h, u, cgit, 1e-6f, 0, kn, tr);
	else
		orc_cg_extend(u, g, f, w, h, u, cgit, 1e-9f, 1, kn, tr);
}

/* per-channel loop over planar data, simpois.c:288-302 */
void orc_solver_separable(float *out, const float *in, const float *dat,
		int w, int h, int pd, float tstep, int niter, int scale,
		float cgrad, const orc_knobs *kn, orc_trace *tr)
{
	size_t wh = (size_t)w * h;
	for (int l = 0; l < pd; l++)
		orc_poisson_rec(out + wh * l, in + wh * l, dat ? dat + wh * l : NULL,
				w, h, tstep, niter, scale, (int)cgrad, kn, tr);
}

/* ------------------------------------------------------ residual (FP64) */

/* ||b - A u|| / ||b|| over the NaN pixels of g, for the masked system of
 * cleant_cgpois.c:225-236 (Poisson when bilap == 0).  A and b are formed in
 * double from float data (no float rounding of the stencil), so this is a
 * measurement of the solution, not a restatement of the reference's
 * float-evaluated operator.  Writes ||b|| to *bnorm when non-NULL. */
double orc_rel_residual(const float *u, const float *g, const float *f,
		int w, int h, int bilap, double *bnorm)
{
	double rr = 0, bb = 0;
	for (int j = 0; j < h; j++)
	for (int i = 0; i < w; i++) {
		size_t ij = (size_t)j * w + i;
		if (!isnan(g[ij])) continue;
		/* A u with all pixels, and A(0) with known pixels only */
		double au = 0, a0 = 0;
		static const int d5[5][3] = {{0,0,-4},{1,0,1},{0,1,1},{-1,0,1},{0,-1,1}};
		static const int d13[13][3] = {{0,0,20},{1,0,-8},{0,1,-8},{-1,0,-8},
			{0,-1,-8},{1,1,2},{1,-1,2},{-1,1,2},{-1,-1,2},{2,0,1},{0,2,1},
			{-2,0,1},{0,-2,1}};
		const int (*d)[3] = bilap ? d13 : d5;
		int nd = bilap ? 13 : 5;
		for (int k = 0; k < nd; k++) {
			int a = i + d[k][0], b = j + d[k][1];
			a = a < 0 ? 0 : (a >= w ? w - 1 : a);
			b = b < 0 ? 0 : (b >= h ? h - 1 : b);
			size_t q = (size_t)b * w + a;
			au += d[k][2] * (double)u[q];
			if (!isnan(g[q])) a0 += d[k][2] * (double)g[q];
		}
		if (!bilap) { au = -au; a0 = -a0; }
		double bv = (f ? -(double)f[ij] : 0) - a0;
		double ax = au - a0; /* A restricted to the unknowns */
		double res = bv - ax;
		rr += res * res;
		bb += bv * bv;
	}
	if (bnorm) *bnorm = sqrt(bb);
	return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}
