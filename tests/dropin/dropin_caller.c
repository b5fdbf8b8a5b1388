/*
 * A reference-side caller of the drop-in library: a plain C program that
 * includes include/simpois_hip.h and links libsimpois_hip.so the way a
 * program built against imscript links simpois.o (simpois.c:247-302) and
 * cleant_cgpois.c (:266-278).  The GPU tests drive it on the reference's
 * golden cases.
 *
 *   dropin_caller separable W H PD in.f32 dat.f32|- out.f32 T N S C
 *       poisson_solver_separable(out, in, dat, W, H, PD, T, N, S, C)
 *   dropin_caller rec W H 1 in.f32 dat.f32|- out.f32 T N S C
 *       poisson_rec(out, in, dat, W, H, T, N, S, C)
 *   dropin_caller pcg|bcg W H 1 in.f32 dat.f32|- out.f32 init.f32 MAXIT EPS
 *       poisson_extension_by_cg / biharmonic_extension_by_cg(out, in, dat,
 *       W, H, init, MAXIT, EPS)
 *
 * Files hold raw native float32 samples, planar.  Exit status 0 on success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "simpois_hip.h"

static float *load(const char *path, size_t n)
{
	if (!strcmp(path, "-")) return NULL;
	FILE *f = fopen(path, "rb");
	if (!f) { perror(path); exit(2); }
	float *x = malloc(n * sizeof *x);
	if (!x || fread(x, sizeof *x, n, f) != n) {
		fprintf(stderr, "%s: short read\n", path);
		exit(2);
	}
	fclose(f);
	return x;
}

static void save(const char *path, const float *x, size_t n)
{
	FILE *f = fopen(path, "wb");
	if (!f || fwrite(x, sizeof *x, n, f) != n) { perror(path); exit(2); }
	fclose(f);
}

int main(int argc, char **argv)
{
	if (argc < 8) {
		fprintf(stderr, "usage: %s separable|rec|pcg|bcg W H PD in dat out ...\n",
				argv[0]);
		return 2;
	}
	const char *fn = argv[1];
	const int w = atoi(argv[2]), h = atoi(argv[3]), pd = atoi(argv[4]);
	const size_t n = (size_t)w * h * pd;
	float *in = load(argv[5], n), *dat = load(argv[6], n);
	float *out = calloc(n, sizeof *out);
	if (!in || !out) return 2;
	if (!strcmp(fn, "separable") || !strcmp(fn, "rec")) {
		if (argc != 12) return 2;
		const float t = atof(argv[8]);
		const int niter = atoi(argv[9]), scale = atoi(argv[10]);
		if (fn[0] == 's')
			poisson_solver_separable(out, in, dat, w, h, pd, t, niter, scale,
					atof(argv[11]));
		else
			poisson_rec(out, in, dat, w, h, t, niter, scale, atoi(argv[11]));
	} else if (!strcmp(fn, "pcg") || !strcmp(fn, "bcg")) {
		if (argc != 11) return 2;
		float *init = load(argv[8], n);
		if (!init) return 2;
		if (fn[0] == 'p')
			poisson_extension_by_cg(out, in, dat, w, h, init, atoi(argv[9]),
					atof(argv[10]));
		else
			biharmonic_extension_by_cg(out, in, dat, w, h, init, atoi(argv[9]),
					atof(argv[10]));
		free(init);
	} else {
		return 2;
	}
	save(argv[7], out, n);
	free(in);
	free(dat);
	free(out);
	return 0;
}
