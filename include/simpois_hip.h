/*
 * simpois_hip.h -- C ABI of the MI355X (gfx950) implementation of imscript's
 * simpois hole-filling path.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements with the reference's exact signatures and
 *     semantics.  Buffers are host memory, planar (plane l at +w*h*l), and
 *     the call is synchronous.  Failures print a message and exit(1), as the
 *     reference's xmalloc does (simpois.c:10-19).
 *
 *  2. sp_* entry points with status codes that never exit, working on host
 *     or device buffers, plus the additions that have no reference
 *     counterpart (batched solve, adjoint, residual).
 *
 * Environment knobs read once per process, as SMART_PARAMETER_SILENT does
 * (smapa.h:44-60): PMSFAC (default 3), PONLIT (0), CG_EPS (-1).
 * SIMPOIS_HIP_QUIET=1 suppresses the per-iteration CG trace on stderr.
 */
#ifndef SIMPOIS_HIP_H
#define SIMPOIS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ drop-in */

/* Fill the NaN pixels of each plane of `in` (w*h*pd floats, planar) and
 * write the result to `out`.  `dat` is the Poisson data term (same layout)
 * or NULL.  tstep > 0: Laplace/Poisson Gauss-Seidel with step tstep;
 * tstep < 0: biharmonic; tstep < -1000: triharmonic.  niter sweeps at each
 * of at most `scale` pyramid levels; cgrad CG iterations per level (-1: as
 * many as unknowns).  `out` must not alias `in`.
 * Replaces simpois.c:288-290 (poisson_solver_separable). */
void poisson_solver_separable(float *out, float *in, float *dat,
		int w, int h, int pd,
		float tstep, int niter, int scale, float cgrad);

/* One plane of the above.  Replaces simpois.c:247-248 (poisson_rec). */
void poisson_rec(float *u, float *g, float *f, int w, int h,
		float tstep, int niter, int scale, int cgit);

/* Masked CG refinement of one plane from the initial guess `init`
 * (Poisson: -Laplacian; biharmonic: 13-point bilaplacian), absolute stop
 * sqrt(r.r) < eps (CG_EPS overrides eps), maxit == -1: as many iterations
 * as unknowns.  `out` may alias `init`.
 * Replace cleant_cgpois.c:266-271 and :273-278. */
void poisson_extension_by_cg(float *out, float *in, float *dat, int w, int h,
		float *init, int maxit, float eps);
void biharmonic_extension_by_cg(float *out, float *in, float *dat, int w, int h,
		float *init, int maxit, float eps);

/* The simpois command line (flags -i -o -m -f -t -n -s -c, NPY/PFM I/O).
 * Replaces simpois.c:351-439 (main_simpois). */
int main_simpois(int argc, char *argv[]);

/* AMLE hole filling of one w x h plane: y = x with the NaN pixels filled by
 * AMLE_NITER (default 100) raster-order sweeps of the AMLE_NN (default 4)
 * neighbour scheme from the plane's maximum, or from the AMLE_INIT image at
 * the holes; the diagnostic line every tenth sweep on stderr
 * (SIMPOIS_HIP_QUIET=1 omits it).  Synchronous; failures print a message
 * and exit.  AMLE_NN outside 1..32 is refused.
 * Replaces amle.c:203-246 (amle). */
void amle(float *y, float *x, int w, int h);

/* The amle command line: amle [in [mask [out]]] (NPY/PFM I/O).
 * Replaces amle.c:325-366 (main_amle). */
int main_amle(int argc, char *argv[]);

/* Explicit linear diffusion of heatd: n steps of y = x + t * L(x) on each
 * of pd planar w x h planes, L the 9-point scheme of the diffusion matrix
 * (a, b, c) of each pixel divided by 4 (A: w*h*3 interleaved, shared by
 * all planes; NULL: the identity, the 5-point Laplacian), clamped borders.
 * Bit-identical to the reference.  As there, x ends equal to y; with
 * n <= 0 neither is written.  Synchronous; failures print a message and
 * exit.  Replace heatd.c:90-103 (linear_diffusion_separable), :72-88
 * (linear_diffusion) and :58-70 (linear_diffusion_step, x unchanged). */
void linear_diffusion_separable(float *y, float *x, int w, int h, int pd,
		float *A, int n, float t);
void linear_diffusion(float *y, float *x, int w, int h, float *A, int n, float t);
void linear_diffusion_step(float *y, float *x, int w, int h, float *A, float t);

/* The heatd command line: heatd [-n 1] [-t 0.25] [-A field] [in [out]].
 * Replaces heatd.c:109-150 (main). */
int main_heatd(int argc, char *argv[]);

/* ------------------------------------------------------- status API */

enum {
	SP_OK = 0,
	SP_ERR_ARG = 1,     /* invalid argument (sizes, aliasing, NULL) */
	SP_ERR_NOMEM = 2,   /* device or host allocation failed */
	SP_ERR_HIP = 3,     /* HIP runtime error */
	SP_ERR_NODEV = 4    /* no usable gfx950 device */
};

typedef struct sp_params {
	float tstep;     /* -t, default 0.25 */
	int niter;       /* -n, default 10 */
	int scale;       /* -s, default 99 */
	int cgit;        /* -c, default 0 */
	double pmsfac;   /* PMSFAC */
	double ponlit;   /* PONLIT */
	double cg_eps;   /* CG_EPS (< 0: built-in 1e-6 / 1e-9) */
	int trace;       /* print the CG trace to stderr like minicg.c:46 */
	/* additions (no reference counterpart) */
	int cg_fp32;     /* 1: CG with float vectors and double reductions,
	                    all planes batched; 0: double vectors (parity) */
	double cg_rel;   /* > 0: CG stops at sqrt(r.r) < cg_rel * ||b|| per plane
	                    (FP32 CG only); 0: absolute threshold as minicg.c:47 */
	int cg_chunk;    /* CG iterations per captured graph (FP32 CG) */
	int cg_mg;       /* 1: precondition the FP32 CG with a multigrid V-cycle
	                    (Poisson only); 0: plain CG as the reference */
	int autotune;    /* 1: time every compiled Gauss-Seidel tile shape on the
	                    first sweep of each level size and keep the fastest
	                    (cached per process; results do not depend on it);
	                    0: built-in shapes.  Default from SIMPOIS_HIP_AUTOTUNE */
	int guidance;    /* 0: f is the data term, as in the reference; 1 or 2: f
	                    holds a source image v and the data term is plambda's
	                    "x,l" (1) or "x,lr" (2) Laplacian of v with the
	                    nearest-pixel rule, computed inside the solve (the
	                    Poisson-editing pipeline plambda v "x,l" | simpois -f) */
	double cg_rel_coarse; /* > 0: relative CG tolerance of the pyramid levels
	                    coarser than the finest, whose solutions only start
	                    the next level (FP32 CG); 0: cg_rel at every level */
} sp_params;

/* one image of a batch: planar w*h*pd device buffers, f may be NULL */
typedef struct sp_image {
	int w, h, pd;
	const float *g;
	const float *f;
	float *out;
} sp_image;

typedef struct sp_stats {
	int levels;            /* pyramid levels of the solve (the deepest plane) */
	int cg_iters_finest;   /* CG iterations at the finest level, the maximum
	                          over the planes of the call */
	double elapsed_ms;     /* device time of the solve */
} sp_stats;

/* Why a plane of a solve stopped (sp_plane_stats.status). */
enum {
	SP_PLANE_SWEPT = 0,      /* no CG asked (-c 0): the fixed sweeps ran */
	SP_PLANE_CONVERGED = 1,  /* the CG met its stop test, or had nothing to
	                            solve (no holes, or a zero initial residual) */
	SP_PLANE_MAXIT = 2,      /* the CG stopped at its iteration limit */
	SP_PLANE_BREAKDOWN = 3,  /* the CG stopped on p.Ap <= 0 (round-off); the
	                            iterate of that point is kept */
	SP_PLANE_NONFINITE = 4   /* the residual is not finite: NaN or Inf in the
	                            data the plane's system reads.  The FP32 CG
	                            stops the plane there and keeps its initial
	                            guess; the parity CG (FP64 vectors) runs on
	                            like the reference and writes what it gets */
};

/* Per-plane outcome of a batched solve (one entry per plane: the planes of
 * image 0, then of image 1, ...).  Planes never interact: a plane that
 * fails leaves the others' bits unchanged. */
typedef struct sp_plane_stats {
	int status;            /* SP_PLANE_* of the finest level; NONFINITE also
	                          when the output's residual is not finite */
	int levels;            /* pyramid levels of this plane */
	int holes;             /* NaN pixels of the plane (unknowns) */
	int sweeps;            /* Gauss-Seidel sweeps run at the finest level */
	int cg_iters;          /* CG iterations at the finest level */
	double cg_rr;          /* the CG's last r.r at the finest level (its own
	                          recurrence); -1 without CG */
	double rel_residual;   /* FP64 ||b - A u|| / ||b|| of the output over
	                          the holes, the system of cleant_cgpois.c:225-236
	                          (Poisson, or biharmonic for tstep < 0); -1 for
	                          the triharmonic operator (tstep < -1000) */
} sp_plane_stats;

/* Defaults of the reference (simpois.c:357-360) with the env knobs. */
void sp_params_default(sp_params *p);

/* Text of the last error of the calling thread. */
const char *sp_last_error(void);

/* Host-buffer solve with a status code (same semantics as
 * poisson_solver_separable). */
int sp_solve_host(float *out, const float *in, const float *dat,
		int w, int h, int pd, const sp_params *p, sp_stats *st);

/* Device-buffer solve on `stream` (a hipStream_t, NULL = default).
 * d_out, d_in, d_dat are device pointers, planar; d_dat may be NULL.
 * The call returns when the work is queued and the result is complete on
 * the stream, except for the CG stage, which synchronises the stream to
 * test for convergence. */
int sp_solve_device(float *d_out, const float *d_in, const float *d_dat,
		int w, int h, int pd, const sp_params *p, void *stream,
		sp_stats *st);

/* Batched device solve: every plane of every image runs its own pyramid,
 * all advancing together level by level (levels aligned at the finest),
 * with per-plane CG convergence.  No reference counterpart: the reference
 * solves one image per process (doc/misc/optimization.txt:44-50). */
int sp_solve_batch_device(int n, const sp_image *imgs, const sp_params *p,
		void *stream, sp_stats *st);

/* sp_solve_batch_device, also filling ps[0 .. total planes) with the
 * outcome of each plane (ps may be NULL).  The residuals cost one more
 * read of the finest level.  No reference counterpart. */
int sp_solve_batch_device_ex(int n, const sp_image *imgs, const sp_params *p,
		void *stream, sp_stats *st, sp_plane_stats *ps);

/* The CG history of the last solve in the process at one plane (0-based,
 * planes of image 0 first) and pyramid level (0 = finest): the r.r values
 * after each iteration, as minicg.c:46 prints them, at most cap of them
 * into rr, and the level's unknowns into *n.  Returns the number of
 * iterations, -1 for a plane or level the last solve did not have.  The
 * FP64 (parity) CG always records it; the FP32 CG when sp_params.trace is
 * set.  No reference counterpart. */
int sp_last_cg_trace(int plane, int level, int *n, double *rr, int cap);

/* Solve plans (buffers, tile lists, captured graphs) are cached by call
 * signature (sizes, buffer addresses, parameters), least recently used
 * first out: at most max_plans plans and max_bytes device bytes (defaults
 * 256 and 16 GiB).  A plan that cannot be allocated empties the cache and
 * is tried once more.  sp_plan_cache_clear frees every cached plan.  No
 * reference counterpart. */
void sp_plan_cache_limit(int max_plans, size_t max_bytes);
void sp_plan_cache_clear(void);

/* Cap on the device bytes the library holds (plans in the cache included);
 * an allocation beyond it fails with SP_ERR_NOMEM as a failed hipMalloc
 * does.  0 (the default): no cap.  Returns the previous cap.
 * sp_device_bytes: the bytes held now.  No reference counterpart. */
size_t sp_set_alloc_cap(size_t bytes);
size_t sp_device_bytes(void);

/* one image of an adjoint batch: planar w*h*pd device buffers */
typedef struct sp_adjoint_image {
	int w, h, pd;
	const float *g;        /* forward input, NaN = hole */
	const float *grad_u;   /* dL/du of the filled output */
	float *grad_g;         /* dL/dg: v at known pixels plus the adjoint
	                          flux from nearby holes; 0 at holes */
	float *grad_f;         /* dL/df: -lambda at holes, 0 elsewhere; may be NULL */
} sp_adjoint_image;

/* Gradient of the fill made with the parameters in p.
 * p->cgit != 0: the converged fill.  tstep > 0 (clamped 5-point operator,
 *   symmetric): the adjoint is the masked Laplacian solve A lambda = grad_u
 *   on the holes, run by the same pyramid and CG as the forward solve.
 *   tstep < 0 (the 13-point bilaplacian, not symmetric within two pixels of
 *   the border): A^T lambda = grad_u with the transposed operator, by the
 *   FP32 CG (x in FP64) at every pyramid level without sweeps, to the
 *   tolerance in p; grad_g = grad_u - A_hk^T lambda on the known pixels.
 * p->cgit == 0 (tstep > 0 only): simpois's own output (niter Gauss-Seidel
 *   sweeps at each of at most `scale` pyramid levels, PMSFAC, PONLIT),
 *   differentiated exactly:
 *   transposed sweeps (run as forward sweeps of the image turned by 180
 *   degrees), transposed prolongation and restrictions.  The sweeps are
 *   linear with coefficients fixed by the mask, so no forward iterate is
 *   stored or recomputed.  Data finite or NaN.
 * No reference counterpart. */
int sp_adjoint_batch_device(int n, const sp_adjoint_image *imgs,
		const sp_params *p, void *stream, sp_stats *st);

/* Profiler: sp_profile(1) clears and enables event timing of the solver
 * phases ("gs_tile", "gs_tile_r23", "cg32", "cg64", "pyramid");
 * sp_profile_read sums device milliseconds, kernel launches and
 * algorithmic bytes of a phase. */
void sp_profile(int on);
int sp_profile_read(const char *name, double *ms, long *launches, double *bytes);

/* FP64 relative residual ||b - A u|| / ||b|| of each plane over the NaN
 * pixels of g, for the masked Poisson (bilap = 0) or biharmonic system of
 * cleant_cgpois.c:225-236.  Device pointers; writes pd doubles to host
 * array `rel`. */
int sp_residual_device(const float *d_u, const float *d_g, const float *d_f,
		int w, int h, int pd, int bilap, double *rel, void *stream);

/* sp_residual_device without any allocation: the same kernels and the same
 * results (bit for bit), with the work-block descriptors, per-block partial
 * sums and per-plane totals in a fixed 4 MiB device array that each GPU the
 * library runs on holds until the process ends (planes up to about 170 M
 * pixels in total; larger calls fail with SP_ERR_ARG).  It calls no
 * hipMalloc, hipFree or other allocating routine.  Every call on one GPU
 * uses that GPU's array, so calls that overlap on one GPU (other threads,
 * other streams) race.  Synchronises `stream` before returning.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc). */
int sp_residual_device_noalloc(const float *d_u, const float *d_g, const float *d_f,
		int w, int h, int pd, int bilap, double *rel, void *stream);

/* plambda's 3x3 Laplacian image operator on pd planar w x h float planes
 * (the producer of simpois's guidance field; replaces `plambda in "x,l"`,
 * plambda.c:2589-2605 stencils and :2680-2689 apply_3x3_stencil, on one
 * plane layout): out = sum over k = 0..8, in that order, of s[k] *
 * P(in, i-1+k%3, j-1+k/3), bit-identical to the reference.  variant 0: "x,l"
 * {0,1,0, 1,-4,1, 0,1,0}; 1: "x,lr" {.5,0,.5, 0,-2,0, .5,0,.5}.  boundary:
 * the PLAMBDA_GETPIXEL rule (plambda.c:2573-2588, getpixel.c): 0 zero,
 * 1 nearest (plambda's default), 2 reflection, 3 periodic, 4 NaN.  out must
 * not alias in.  The device form is asynchronous on `stream`. */
int sp_plambda_lap_device(float *d_out, const float *d_in, int w, int h, int pd,
		int variant, int boundary, void *stream);
int sp_plambda_lap_host(float *out, const float *in, int w, int h, int pd,
		int variant, int boundary);

/* AMLE hole filling (amle.c:203-246) of pd planar w x h planes, each on its
 * own: the NaN pixels of d_in get niter raster-order sweeps of the
 * nn-neighbour AMLE update (nn = AMLE_NN, 1..32; the neighbour table of
 * amle.c:16-29), started from the plane's maximum or, where d_init is not
 * NULL, from d_init at the holes (AMLE_INIT).  Bit-identical to the
 * reference's output.  trace = 1 also sweeps the lower bound tinf and prints
 * the reference's diagnostic line every tenth sweep (amle.c:230-233, one
 * line per plane); the output does not depend on it.  d_out may alias d_in.
 * Device pointers; asynchronous on `stream` without the trace.  The host
 * form copies through device memory and returns when done. */
int sp_amle_device(float *d_out, const float *d_in, const float *d_init,
		int w, int h, int pd, int nn, int niter, int trace, void *stream);
int sp_amle_host(float *out, const float *in, const float *init,
		int w, int h, int pd, int nn, int niter, int trace);

/* heatd's diffusion (as linear_diffusion_separable) with a status code:
 * d_out receives n steps from d_in, which is not written (n = 0 copies
 * it); d_A is the w*h*3 interleaved field or NULL.  Device pointers,
 * asynchronous on `stream` (scratch for n > 1 comes from the stream-ordered
 * allocator).  The host form copies through device memory. */
int sp_heatd_device(float *d_out, const float *d_in, const float *d_A,
		int w, int h, int pd, int n, float t, void *stream);
int sp_heatd_host(float *out, const float *in, const float *A,
		int w, int h, int pd, int n, float t);

/* dst[0..n) = src[0..n) on the device with a plain streaming kernel
 * (k_stream_copy): the bandwidth ceiling the solver kernels are compared
 * with, and the known-size reference of the HBM traffic counters.  n a
 * multiple of 4, both buffers 16-byte aligned; asynchronous on `stream`.
 * No reference counterpart. */
int sp_stream_copy(float *d_dst, const float *d_src, size_t n, void *stream);

/* Choices of the tile-shape autotuner so far in this process, one text
 * line per tuned level size: "gs R=<radius> T=<sweeps> planes=<n>
 * px=<pixels> choice=<BIxBJxBT:ms>,..." with the chosen shape prefixed '*'.
 * Writes at most cap bytes (NUL-terminated) and returns the size needed.
 * No reference counterpart. */
int sp_autotune_report(char *buf, int cap);

/* FP64 CG levels solved in this process by the persistent launch (all
 * iterations of a level in one launch), and those whose launch found its
 * work-groups not all resident at its start and left the iterations to
 * separate launches (same results).  Either pointer may be NULL.
 * No reference counterpart. */
void sp_cg_persist_counts(long *ran, long *gave_up);

#ifdef __cplusplus
}
#endif

#endif /* SIMPOIS_HIP_H */
