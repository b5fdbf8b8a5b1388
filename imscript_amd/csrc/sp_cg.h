// Batched FP32-vector CG (sp_cg.hip): device-side descriptors.
#pragma once
#include "sp_internal.h"

// one plane of the batched CG: full-grid float vectors, zero at known pixels
struct Cg32Plane {
	const float *g;    // boundary data, NaN = unknown
	const float *f;    // data term or nullptr
	float *u;          // initial guess in, solution out (unknowns only)
	double *x;         // solution accumulates in double: ~1e3 float updates
	                   // would otherwise cost ~1e-5 of relative residual
	float *r, *p, *q;
	float *p2;         // second direction buffer: the fused direction +
	                   // product kernel reads p_old from one, writes p_new
	                   // to the other, swapping every iteration
	float *z;          // preconditioned residual (== r without multigrid)
	unsigned char *code;   // per pixel: bit 0 unknown, bits 1-4 the E, S, W,
	                       // N neighbour is an in-image unknown, bits 5-7 the
	                       // number of in-image neighbours (set by cg32_setup)
	const int *cnt;    // number of unknowns (device)
	int w, h, bilap;   // 0 Poisson, 1 biharmonic, 2 the transposed biharmonic
	                   // operator (the adjoint's solve)
	int u0;            // the initial guess is 0 at every unknown (coarsest
	                   // pyramid level, no sweeps before the CG): u is not read
	int blk0, nblk;    // this plane's work blocks / partial-sum slots
	int t0, nt;        // this plane's level-0 multigrid tiles / their slots
};

// per-plane scalars (device)
struct Cg32Scal {
	double rr, rz, alpha, beta, stop, eps, bnorm;
	int rel;           // stop = eps * ||b|| instead of eps
	int active, iters, maxit;
	int maxit_from_count;   // maxit = number of unknowns (the -c -1 rule)
	int trace0, trace_cap;  // slots of this plane in the rr trace
	int upd;           // this iteration updated r: the direction kernel
	                   // adds alpha p to x before it replaces p
	int fin;           // this iteration stopped the plane: the fused
	                   // direction kernel writes the solution to u
	int status;        // why the plane stopped: 0 still running, else
	                   // SP_PLANE_CONVERGED / _MAXIT / _BREAKDOWN / _NONFINITE
	double alpha_prev; // the alpha before the current one (the fused
	                   // direction kernel adds alpha p to x in pairs)
};

// Fused multigrid tile kernels: levels 0 and 1 run one of the fine tile
// shapes, levels 2..MG_FUSE_MAX one of the coarse shapes; a shape fixes the
// number nu of red-black sweeps before and after the coarse correction
// (the halo is 2 nu cells).  A launch moves its region through HBM once
// whatever nu is, so stronger smoothing there costs compute only and saves
// V-cycles; the levels below them (per-level kernels, tail) smooth
// MG_NU_TAIL times, where more sweeps change nothing.
constexpr int MG_NU_TAIL = 2;
// tile shapes (width, height, nu, threads).  The fewer cells a thread owns,
// the shorter a tile's chain of dependent half sweeps, which is what bounds
// the small levels (few tiles: latency); larger tiles recompute less halo.
// A tile's interior is exact whatever its size (it recomputes its halo),
// so shapes of the same nu give the same bits.
constexpr int MG_NFINE = 6;
constexpr int MG_FINE_TX[MG_NFINE] = {120, 120, 116, 112, 108, 104};
constexpr int MG_FINE_TY[MG_NFINE] = {56, 56, 52, 48, 44, 40};
constexpr int MG_FINE_NU[MG_NFINE] = {2, 2, 3, 4, 5, 6};
constexpr int MG_FINE_T[MG_NFINE] = {512, 1024, 1024, 1024, 1024, 1024};
constexpr int MG_NGEO = 6;
constexpr int MG_GEO_TX[MG_NGEO] = {112, 112, 112, 48, 104, 40};
constexpr int MG_GEO_TY[MG_NGEO] = {32, 32, 48, 16, 40, 8};
constexpr int MG_GEO_NU[MG_NGEO] = {4, 4, 4, 4, 6, 6};
constexpr int MG_GEO_T[MG_NGEO] = {256, 512, 1024, 256, 1024, 256};

// one plane at one level of the multigrid preconditioner
struct MgLevel {
	const float *g;        // level 0: boundary data, NaN = unknown; else null
	unsigned char *code;   // level 0: neighbour code per pixel (see sp_cg.hip)
	float *D, *wE, *wS;    // levels > 0: Galerkin coefficients, D = 0 inactive
	unsigned *gcode;       // levels 1..MG_FUSE_MAX: the same coefficients packed
	                       // (one word per cell to MG_G32_MAX, then two)
	float *z, *b;          // correction and right-hand side
	float *zd;             // fused levels: z after the pre-smoothing
	int w, h;
};

// levels 1..MG_FUSE_MAX run the fused tile kernels.  Their coefficients
// are integers: a coupling of level l is at most 2^l (the edges between two
// 2^l x 2^l aggregates), the diagonal at most 4^(l+1).  Levels to
// MG_G32_MAX pack them in 32 bits (diagonal in 11 bits, couplings in 5
// bits each); the levels after them in 64 bits (couplings in a byte each,
// <= 2^7, and the diagonal as a float).
constexpr int MG_G32_MAX = 4;
constexpr int MG_FUSE_MAX = 7;

// work lists of one multigrid level (device arrays indexed by batch plane)
struct MgLevelWork {
	MgLevel *lv;           // descriptor of each plane at this level
	MgLevel *coarse;       // descriptor at the next level (planes that have it)
	BlockWork *mid;        // cells of planes that have a coarser level
	BlockWork *last;       // cells of planes for which this is the coarsest
	BlockWork *cells_c;    // coarse cells (next level) of the mid planes
	BlockWork *build_c;    // coarse cells (next level) of every plane that has it
	int4 *tiles;           // fused levels: tiles (plane, x0, y0, -) of the
	                       // planes that run this level outside the tail
	int geo;               // fused levels: tile shape (MG_FINE_* at level 1,
	                       // MG_GEO_* from level 2)
	int nmid, nlast, ncells_c, nbuild_c, ntiles;
	double cells;          // cells covered by `tiles` (profiler bytes)
};

// per plane: the levels [first, depth) run inside one work-group
constexpr int MG_MAXLEV = 16;
constexpr int MG_TAIL_CELLS = 4096;    // first level at or below this size
constexpr int MG_TAIL_T = 1024;
// LDS budget of the tail kernel: z, b, D, wE, wS of every tail level
constexpr int MG_TAIL_LDS_MAX = 155000;

// optional timer around individual launches (profiling runs only; never
// set while a graph is being captured)
typedef void (*ProfFn)(const char *name, int begin, double bytes, hipStream_t s);

struct Cg32Batch {
	Cg32Plane *planes;
	Cg32Scal *scal;
	BlockWork *bw;
	double *part;      // 2 per work block
	double *trace;     // optional rr_new trace
	int *nactive;      // planes still iterating
	int nplanes, nblk;
	int bilap;         // biharmonic: separate direction and product kernels
	// multigrid preconditioner (nlev == 0: plain CG)
	MgLevelWork *mg;   // host array of nlev levels
	int nlev, nu, coarse_sweeps;   // nu: per-level kernels and the tail
	float wc;          // coarse-correction factor
	MgLevel *tail_lv;  // [plane * MG_MAXLEV + level]
	int4 *tail_rng;    // per plane: (first tail level, depth, runs in LDS,
	                   // tail cells of all its levels)
	int tail_lds;      // dynamic LDS bytes of the LDS tail kernel (0: none)
	int tail_global;   // some plane runs the global-memory tail
	int4 *tiles0;      // level-0 tiles (plane, x0, y0, -), contiguous per plane
	int ntiles0;
	int geo0;          // their shape (MG_FINE_*)
	double *tpart;     // one partial per level-0 tile
	double px0;        // level-0 pixels of the batch (profiler bytes)
	ProfFn prof;
};

namespace spk {
void cg32_setup(const Cg32Batch &B, hipStream_t s);
// one iteration; par: parity of the iteration in its captured chunk (the
// direction buffer it reads), chunks are even
void cg32_iter(const Cg32Batch &B, hipStream_t s, int par);
void cg32_finish(const Cg32Batch &B, hipStream_t s);
void mg_build(const Cg32Batch &B, hipStream_t s);
// allow `bytes` of dynamic LDS for the LDS tail kernel; 0 on success
int mg_tail_lds_enable(int bytes);
}
