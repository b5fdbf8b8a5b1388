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
	float *z;          // preconditioned residual (== r without multigrid)
	unsigned char *code;   // per pixel: bit 0 unknown, bits 1-4 the E, S, W,
	                       // N neighbour is an in-image unknown, bits 5-7 the
	                       // number of in-image neighbours (set by cg32_setup)
	const int *cnt;    // number of unknowns (device)
	int w, h, bilap;
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
};

// Fused multigrid tile kernels: levels 0 and 1 smooth nu = 2 times on
// 120x56 tiles (512 threads), levels 2..MG_FUSE_MAX nu = 4 times on 112x32
// tiles (256 threads); the halo is 2 nu cells, the region 128 cells wide.  Those levels are
// cheap (a sixteenth of the pixels and less), and the stronger smoothing
// there saves V-cycles (14 -> 11 CG iterations to 5e-6 on 50% rectangle
// holes); the levels below them (per-level kernels, tail) smooth nu = 2
// times, where more sweeps change nothing.
constexpr int MG_NU_FINE = 2, MG_NU_COARSE = 4;
constexpr int MG_TX_FINE = 120, MG_TY_FINE = 56, MG_T_FINE = 512;
constexpr int MG_TX_COARSE = 112, MG_TY_COARSE = 32, MG_T_COARSE = 256;

// one plane at one level of the multigrid preconditioner
struct MgLevel {
	const float *g;        // level 0: boundary data, NaN = unknown; else null
	unsigned char *code;   // level 0: neighbour code per pixel (see sp_cg.hip)
	float *D, *wE, *wS;    // levels > 0: Galerkin coefficients, D = 0 inactive
	unsigned *gcode;       // levels 1..MG_FUSE_MAX: the same coefficients packed
	float *z, *b;          // correction and right-hand side
	int w, h;
};

// levels 1..MG_FUSE_MAX run the fused tile kernels: their coefficients are
// integers small enough to pack in 32 bits (diagonal <= 4^(l+1) <= 1024 in
// 11 bits, couplings <= 2^l <= 16 in 5 bits each)
constexpr int MG_FUSE_MAX = 4;

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
	double *tpart;     // one partial per level-0 tile
	double px0;        // level-0 pixels of the batch (profiler bytes)
	ProfFn prof;
};

namespace spk {
void cg32_setup(const Cg32Batch &B, hipStream_t s);
void cg32_iter(const Cg32Batch &B, hipStream_t s);
void cg32_finish(const Cg32Batch &B, hipStream_t s);
void mg_build(const Cg32Batch &B, hipStream_t s);
// allow `bytes` of dynamic LDS for the LDS tail kernel; 0 on success
int mg_tail_lds_enable(int bytes);
}
