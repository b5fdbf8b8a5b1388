// Host driver of the simpois path: pyramid plan, Gauss-Seidel scheduling,
// CG drivers, profiler and the C ABI (include/simpois_hip.h).
//
// Call structure restated from the reference (simpois.c:247-302):
//   per plane: restrict g (and f, times PMSFAC) down the pyramid; from the
//   coarsest level up: u = isfinite(g) ? g : prolong(coarse output);
//   niter Gauss-Seidel sweeps; CG when -c is non-zero; the level's output
//   is the input at non-NaN pixels when the level has holes.
// Here every plane of every image of a call advances together, level by
// level (levels aligned at the finest), on one HIP stream.  Planes never
// interact, so each result equals the reference's channel-by-channel loop.
#include "sp_internal.h"
#include "sp_cg.h"
#include "sp_tiles.h"
#include "sp_amle.h"
#include "../../include/simpois_hip.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <mutex>
#include <string>
#include <vector>

// ------------------------------------------------------------- errors

static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

extern "C" const char *sp_last_error(void) { return g_err.c_str(); }

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	return set_err(e_ == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, \
		"%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
	} while (0)
#define TRY(x) do { int r_ = (x); if (r_) return r_; } while (0)

// ------------------------------------------------------ env knobs (once)

// getenv + sscanf("%lf") on first use, kept for the process, as
// SMART_PARAMETER_SILENT does (smapa.h:44-60)
static double env_knob(const char *name, double dflt)
{
	static std::mutex mu;
	static std::vector<std::pair<std::string, double>> known;
	std::lock_guard<std::mutex> lk(mu);
	for (auto &kv : known)
		if (kv.first == name) return kv.second;
	double v = dflt, y;
	const char *s = getenv(name);
	if (s && sscanf(s, "%lf", &y) == 1) v = y;
	known.emplace_back(name, v);
	return v;
}

extern "C" void sp_params_default(sp_params *p)
{
	memset(p, 0, sizeof *p);
	p->tstep = 0.25f;
	p->niter = 10;
	p->scale = 99;
	p->cgit = 0;
	p->pmsfac = env_knob("PMSFAC", 3);
	p->ponlit = env_knob("PONLIT", 0);
	p->cg_eps = env_knob("CG_EPS", -1);
	const char *q = getenv("SIMPOIS_HIP_QUIET");
	p->trace = !(q && *q && strcmp(q, "0"));
	p->cg_fp32 = 0;
	p->cg_rel = 0;
	p->cg_rel_coarse = 0;
	p->cg_chunk = 32;
	p->cg_mg = 0;
	const char *at = getenv("SIMPOIS_HIP_AUTOTUNE");
	p->autotune = at && *at && strcmp(at, "0");
}

// ------------------------------------------------------------- device

static int g_dev_checked = 0, g_dev_ok = 0, g_max_lds = 65536;
// LDS the whole-level R = 1 kernel by unknown pixel may use (0: path off);
// SIMPOIS_HIP_GS_CELLS=0 turns it off (such levels then go to k_gs_small
// or to tiles)
static int g_gsc_lds = 0;
static constexpr int GSC_LDS_MAX = 155000;
static const int g_gsc_off = [] {
	const char *e = getenv("SIMPOIS_HIP_GS_CELLS");
	return e && !strcmp(e, "0") ? 1 : 0;
}();
// Gauss-Seidel tile shape of a plan without the autotuner: variant 0 of
// the radius, or SIMPOIS_HIP_GS_SHAPE=v (read when a plan is built) when v
// is a compiled variant of it
static int gs_default_var(int R)
{
	const char *e = getenv("SIMPOIS_HIP_GS_SHAPE");
	if (!e || !*e) return 0;
	char *end = nullptr;
	const long v = strtol(e, &end, 10);
	if (*end || v < 0 || v >= sp_shape_count(R)) return 0;
	return (int)v;
}

// A level's Gauss-Seidel tiles run in one launch ordered by their
// dependences (k_gs_flow); SIMPOIS_HIP_GS_FLOW=0 makes it one launch per
// wavefront of tiles (k_gs_tile)
static const int g_gs_flow = [] {
	const char *e = getenv("SIMPOIS_HIP_GS_FLOW");
	return e && !strcmp(e, "0") ? 0 : 1;
}();
// SIMPOIS_HIP_CG_PERSIST=0 (read once) runs the FP64 CG's iterations as
// three launches each (captured in graphs) instead of one persistent launch
static const int g_cg_persist = [] {
	const char *e = getenv("SIMPOIS_HIP_CG_PERSIST");
	return e && !strcmp(e, "0") ? 0 : 1;
}();
// polls of one predecessor flag before k_gs_flow gives up (a poll is a
// read from memory and a 128-cycle sleep, about 1 us: some 0.25 s in all,
// far beyond the few tile run-times a predecessor can take); also the bound
// of each grid barrier wait of the persistent FP64 CG.
// SIMPOIS_HIP_GS_FLOW_SPIN=n (read once) sets it, for tests of the error
// path: with n = 0 any wait that is not met at once gives up.
// log2 of the arrival groups of the persistent FP64 CG's grid barriers:
// SIMPOIS_HIP_CG_GROUPS=1, 2, 4, 8, 16 or 32 (read once; default 8, the XCDs)
static const int g_cg_groups = [] {
	const char *e = getenv("SIMPOIS_HIP_CG_GROUPS");
	const int v = e ? atoi(e) : 8;
	for (int lg = 0; lg <= 5; lg++)
		if (v == 1 << lg) return lg;
	return 3;
}();
// polls of the persistent FP64 CG's start gate before a work-group gives
// the launch up (default 2^12, about 4 ms: the work-groups of a launch that
// can all be resident start within microseconds of each other);
// SIMPOIS_HIP_CG_GATE_SPIN=n (read once) sets it, for tests of the fallback
// (with 0 a launch goes ahead only if its last work-group counts in before
// any other polls)
static const int g_cg_gate_spin = [] {
	const char *e = getenv("SIMPOIS_HIP_CG_GATE_SPIN");
	if (!e || !*e) return 1 << 12;
	const long v = strtol(e, nullptr, 10);
	return v < 0 ? 0 : v > (1L << 30) ? 1 << 30 : (int)v;
}();
// k_cg_persist launches in this process that ran the iterations / gave
// the launch up at their start gate (sp_cg_persist_counts)
static std::atomic<long> g_cg_persisted{0}, g_cg_gave_up{0};
static const int GS_FLOW_SPIN = [] {
	const char *e = getenv("SIMPOIS_HIP_GS_FLOW_SPIN");
	if (!e || !*e) return 1 << 18;
	const long v = strtol(e, nullptr, 10);
	return v < 0 ? 0 : v > (1L << 30) ? 1 << 30 : (int)v;
}();

// SIMPOIS_HIP_TAIL_LDS=0 keeps every multigrid tail in global memory
static const int g_tail_lds_off = [] {
	const char *e = getenv("SIMPOIS_HIP_TAIL_LDS");
	return e && !strcmp(e, "0") ? 1 : 0;
}();

// tile shape of the fused levels 0..MG_FUSE_MAX (index into MG_FINE_* at
// levels 0 and 1, into MG_GEO_* from level 2); -1: the default rule.
// SIMPOIS_HIP_MG_GEO="g0,g1,g2,..." sets them level by level.
static int g_mg_geo[MG_FUSE_MAX + 1] = {-1, -1, -1, -1, -1, -1, -1, -1};
static const int g_mg_geo_env = [] {
	const char *e = getenv("SIMPOIS_HIP_MG_GEO");
	for (int l = 0; e && *e && l <= MG_FUSE_MAX; l++) {
		char *end;
		const long v = strtol(e, &end, 10);
		if (end == e) break;
		if (v >= 0 && v < (l < 2 ? MG_NFINE : MG_NGEO)) g_mg_geo[l] = (int)v;
		e = *end == ',' ? end + 1 : end;
	}
	return 0;
}();

// smoothing of the per-level kernels and the tail, sweeps at the coarsest
// level, coarse-correction factor; SIMPOIS_HIP_MG_TAIL="nu,sweeps,wc" sets them
static int g_mg_nu = MG_NU_TAIL, g_mg_csw = 4;
static float g_mg_wc = 1.8f;
static const int g_mg_tail_env = [] {
	const char *e = getenv("SIMPOIS_HIP_MG_TAIL");
	int nu, csw;
	float wc;
	if (e && sscanf(e, "%d,%d,%f", &nu, &csw, &wc) == 3 && nu >= 1 && nu <= 16
			&& csw >= 1 && csw <= 256 && wc > 0.0f && wc < 4.0f) {
		g_mg_nu = nu;
		g_mg_csw = csw;
		g_mg_wc = wc;
	}
	return 0;
}();

// default tile shapes.  The fused kernels move their region through HBM
// once per launch whatever the number of sweeps, so more smoothing there
// costs compute only and saves V-cycles: level 0 smooths nu = 5 times
// (108x44 tiles), level 1 nu = 4 (112x48), the coarse levels nu = 6, on
// 104x40 tiles with 1024 threads while the level (all planes of the batch)
// has 2^19 cells or more, below that on 40x8 tiles (many short tiles: those
// levels are bound by a tile's latency, not by halo bytes).  Measured on
// 4096x4096x3 to 5e-6: 8 CG iterations, 10.2 ms, against 11 iterations and
// 11.6 ms with nu = 2, 2, 4.
static int mg_geo_of(int l, double cells)
{
	if (g_mg_geo[l] >= 0) return g_mg_geo[l];
	if (l < 2) return l == 0 ? 4 : 3;
	return cells >= (double)(1 << 19) ? 4 : 5;
}

static int ensure_device()
{
	static std::mutex mu;
	std::lock_guard<std::mutex> lk(mu);
	if (g_dev_checked)
		return g_dev_ok ? SP_OK : set_err(SP_ERR_NODEV, "no usable gfx950 device");
	g_dev_checked = 1;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
		return set_err(SP_ERR_NODEV, "no HIP device visible");
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return set_err(SP_ERR_NODEV, "hipGetDevice failed");
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
		return set_err(SP_ERR_NODEV, "hipGetDeviceProperties failed");
	if (!strstr(prop.gcnArchName, "gfx950"))
		return set_err(SP_ERR_NODEV, "device %s is not gfx950", prop.gcnArchName);
	int lds = 0;
	if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)
			== hipSuccess && lds > 0)
		g_max_lds = std::min(lds, 65536);
	if (!g_gsc_off)
		g_gsc_lds = spk::gs_cells_enable(GSC_LDS_MAX) == 0 ? GSC_LDS_MAX : g_max_lds;
	(void)hipGetLastError();
	g_dev_ok = 1;
	return SP_OK;
}

// ------------------------------------------------------------ profiler
//
// When enabled, each phase records an event pair around its launches on the
// solve's stream, with the launch count and the algorithmic bytes of the
// phase (definitions next to each use).  sp_profile_read sums them.

struct ProfRec { std::string name; hipEvent_t a, b; long launches; double bytes; };
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

struct ProfScope {
	bool on;
	ProfRec r;
	hipStream_t s;
	ProfScope(const char *name, hipStream_t st) : on(false), s(st)
	{
		std::lock_guard<std::mutex> lk(g_prof_mu);
		if (!g_prof_on) return;
		on = hipEventCreate(&r.a) == hipSuccess && hipEventCreate(&r.b) == hipSuccess;
		if (!on) return;
		r.name = name; r.launches = 0; r.bytes = 0;
		(void)hipEventRecord(r.a, s);
	}
	void add(long launches, double bytes) { r.launches += launches; r.bytes += bytes; }
	~ProfScope()
	{
		if (!on) return;
		(void)hipEventRecord(r.b, s);
		std::lock_guard<std::mutex> lk(g_prof_mu);
		g_prof.push_back(r);
	}
};

// per-launch timer used by the CG engine in profiling runs
static std::vector<ProfRec> g_open;
static void prof_fn(const char *name, int begin, double bytes, hipStream_t s)
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	if (!g_prof_on) return;
	if (begin) {
		ProfRec r;
		r.name = name; r.launches = 1; r.bytes = 0;
		if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess)
			return;
		(void)hipEventRecord(r.a, s);
		g_open.push_back(r);
		return;
	}
	for (size_t k = g_open.size(); k-- > 0;)
		if (g_open[k].name == name) {
			ProfRec r = g_open[k];
			g_open.erase(g_open.begin() + k);
			(void)hipEventRecord(r.b, s);
			r.bytes = bytes;
			g_prof.push_back(r);
			return;
		}
}

static bool prof_enabled()
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	return g_prof_on;
}

extern "C" void sp_profile(int on)
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	for (auto &r : g_prof) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
	g_prof.clear();
	g_prof_on = on != 0;
}

extern "C" int sp_profile_read(const char *name, double *ms, long *launches,
		double *bytes)
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	double t = 0, by = 0;
	long n = 0;
	for (auto &r : g_prof) {
		if (r.name != name) continue;
		HIPCHK(hipEventSynchronize(r.b));
		float e = 0;
		HIPCHK(hipEventElapsedTime(&e, r.a, r.b));
		t += e; n += r.launches; by += r.bytes;
	}
	if (ms) *ms = t;
	if (launches) *launches = n;
	if (bytes) *bytes = by;
	return SP_OK;
}

// ------------------------------------------------ device memory account
//
// Every device allocation of the library goes through dev_alloc / dev_free,
// which keep the bytes held (cached plans included) and refuse one that
// would take them over the cap of sp_set_alloc_cap (0: no cap) with
// SP_ERR_NOMEM, as a failed hipMalloc returns.

static std::atomic<size_t> g_live_bytes{0};
static std::atomic<size_t> g_alloc_cap{0};

static int dev_alloc(void **p, size_t bytes)
{
	*p = nullptr;
	const size_t cap = g_alloc_cap.load(), live = g_live_bytes.load();
	if (cap && live + bytes > cap)
		return set_err(SP_ERR_NOMEM, "device allocation of %zu bytes refused: the "
				"library holds %zu of its %zu-byte cap", bytes, live, cap);
	HIPCHK(hipMalloc(p, bytes));
	g_live_bytes += bytes;
	return SP_OK;
}

static void dev_free(void *p, size_t bytes)
{
	if (!p) return;
	(void)hipFree(p);
	g_live_bytes -= bytes;
}

extern "C" size_t sp_set_alloc_cap(size_t bytes)
{
	return g_alloc_cap.exchange(bytes);
}

extern "C" size_t sp_device_bytes(void) { return g_live_bytes.load(); }

// --------------------------------------------------------------- plan

struct PlaneIn {            // one plane of one image (device pointers)
	int w, h;
	const float *g, *f;
	float *out;
};

struct LevelPlan {
	std::vector<int> members;        // plane ids present at this level
	std::vector<SpPlane> sp;         // their descriptors at this level
	std::vector<int> nsweep, cgit;   // after the PONLIT rule
	SpPlane *d_sp = nullptr;
	SpPlane *d_parent = nullptr;     // finer-level descriptor of each member
	SpPlane *d_coarse = nullptr;     // coarser-level descriptor (g = null: none)
	BlockWork *d_bw = nullptr;
	int nbw = 0;
	int *d_first = nullptr;          // first work block of each member, + end
	int *d_bcnt = nullptr;           // holes per work block
	GsInst *d_insts = nullptr;
	std::vector<GsInst> insts;       // host copy of d_insts
	std::vector<int> tiled;          // members swept in tiles (too big for LDS)
	int gs_var = 0;                  // tile shape variant (sp_gs.h)
	bool gs_tuned = false;
	std::vector<int> wf_off;
	GsTile *d_tiles = nullptr;
	long tile_pass_px = 0;           // for the profiler
	int *d_small = nullptr;
	int nsmall = 0, small_lds = 0, small_threads = 64;
	int cells_slots = 0;             // slots per thread of k_gs_cells (0: k_gs_small)
	// dependence-ordered tiles (k_gs_flow): member tile grids, one flag per
	// grid cell, resident work-groups, flag value of the last launch
	GsFlowInst *d_flow = nullptr;
	int *d_flags = nullptr;
	int flow_blocks = 0, epoch = 0;
	bool any_cg = false;
	// FP32 CG of this level, built on first use and kept with the plan
	struct {
		bool built = false;
		Cg32Batch B{};
		std::vector<MgLevelWork> mgw;
		std::vector<Cg32Scal> sc0;      // initial scalars, uploaded per solve
		std::vector<int> ms;            // members in the batch
		long maxit_host = 0;
		int trace_len = 0, chunk = 1;
		int last_chunks = 1;            // chunks the last solve ran
		hipGraphExec_t exec = nullptr;
	} cg;
};

struct CgTrace { int n = 0; std::vector<double> rr; };

struct Plan {
	sp_params P{};
	int R = 1;
	float teff = 0;
	std::vector<PlaneIn> planes;
	std::vector<int> depth;          // pyramid levels of each plane
	std::vector<LevelPlan> lv;
	std::vector<std::pair<void *, size_t>> allocs;
	size_t bytes = 0;                // device bytes held by allocs
	int *d_cnt = nullptr;
	float **d_outs = nullptr;
	// FP64 CG scratch, sized for the largest plane
	float *cg_q = nullptr;
	double *cg_x = nullptr, *cg_r = nullptr, *cg_p = nullptr,
		*cg_part = nullptr, *cg_rr = nullptr, *cg_rrp = nullptr;
	unsigned *cg_idx = nullptr;         // pixel of each unknown of the plane solved
	int *cg_bcnt = nullptr;             // unknowns per setup block, prefix sums
	float *cg_pf = nullptr;             // direction rounded to float (full grid)
	int cg_rr_cap = 0;
	int cg_cap[2] = {-1, -1};           // resident k_cg_persist work-groups (-1: not asked)
	int *cg_done = nullptr, *h_done = nullptr;
	unsigned *cg_bar = nullptr;         // k_cg_persist's barrier words
	spk::CgDev *cg_dev = nullptr;       // descriptor the FP64 CG kernels read
	// FP64 CG iterations captured in chunks of 16 << i (i = 0..2), one set
	// per operator and grid (the work blocks of a level's unknowns), shared
	// by the levels and planes of that grid (they read *cg_dev); built on
	// first use
	struct Cg64Graphs { int bilap, grid; hipGraphExec_t exec[3]; };
	std::vector<Cg64Graphs> cg64;
	// pinned staging of small per-solve copies (descriptors in, results out):
	// copies to and from pageable memory each cost a host round trip
	void *h_stage = nullptr;
	size_t h_stage_cap = 0;
	// FP32 batched CG scratch: r p q (float) and x (double) per pixel
	float *cg32_vec = nullptr;
	double *cg32_x = nullptr;
	unsigned char *cg32_code = nullptr;
	size_t cg32_cap = 0;
	std::vector<std::vector<CgTrace>> traces;   // [plane][level]
	std::vector<int> cg_iters_finest;
	std::vector<int> status_finest;     // SP_PLANE_* of each plane's finest level
	std::vector<double> rr_finest;      // last r.r of that CG (-1: none)
	// FP64 residual of the output, per plane (sp_solve_batch_device_ex)
	double *res_part = nullptr, *res_out = nullptr;
	hipStream_t cap_stream = nullptr;

	// claim counter of k_gs_flow and its error flag (read after a solve)
	unsigned *gs_claim = nullptr;
	int *gs_err = nullptr, *h_gs_err = nullptr;
	bool gs_flow_used = false;

	int transpose = 0;               // the CG applies the transposed biharmonic
	                                 // operator (the adjoint's solve)
	std::vector<char> key;           // identity of the call this plan serves
	// guidance from a source image: plambda's Laplacian of each plane's
	// caller-side f (the image) into plan memory, which the solve uses as f
	std::vector<std::pair<const float *, float *>> guide;
	std::vector<std::pair<int, int>> guide_wh;

	~Plan()
	{
		for (auto &L : lv)
			if (L.cg.exec) (void)hipGraphExecDestroy(L.cg.exec);
		for (auto &g : cg64)
			for (auto &e : g.exec)
				if (e) (void)hipGraphExecDestroy(e);
		for (auto &a : allocs) dev_free(a.first, a.second);
		if (h_done) (void)hipHostFree(h_done);
		if (h_gs_err) (void)hipHostFree(h_gs_err);
		if (h_stage) (void)hipHostFree(h_stage);
		if (cap_stream) (void)hipStreamDestroy(cap_stream);
	}
	// pinned staging area of at least n bytes (no copy may be in flight)
	int stage(size_t n)
	{
		if (n <= h_stage_cap) return SP_OK;
		if (h_stage) (void)hipHostFree(h_stage);
		h_stage = nullptr;
		h_stage_cap = 0;
		if (hipHostMalloc(&h_stage, n) != hipSuccess)
			return set_err(SP_ERR_NOMEM, "pinned host allocation of %zu bytes failed", n);
		h_stage_cap = n;
		return SP_OK;
	}
	template <typename T> int alloc(T **p, size_t n)
	{
		void *q = nullptr;
		const size_t by = std::max<size_t>(n, 1) * sizeof(T);
		TRY(dev_alloc(&q, by));
		allocs.push_back({q, by});
		bytes += by;
		*p = (T *)q;
		return SP_OK;
	}
	template <typename T> int upload(T **p, const std::vector<T> &v)
	{
		TRY(alloc(p, v.size()));
		if (!v.empty())
			HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
		return SP_OK;
	}
};

// per-plane arrays start on 256-byte boundaries (paired loads need 8)
static size_t rup64(size_t n) { return (n + 63) & ~(size_t)63; }

static int op_radius(float tstep, float *teff)
{
	if (!(tstep < 0)) { *teff = tstep; return 1; }
	if (tstep < -1000) {
		float t = tstep;
		t += 1000;
		t *= -1;
		*teff = t;
		return 3;
	}
	*teff = tstep;
	return 2;
}

static constexpr int EW_SPAN = 256 * 4;   // pixels per elementwise work block

// Tiles of the level's tiled members for shape variant `var`, merged by
// wavefront index (tiles of one wavefront are independent across members)
static int gs_tiles(const Plan &pl, const LevelPlan &L, int var,
		std::vector<GsTile> &tiles, std::vector<int> &wf_off, long *pass_px)
{
	std::vector<std::vector<GsTile>> wf;
	long px = 0;
	for (int m : L.tiled) {
		const int w = L.sp[m].w, h = L.sp[m].h, T = L.nsweep[m];
		std::vector<GsTile> t;
		std::vector<int> o;
		const int bt = build_tiles_var(pl.R, var, w, h, T, 1, t, o);
		if (!bt) return set_err(SP_ERR_ARG, "no tile shape %d for radius %d", var, pl.R);
		if (wf.size() + 1 < o.size()) wf.resize(o.size() - 1);
		for (size_t z = 0; z + 1 < o.size(); z++)
			for (int e = o[z]; e < o[z + 1]; e++) {
				GsTile tl = t[e];
				tl.inst = m;
				wf[z].push_back(tl);
			}
		px += (long)((T + bt - 1) / bt) * w * h;
	}
	tiles.clear();
	wf_off.assign(1, 0);
	for (auto &v : wf) {
		tiles.insert(tiles.end(), v.begin(), v.end());
		wf_off.push_back((int)tiles.size());
	}
	if (pass_px) *pass_px = px;
	return SP_OK;
}

// the dependence-ordered launch's data for the tiles of variant var: each
// member's tile grid and one flag per grid cell, 0 for a tile (below every
// epoch), INT_MAX where the grid has no tile (never waited for); *d_flags
// stays null when the grids exceed the flag index range
static int gs_flow_build(Plan &pl, const LevelPlan &L, int var,
		const std::vector<GsTile> &tiles, GsFlowInst **d_fi, int **d_flags)
{
	*d_flags = nullptr;
	std::vector<GsFlowInst> fi(L.sp.size(), GsFlowInst{0, 0, 0, 0});
	long total = 0;
	for (int m : L.tiled) {
		GsFlowInst &f = fi[m];
		tile_grid_var(pl.R, var, L.sp[m].w, L.sp[m].h, L.nsweep[m], f.nI, f.nJ, f.nT);
		f.base = (int)std::min<long>(total, 0x7fffffffL);
		total += (long)f.nI * f.nJ * f.nT;
	}
	if (total > 0x7fffffffL) return SP_OK;
	std::vector<int> flags((size_t)total, 0x7fffffff);
	for (const GsTile &t : tiles) {
		const GsFlowInst &f = fi[t.inst];
		flags[(size_t)f.base + ((size_t)t.tt * f.nJ + t.tj) * f.nI + t.ti] = 0;
	}
	TRY(pl.upload(d_fi, fi));
	TRY(pl.upload(d_flags, flags));
	if (!pl.gs_claim) {
		TRY(pl.alloc(&pl.gs_claim, 1));
		TRY(pl.alloc(&pl.gs_err, 1));
		HIPCHK(hipMemset(pl.gs_err, 0, sizeof(int)));
		if (hipHostMalloc(&pl.h_gs_err, sizeof(int)) != hipSuccess)
			return set_err(SP_ERR_NOMEM, "pinned host allocation of 4 bytes failed");
	}
	return SP_OK;
}

static int gs_build(Plan &pl, LevelPlan &L, int var)
{
	std::vector<GsTile> tiles;
	TRY(gs_tiles(pl, L, var, tiles, L.wf_off, &L.tile_pass_px));
	TRY(pl.upload(&L.d_tiles, tiles));
	L.gs_var = var;
	L.d_flags = nullptr;
	if (!g_gs_flow || tiles.empty()) return SP_OK;
	TRY(gs_flow_build(pl, L, var, tiles, &L.d_flow, &L.d_flags));
	L.flow_blocks = spk::gs_flow_blocks(pl.R, var);
	L.epoch = 0;
	return SP_OK;
}

// ---------------------------------------------------------- autotuner
//
// With sp_params.autotune set, the first sweep of a level whose members run
// in tiles times every compiled tile shape of its radius (sp_gs.h) on a
// scratch copy of the level's current data, keeps the fastest, and caches
// the choice for the process under (radius, sweeps, member sizes).  Every
// shape runs the same lexicographic schedule, so the choice never changes
// the result, only the time.

struct TuneRec {
	std::vector<long> key;
	int var;
	std::vector<float> ms;   // per variant
};
static std::mutex g_tune_mu;
static std::vector<TuneRec> g_tune;

static std::vector<long> gs_key(const Plan &pl, const LevelPlan &L)
{
	std::vector<long> k = {pl.R};
	for (int m : L.tiled) {
		k.push_back(L.sp[m].w);
		k.push_back(L.sp[m].h);
		k.push_back(L.nsweep[m]);
	}
	return k;
}

static int gs_launch_all(const Plan &pl, int var, const GsTile *d_tiles,
		const std::vector<int> &wf_off, const GsInst *d_insts, hipStream_t s)
{
	for (size_t k = 0; k + 1 < wf_off.size(); k++)
		spk::launch_gs_tiles(pl.R, var, d_tiles + wf_off[k], wf_off[k + 1] - wf_off[k],
				d_insts, s);
	HIPCHK(hipGetLastError());
	return SP_OK;
}

static int gs_tune(Plan &pl, LevelPlan &L, hipStream_t s)
{
	L.gs_tuned = true;
	const std::vector<long> key = gs_key(pl, L);
	{
		std::lock_guard<std::mutex> lk(g_tune_mu);
		for (const TuneRec &r : g_tune)
			if (r.key == key) {
				if (r.var != L.gs_var) TRY(gs_build(pl, L, r.var));
				return SP_OK;
			}
	}
	// scratch copy of the tiled members' current values
	size_t need = 0;
	for (int m : L.tiled) need += rup64((size_t)L.sp[m].w * L.sp[m].h);
	float *scratch = nullptr;
	TRY(pl.alloc(&scratch, need));
	std::vector<GsInst> gi = L.insts;
	size_t off = 0;
	for (int m : L.tiled) {
		gi[m].u = scratch + off;
		off += rup64((size_t)L.sp[m].w * L.sp[m].h);
	}
	GsInst *d_gi = nullptr;
	TRY(pl.upload(&d_gi, gi));
	hipEvent_t a, b;
	HIPCHK(hipEventCreate(&a));
	HIPCHK(hipEventCreate(&b));
	const int nv = sp_shape_count(pl.R);
	std::vector<float> ms(nv, 0.0f);
	int best = 0;
	int rc = SP_OK;
	for (int v = 0; v < nv && rc == SP_OK; v++) {
		std::vector<GsTile> tiles;
		std::vector<int> wo;
		rc = gs_tiles(pl, L, v, tiles, wo, nullptr);
		GsTile *d_t = nullptr;
		if (rc == SP_OK) rc = pl.upload(&d_t, tiles);
		// time the launch form the solve uses
		GsFlowInst *d_fi = nullptr;
		int *d_fl = nullptr, nb = 0;
		if (rc == SP_OK && g_gs_flow && !tiles.empty()) {
			rc = gs_flow_build(pl, L, v, tiles, &d_fi, &d_fl);
			nb = std::min((int)tiles.size(), spk::gs_flow_blocks(pl.R, v));
		}
		float t = 1e30f;
		for (int rep = 0; rep < 2 && rc == SP_OK; rep++) {   // the first warms up
			for (int m : L.tiled)
				if (hipMemcpyAsync(gi[m].u, L.sp[m].u,
						(size_t)L.sp[m].w * L.sp[m].h * 4,
						hipMemcpyDeviceToDevice, s) != hipSuccess)
					rc = set_err(SP_ERR_HIP, "autotune copy failed");
			if (rc != SP_OK) break;
			(void)hipEventRecord(a, s);
			if (d_fl) {
				spk::launch_gs_flow(pl.R, v, d_t, (int)tiles.size(), nb, d_gi, d_fi,
						pl.gs_claim, d_fl, rep + 1, GS_FLOW_SPIN, pl.gs_err, s);
				if (hipGetLastError() != hipSuccess)
					rc = set_err(SP_ERR_HIP, "autotune launch failed");
			} else {
				rc = gs_launch_all(pl, v, d_t, wo, d_gi, s);
			}
			(void)hipEventRecord(b, s);
			if (hipEventSynchronize(b) != hipSuccess)
				rc = set_err(SP_ERR_HIP, "autotune run failed");
			float e = 0;
			(void)hipEventElapsedTime(&e, a, b);
			if (rep) t = e;
		}
		ms[v] = t;
		if (t < ms[best]) best = v;
	}
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	TRY(rc);
	if (best != L.gs_var) TRY(gs_build(pl, L, best));
	std::lock_guard<std::mutex> lk(g_tune_mu);
	g_tune.push_back({key, best, ms});
	return SP_OK;
}

extern "C" void sp_cg_persist_counts(long *ran, long *gave_up)
{
	if (ran) *ran = g_cg_persisted.load();
	if (gave_up) *gave_up = g_cg_gave_up.load();
}

extern "C" int sp_autotune_report(char *buf, int cap)
{
	std::string out;
	std::lock_guard<std::mutex> lk(g_tune_mu);
	for (const TuneRec &r : g_tune) {
		const int R = (int)r.key[0];
		char line[256];
		long px = 0;
		for (size_t z = 1; z + 2 < r.key.size(); z += 3) px += r.key[z] * r.key[z + 1];
		snprintf(line, sizeof line, "gs R=%d T=%ld planes=%d px=%ld choice=", R,
				r.key.size() > 3 ? r.key[3] : 0L, (int)(r.key.size() - 1) / 3, px);
		out += line;
		for (int v = 0; v < (int)r.ms.size(); v++) {
			sp_with_shape(R, v, [&](auto sh) {
				using S = decltype(sh);
				snprintf(line, sizeof line, "%s%s%dx%dx%d:%.3f", v ? "," : "",
						v == r.var ? "*" : "", S::BI, S::BJ, S::BT, r.ms[v]);
			});
			out += line;
		}
		out += "\n";
	}
	if (buf && cap > 0) {
		const size_t n = std::min(out.size(), (size_t)cap - 1);
		memcpy(buf, out.data(), n);
		buf[n] = 0;
	}
	return (int)out.size() + 1;
}

static int plan_create(Plan &pl, const std::vector<PlaneIn> &planes,
		const sp_params &P, bool has_f)
{
	pl.P = P;
	pl.planes = planes;
	if (P.guidance) {
		if (P.guidance != 1 && P.guidance != 2)
			return set_err(SP_ERR_ARG, "guidance must be 0, 1 (plambda x,l) or 2 (x,lr)");
		for (PlaneIn &q : pl.planes) {
			if (!q.f) continue;
			float *d = nullptr;
			TRY(pl.alloc(&d, (size_t)q.w * q.h));
			pl.guide.push_back({q.f, d});
			pl.guide_wh.push_back({q.w, q.h});
			q.f = d;
		}
	}
	pl.R = op_radius(P.tstep, &pl.teff);
	const int np = (int)planes.size();
	// pyramid depth of each plane, simpois.c:253-256
	std::vector<std::vector<std::pair<int, int>>> shapes(np);
	int nlev = 0;
	for (int p = 0; p < np; p++) {
		int cw = pl.planes[p].w, ch = pl.planes[p].h, sc = P.scale;
		for (;;) {
			shapes[p].push_back({cw, ch});
			if (!(sc > 1 && (cw > 1 || ch > 1))) break;
			cw = (int)ceil(cw / 2.0);
			ch = (int)ceil(ch / 2.0);
			sc--;
		}
		pl.depth.push_back((int)shapes[p].size());
		nlev = std::max(nlev, (int)shapes[p].size());
	}
	pl.lv.resize(nlev);
	size_t ncnt = 0;
	for (int k = 0; k < nlev; k++)
		for (int p = 0; p < np; p++)
			if (pl.depth[p] > k) { pl.lv[k].members.push_back(p); ncnt++; }
	TRY(pl.alloc(&pl.d_cnt, ncnt));
	HIPCHK(hipHostMalloc((void **)&pl.h_done, 8 * sizeof(int)));
	pl.traces.assign(np, std::vector<CgTrace>(nlev));
	pl.cg_iters_finest.assign(np, 0);
	pl.status_finest.assign(np, SP_PLANE_SWEPT);
	pl.rr_finest.assign(np, -1.0);
	{
		std::vector<float *> outs(np);
		for (int p = 0; p < np; p++) outs[p] = pl.planes[p].out;
		TRY(pl.upload(&pl.d_outs, outs));
	}
	size_t cnt_off = 0;
	for (int k = 0; k < nlev; k++) {
		LevelPlan &L = pl.lv[k];
		const int nm = (int)L.members.size();
		size_t pool_px = 0;
		for (int p : L.members)
			pool_px += (size_t)shapes[p][k].first * shapes[p][k].second;
		float *pool = nullptr;
		const int arrays = k > 0 ? (has_f ? 4 : 3) : 1;   // g u c [f] | c
		TRY(pl.alloc(&pool, pool_px * arrays));
		size_t off = 0;
		L.sp.resize(nm);
		for (int m = 0; m < nm; m++) {
			const int p = L.members[m];
			const int w = shapes[p][k].first, h = shapes[p][k].second;
			const size_t n = (size_t)w * h;
			SpPlane &S = L.sp[m];
			S.w = w; S.h = h;
			S.cnt = pl.d_cnt + cnt_off + m;
			if (k == 0) {
				S.g = const_cast<float *>(pl.planes[p].g);
				S.f = const_cast<float *>(pl.planes[p].f);
				S.u = pl.planes[p].out;
				S.c = pool + off; off += n;
			} else {
				S.g = pool + off; off += n;
				S.u = pool + off; off += n;
				S.c = pool + off; off += n;
				S.f = nullptr;
				if (has_f) { S.f = pool + off; off += n; }
			}
			// only the Gauss-Seidel sweeps read the coefficient plane
			if (P.niter <= 0) S.c = nullptr;
		}
		cnt_off += nm;
		TRY(pl.upload(&L.d_sp, L.sp));
		// work blocks, contiguous per member
		std::vector<BlockWork> bw;
		std::vector<int> first;
		for (int m = 0; m < nm; m++) {
			first.push_back((int)bw.size());
			const size_t n = (size_t)L.sp[m].w * L.sp[m].h;
			for (size_t s = 0; s < n; s += EW_SPAN) bw.push_back({m, (int)s});
		}
		first.push_back((int)bw.size());
		L.nbw = (int)bw.size();
		TRY(pl.upload(&L.d_bw, bw));
		TRY(pl.upload(&L.d_first, first));
		TRY(pl.alloc(&L.d_bcnt, bw.size()));
		// PONLIT restricts the work to the level of that width (:272-273)
		L.nsweep.assign(nm, std::max(P.niter, 0));
		L.cgit.assign(nm, P.cgit);
		for (int m = 0; m < nm; m++)
			if (P.ponlit != 0 && P.ponlit != (double)L.sp[m].w)
				L.nsweep[m] = L.cgit[m] = 0;
		for (int m = 0; m < nm; m++) L.any_cg |= L.cgit[m] != 0;
		std::vector<GsInst> gi(nm);
		for (int m = 0; m < nm; m++)
			gi[m] = {L.sp[m].u, L.sp[m].c, L.sp[m].w, L.sp[m].h, L.nsweep[m],
				pl.teff, nullptr};
		TRY(pl.upload(&L.d_insts, gi));
		L.insts = gi;
		// Gauss-Seidel: members that fit in LDS run whole; others in tiles.
		// For R = 1 the whole-level kernel deals the unknown pixels to the
		// threads (k_gs_cells), for levels of up to GSC_MAXCELLS pixels
		std::vector<int> small;
		const bool cells = pl.R == 1 && g_gsc_lds > 0;
		if (cells) {
			int th = 64, maxs = 1, lds = 0;
			for (int m = 0; m < nm; m++) {
				const int w = L.sp[m].w, h = L.sp[m].h, T = L.nsweep[m];
				if (T <= 0) continue;
				if ((long)w * h <= GSC_MAXCELLS && gsc_lds_bytes(w, h) <= g_gsc_lds
						&& small_nsteps(1, w, h, T) < 0x7fffffffL) {
					small.push_back(m);
					th = std::max(th, gsc_threads(w, h));
					lds = std::max<int>(lds, (int)gsc_lds_bytes(w, h));
					continue;
				}
				L.tiled.push_back(m);
			}
			for (int m : small) maxs = std::max(maxs, gsc_slots(L.sp[m].w, L.sp[m].h, th));
			L.small_threads = th;
			L.small_lds = lds;
			L.cells_slots = maxs;
		}
		for (int m = 0; m < nm && !cells; m++) {
			const int w = L.sp[m].w, h = L.sp[m].h, T = L.nsweep[m];
			if (T <= 0) continue;
			// whole level in one work-group: it fits in LDS, every thread
			// holds at most SMALL_MAXSLOT slots, and the step count fits int
			const long slots = (long)h * ((w - 1) / (pl.R * pl.R + 1) + 1);
			if ((long)8 * small_cells(pl.R, w, h) <= g_max_lds
					&& slots <= 1024L * SMALL_MAXSLOT
					&& small_nsteps(pl.R, w, h, T) < 0x7fffffffL) {
				small.push_back(m);
				const long th = (slots + 63) / 64 * 64;
				L.small_threads = (int)std::max<long>(L.small_threads,
						std::min<long>(1024, th));
				L.small_lds = std::max(L.small_lds, spk::small_lds_bytes(pl.R, w, h));
				continue;
			}
			L.tiled.push_back(m);
		}
		L.nsmall = (int)small.size();
		TRY(pl.upload(&L.d_small, small));
		TRY(gs_build(pl, L, gs_default_var(pl.R)));
	}
	// parent (finer) and coarse descriptors aligned with each level's members
	for (int k = 0; k < nlev; k++) {
		LevelPlan &L = pl.lv[k];
		std::vector<SpPlane> par, co;
		for (size_t m = 0; m < L.members.size(); m++) {
			const int p = L.members[m];
			if (k > 0) {
				const LevelPlan &F = pl.lv[k - 1];
				const size_t fm = std::find(F.members.begin(), F.members.end(), p)
					- F.members.begin();
				par.push_back(F.sp[fm]);
			}
			SpPlane none{};
			if (k + 1 < nlev && pl.depth[p] > k + 1) {
				const LevelPlan &C = pl.lv[k + 1];
				const size_t cm = std::find(C.members.begin(), C.members.end(), p)
					- C.members.begin();
				co.push_back(C.sp[cm]);
			} else {
				co.push_back(none);
			}
		}
		if (k > 0) TRY(pl.upload(&L.d_parent, par));
		TRY(pl.upload(&L.d_coarse, co));
	}
	return SP_OK;
}

static int run_gs(Plan &pl, LevelPlan &L, hipStream_t s)
{
	if (pl.P.autotune && !L.gs_tuned && L.wf_off.size() > 1) TRY(gs_tune(pl, L, s));
	if (L.nsmall && L.cells_slots)
		spk::launch_gs_cells(L.cells_slots, L.d_insts, L.d_small, L.nsmall, L.small_lds,
				L.small_threads, s);
	else if (L.nsmall)
		spk::launch_gs_small(pl.R, L.d_insts, L.d_small, L.nsmall, L.small_lds,
				L.small_threads, s);
	if (L.wf_off.size() > 1) {
		// algorithmic bytes: per pass of BT sweeps, read u and the
		// coefficient plane and write u once (12 B per pixel)
		ProfScope pr(pl.R == 1 ? "gs_tile" : "gs_tile_r23", s);
		if (L.d_flags && L.epoch >= 0x7ffffffe)
			L.d_flags = nullptr;   // flags exhausted: one launch per wavefront
		if (L.d_flags) {
			const int nt = L.wf_off.back(), nb = std::min(nt, L.flow_blocks);
			spk::launch_gs_flow(pl.R, L.gs_var, L.d_tiles, nt, nb, L.d_insts, L.d_flow,
					pl.gs_claim, L.d_flags, ++L.epoch, GS_FLOW_SPIN, pl.gs_err, s);
			HIPCHK(hipGetLastError());
			pl.gs_flow_used = true;
			pr.add(1, 12.0 * L.tile_pass_px);
		} else {
			TRY(gs_launch_all(pl, L.gs_var, L.d_tiles, L.wf_off, L.d_insts, s));
			pr.add((long)L.wf_off.size() - 1, 12.0 * L.tile_pass_px);
		}
	}
	HIPCHK(hipGetLastError());
	return SP_OK;
}

// ---------------------------------------------------------- FP64 CG

static int cg64_ensure(Plan &pl, int maxit)
{
	if (!pl.cg_x) {
		size_t n0 = 0;
		for (auto &p : pl.planes) n0 = std::max(n0, (size_t)p.w * p.h);
		TRY(pl.alloc(&pl.cg_x, n0));
		TRY(pl.alloc(&pl.cg_r, n0));
		TRY(pl.alloc(&pl.cg_p, n0));
		TRY(pl.alloc(&pl.cg_q, n0));
		int nb0 = 0;
		for (auto &p : pl.planes) nb0 = std::max(nb0, spk::cg_setup_blocks(p.w, p.h));
		TRY(pl.alloc(&pl.cg_idx, n0));
		TRY(pl.alloc(&pl.cg_bcnt, (size_t)nb0 + 1));
		TRY(pl.alloc(&pl.cg_pf, n0));
		TRY(pl.alloc(&pl.cg_part, std::max<size_t>(2 * 1024, nb0)));
		TRY(pl.alloc(&pl.cg_done, 8));
		TRY(pl.alloc(&pl.cg_bar, 64 * 36));
		TRY(pl.alloc(&pl.cg_rrp, 2));
		TRY(pl.alloc(&pl.cg_dev, 1));
	}
	if (maxit + 2 > pl.cg_rr_cap) {
		const int cap = std::max(maxit + 2, 1024);
		TRY(pl.alloc(&pl.cg_rr, cap));
		pl.cg_rr_cap = cap;
	}
	return SP_OK;
}

// The FP64 CG's captured chunks: 16, 32 and 64 iterations on grid work
// blocks (a level's own count: a grid sized for the largest plane left
// most blocks of the smaller counts returning at once, three launches per
// iteration)
static int cg64_graphs(Plan &pl, int bilap, int grid, hipGraphExec_t **out)
{
	for (auto &g : pl.cg64)
		if (g.bilap == bilap && g.grid == grid) { *out = g.exec; return SP_OK; }
	if (!pl.cap_stream)
		HIPCHK(hipStreamCreateWithFlags(&pl.cap_stream, hipStreamNonBlocking));
	Plan::Cg64Graphs G{bilap, grid, {}};
	for (int i = 0; i < 3; i++) {
		hipGraph_t graph;
		HIPCHK(hipStreamBeginCapture(pl.cap_stream, hipStreamCaptureModeThreadLocal));
		for (int z = 0; z < (16 << i); z++)
			spk::launch_cg_iter(pl.cg_dev, bilap, grid, z & 1, pl.cap_stream);
		HIPCHK(hipStreamEndCapture(pl.cap_stream, &graph));
		const hipError_t ie = hipGraphInstantiate(&G.exec[i], graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		if (ie != hipSuccess) {
			for (auto &e : G.exec)
				if (e) (void)hipGraphExecDestroy(e);
			HIPCHK(ie);
		}
	}
	pl.cg64.push_back(G);
	*out = pl.cg64.back().exec;
	return SP_OK;
}

// masked CG on one plane of one level in FP64 (cleant_cgpois.c:197-264)
static int run_cg64(Plan &pl, int k, int m, int cgit, int bilap, float eps,
		hipStream_t s)
{
	LevelPlan &L = pl.lv[k];
	const SpPlane &S = L.sp[m];
	int nmask = 0;
	HIPCHK(hipMemcpyAsync(&nmask, S.cnt, sizeof(int), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	const int plane = L.members[m];
	if (k == 0) { pl.status_finest[plane] = SP_PLANE_CONVERGED; pl.rr_finest[plane] = 0; }
	if (nmask == 0) return SP_OK;   // :203-207, the level output rule covers it
	const float e = pl.P.cg_eps >= 0 ? (float)pl.P.cg_eps : eps;
	const float fm = cgit == -1 ? (float)nmask : (float)cgit;
	const int maxit = (int)fm;
	if (k == 0) pl.status_finest[plane] = SP_PLANE_MAXIT;
	if (maxit <= 0) return SP_OK;   // x = x0: u is unchanged
	// iteration kernels: 4 unknowns per thread and pass, at most 1024 blocks
	const int span = spk::cg_iter_span();
	const int nblk = std::min(1024, std::max(1, (nmask + span - 1) / span));
	TRY(cg64_ensure(pl, maxit));
	spk::CgDev d;
	d.g = S.g; d.f = S.f; d.u = S.u;
	d.x = pl.cg_x; d.r = pl.cg_r; d.p = pl.cg_p; d.q = pl.cg_q;
	d.idx = pl.cg_idx; d.bcnt = pl.cg_bcnt; d.pf = pl.cg_pf;
	d.part = pl.cg_part; d.rr = pl.cg_rr; d.rrp = pl.cg_rrp; d.done = pl.cg_done;
	d.bar = pl.cg_bar;
	d.w = S.w; d.h = S.h; d.bilap = bilap; d.nblk = nblk;
	d.nb = spk::cg_setup_blocks(S.w, S.h);
	d.nc = nmask;
	d.stop = (double)e;
	// the iterations in one persistent launch when the unknowns take one
	// pass of its work-groups and those can all be resident
	if (pl.cg_cap[bilap] < 0) pl.cg_cap[bilap] = spk::cg_persist_capacity(bilap);
	const bool persist = g_cg_persist && (long)nmask <= (long)nblk * span
		&& nblk <= pl.cg_cap[bilap];
	hipGraphExec_t *exec = nullptr;
	if (!persist) TRY(cg64_graphs(pl, bilap, nblk, &exec));
	// the descriptor into device memory, in stream order (no copy in
	// flight uses the staging area: the hole count above was waited for)
	TRY(pl.stage(sizeof d));
	memcpy(pl.h_stage, &d, sizeof d);
	HIPCHK(hipMemcpyAsync(pl.cg_dev, pl.h_stage, sizeof d, hipMemcpyHostToDevice, s));
	ProfScope pr("cg64", s);
	spk::launch_cg_setup(pl.cg_dev, d.nb, s);
	int it = 0, ci = 0, ran = maxit;
	bool persisted = false;
	if (persist) {
		spk::launch_cg_persist(pl.cg_dev, bilap, nblk, maxit, GS_FLOW_SPIN, g_cg_gate_spin,
				g_cg_groups, s);
		HIPCHK(hipGetLastError());
		HIPCHK(hipMemcpyAsync(pl.h_done, pl.cg_done, 8 * sizeof(int),
				hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (pl.h_done[7]) {
			// its work-groups were not all resident (other work held CUs):
			// nothing was done, the launches below instead
			g_cg_gave_up++;
			TRY(cg64_graphs(pl, bilap, nblk, &exec));
		} else {
			if (pl.h_done[5])
				return set_err(SP_ERR_HIP, "FP64 CG grid barrier wait timed out");
			it = maxit;
			persisted = true;
			g_cg_persisted++;
			if (pl.h_done[0]) ran = pl.h_done[1];
		}
	}
	// otherwise chunks of 16, 32, then 64 iterations from the captured
	// graphs while a whole chunk fits under maxit, the rest launched one
	// iteration at a time; the stop flag is read after each chunk
	// (iterations after the stop return at once: 64 keeps those few against
	// one host round trip per chunk)
	while (it < maxit) {
		while (ci > 0 && (16 << ci) > maxit - it) ci--;
		int c;
		if ((16 << ci) <= maxit - it) {
			c = 16 << ci;
			HIPCHK(hipGraphLaunch(exec[ci], s));
			if (ci < 2) ci++;
		} else {
			c = maxit - it;
			for (int z = 0; z < c; z++)
				spk::launch_cg_iter(pl.cg_dev, bilap, nblk, (it + z) & 1, s);
		}
		it += c;
		HIPCHK(hipGetLastError());
		HIPCHK(hipMemcpyAsync(pl.h_done, pl.cg_done, 2 * sizeof(int),
				hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (pl.h_done[0]) { ran = pl.h_done[1]; break; }
	}
	spk::launch_cg_finish(pl.cg_dev, nblk, s);
	HIPCHK(hipGetLastError());
	// FP64 vectors: 88 B per unknown and iteration (twice the FP32 vectors)
	pr.add(persisted ? 1 : 3L * it, 88.0 * ran * nmask);
	CgTrace &tr = pl.traces[L.members[m]][k];
	tr.n = nmask;
	tr.rr.resize(ran);
	HIPCHK(hipMemcpyAsync(tr.rr.data(), pl.cg_rr + 1, ran * sizeof(double),
			hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	if (k == 0) {
		pl.cg_iters_finest[plane] = ran;
		const double rr = ran > 0 ? tr.rr[ran - 1] : 0.0;
		pl.rr_finest[plane] = rr;
		// the stop test of minicg.c:47 on the last r.r
		pl.status_finest[plane] = !std::isfinite(rr) ? SP_PLANE_NONFINITE
			: sqrt(rr) < (double)e ? SP_PLANE_CONVERGED : SP_PLANE_MAXIT;
	}
	return SP_OK;
}

// ---------------------------------------------------------- FP32 CG

// Multigrid hierarchy of the preconditioner for the planes of one CG batch:
// aggregation by 2x2 down to a smallest side of 8 cells (at most 16 levels);
// coarse coefficients are computed on the device by mg_build.
static int mg_plan(Plan &pl, std::vector<Cg32Plane> &cp, int bilap,
		std::vector<MgLevelWork> &mgw, Cg32Batch &B)
{
	if (bilap)
		return set_err(SP_ERR_ARG, "multigrid preconditioning covers the "
				"Poisson system only");
	const int np = (int)cp.size();
	std::vector<std::vector<std::pair<int, int>>> shp(np);
	int nlev = 0;
	for (int z = 0; z < np; z++) {
		int w = cp[z].w, h = cp[z].h;
		shp[z].push_back({w, h});
		// level 0 is never the coarsest (its fused kernels restrict)
		while (shp[z].size() < 2 || (std::min(w, h) > 8 && shp[z].size() < (size_t)MG_MAXLEV)) {
			w = (w + 1) / 2;
			h = (h + 1) / 2;
			shp[z].push_back({w, h});
		}
		nlev = std::max(nlev, (int)shp[z].size());
	}
	// storage: level 0 uses the CG vectors; levels > 0 own D wE wS z b;
	// levels 0..MG_FUSE_MAX also zd
	size_t cells = 0;
	for (int z = 0; z < np; z++)
		for (size_t l = 0; l < shp[z].size(); l++) {
			const size_t n = rup64((size_t)shp[z][l].first * shp[z][l].second);
			cells += (l > 0 ? 5 : 0) * n + (l <= (size_t)MG_FUSE_MAX ? n : 0);
		}
	float *pool = nullptr;
	TRY(pl.alloc(&pool, std::max<size_t>(cells, 1)));
	// packed coefficients of the levels that may run fused
	size_t gcells = 0;
	for (int z = 0; z < np; z++)
		for (size_t l = 1; l < shp[z].size() && l <= (size_t)MG_FUSE_MAX; l++)
			gcells += (l > (size_t)MG_G32_MAX ? 2 : 1) *
				rup64((size_t)shp[z][l].first * shp[z][l].second);
	unsigned *gcodes = nullptr;
	TRY(pl.alloc(&gcodes, std::max<size_t>(gcells, 1)));
	size_t goff = 0;
	std::vector<int4> tiles0;
	std::vector<std::vector<MgLevel>> lv(nlev, std::vector<MgLevel>(np));
	size_t off = 0;
	for (int z = 0; z < np; z++)
		for (size_t l = 0; l < shp[z].size(); l++) {
			MgLevel &L = lv[l][z];
			L.w = shp[z][l].first;
			L.h = shp[z][l].second;
			const size_t n = rup64((size_t)L.w * L.h);
			if (l <= (size_t)MG_FUSE_MAX) { L.zd = pool + off; off += n; }
			if (l == 0) {
				L.g = cp[z].g;
				L.z = cp[z].z;
				L.b = cp[z].r;
				L.code = cp[z].code;
				cp[z].t0 = (int)tiles0.size();
				const int g0 = mg_geo_of(0, 0.0);
				for (int y0 = 0; y0 < L.h; y0 += MG_FINE_TY[g0])
					for (int x0 = 0; x0 < L.w; x0 += MG_FINE_TX[g0])
						tiles0.push_back(make_int4(z, x0, y0, 0));
				cp[z].nt = (int)tiles0.size() - cp[z].t0;
				continue;
			}
			if (l <= (size_t)MG_FUSE_MAX) {   // 256-byte aligned: n is a multiple of 64
				L.gcode = gcodes + goff;
				goff += (l > (size_t)MG_G32_MAX ? 2 : 1) * n;
			}
			L.D = pool + off; off += n;
			L.wE = pool + off; off += n;
			L.wS = pool + off; off += n;
			L.z = pool + off; off += n;
			L.b = pool + off; off += n;
		}
	// first level of each plane handled by the single-work-group tail
	// (planes whose tail levels fit in LDS run the LDS version of it)
	std::vector<int4> rng(np);
	int tail_lds = 0;
	for (int z = 0; z < np; z++) {
		const int d = (int)shp[z].size();
		int t = d - 1;
		for (int l = 1; l < d; l++)
			if ((long)shp[z][l].first * shp[z][l].second <= MG_TAIL_CELLS) { t = l; break; }
		t = std::max(1, std::min(t, d - 1));
		long cells = 0, framed = 0;
		for (int l = t; l < d; l++) {
			cells += (long)shp[z][l].first * shp[z][l].second;
			framed += (long)(shp[z][l].first + 2) * (shp[z][l].second + 2);
		}
		const long bytes = 20 * framed;   // z b D wE wS with a zero frame
		const int in_lds = bytes <= MG_TAIL_LDS_MAX && !g_tail_lds_off;
		if (in_lds) tail_lds = std::max<int>(tail_lds, (int)bytes);
		rng[z] = make_int4(t, d, in_lds, (int)cells);
	}
	if (tail_lds > 65536 && spk::mg_tail_lds_enable(tail_lds)) {
		// the device refused the LDS size: every plane takes the global tail
		for (auto &r : rng) r.z = 0;
		tail_lds = 0;
	}
	mgw.assign(nlev, MgLevelWork{});
	for (int l = 0; l < nlev; l++) {
		std::vector<BlockWork> mid, cc, bc;
		std::vector<int4> tl;
		double tcells = 0;
		// the level's tile shape from the cells of its fused planes
		for (int z = 0; z < np; z++)
			if ((int)shp[z].size() > l && l >= 1 && l < rng[z].x && l <= MG_FUSE_MAX)
				tcells += (double)lv[l][z].w * lv[l][z].h;
		const int geo = l <= MG_FUSE_MAX ? mg_geo_of(l, tcells) : 0;
		for (int z = 0; z < np; z++) {
			const int d = (int)shp[z].size();
			if (d <= l) continue;
			const MgLevel &Lz = lv[l][z];
			const size_t n = (size_t)Lz.w * Lz.h;
			const size_t nc = d > l + 1 ? (size_t)lv[l + 1][z].w * lv[l + 1][z].h : 0;
			for (size_t st = 0; st < nc; st += EW_SPAN) bc.push_back({z, (int)st});
			if (l < 1 || l >= rng[z].x) continue;   // level 0 or in the tail
			if (l <= MG_FUSE_MAX) {   // fused tiles
				const int tx = l == 1 ? MG_FINE_TX[geo] : MG_GEO_TX[geo];
				const int ty = l == 1 ? MG_FINE_TY[geo] : MG_GEO_TY[geo];
				for (int y0 = 0; y0 < Lz.h; y0 += ty)
					for (int x0 = 0; x0 < Lz.w; x0 += tx)
						tl.push_back(make_int4(z, x0, y0, 0));
			} else {                  // per-level kernels
				for (size_t st = 0; st < n; st += EW_SPAN) mid.push_back({z, (int)st});
				for (size_t st = 0; st < nc; st += EW_SPAN) cc.push_back({z, (int)st});
			}
		}
		MgLevelWork &W = mgw[l];
		TRY(pl.upload(&W.lv, lv[l]));
		if (l + 1 < nlev) TRY(pl.upload(&W.coarse, lv[l + 1]));
		TRY(pl.upload(&W.mid, mid));
		TRY(pl.upload(&W.cells_c, cc));
		TRY(pl.upload(&W.build_c, bc));
		TRY(pl.upload(&W.tiles, tl));
		W.ntiles = (int)tl.size();
		W.geo = geo;
		W.cells = tcells;
		W.nmid = (int)mid.size();
		W.nlast = 0;
		W.ncells_c = (int)cc.size();
		W.nbuild_c = (int)bc.size();
	}
	std::vector<MgLevel> all((size_t)np * MG_MAXLEV);
	for (int z = 0; z < np; z++)
		for (size_t l = 0; l < shp[z].size(); l++) all[(size_t)z * MG_MAXLEV + l] = lv[l][z];
	TRY(pl.upload(&B.tail_lv, all));
	TRY(pl.upload(&B.tail_rng, rng));
	B.tail_lds = tail_lds;
	B.tail_global = 0;
	for (const int4 &r : rng) B.tail_global |= !r.z;
	TRY(pl.upload(&B.tiles0, tiles0));
	B.ntiles0 = (int)tiles0.size();
	B.geo0 = mg_geo_of(0, 0.0);
	TRY(pl.alloc(&B.tpart, tiles0.size()));
	B.mg = mgw.data();
	B.nlev = nlev;
	B.nu = g_mg_nu;
	B.coarse_sweeps = g_mg_csw;
	B.wc = g_mg_wc;
	return SP_OK;
}

// batched FP32 CG over the members of level k with a non-zero -c: device
// descriptors, multigrid hierarchy and the captured iteration graph are
// built once per plan
static int cg32_build(Plan &pl, int k, int bilap, float eps)
{
	LevelPlan &L = pl.lv[k];
	auto &C = L.cg;
	C.ms.clear();
	for (size_t m = 0; m < L.members.size(); m++)
		if (L.cgit[m] != 0) C.ms.push_back((int)m);
	if (C.ms.empty()) { C.built = true; return SP_OK; }
	size_t need = 0;
	for (int m : C.ms) need += rup64((size_t)L.sp[m].w * L.sp[m].h);
	if (need > pl.cg32_cap) {
		TRY(pl.alloc(&pl.cg32_vec, need * 5));
		TRY(pl.alloc(&pl.cg32_x, need));
		TRY(pl.alloc(&pl.cg32_code, need));
		pl.cg32_cap = need;
	}
	std::vector<Cg32Plane> cp;
	std::vector<BlockWork> bw;
	size_t off = 0, offx = 0;
	C.maxit_host = 0;
	C.trace_len = 0;
	C.sc0.clear();
	for (size_t z = 0; z < C.ms.size(); z++) {
		const int m = C.ms[z];
		const SpPlane &S = L.sp[m];
		const size_t n = (size_t)S.w * S.h, na = rup64(n);
		Cg32Plane c{};
		c.g = S.g; c.f = S.f; c.u = S.u; c.cnt = S.cnt;
		c.code = pl.cg32_code + offx;
		c.x = pl.cg32_x + offx; offx += na;
		c.r = pl.cg32_vec + off; off += na;
		c.p = pl.cg32_vec + off; off += na;
		c.q = pl.cg32_vec + off; off += na;
		c.p2 = pl.cg32_vec + off; off += na;
		c.z = c.r;
		if (pl.P.cg_mg) { c.z = pl.cg32_vec + off; off += na; }
		c.w = S.w; c.h = S.h; c.bilap = bilap && pl.transpose ? 2 : bilap;
		// the coarsest level starts from 0 at the unknowns (k_init) unless
		// sweeps ran before the CG
		c.u0 = k == (int)pl.lv.size() - 1 && L.nsweep[m] <= 0;
		c.blk0 = (int)bw.size();
		for (size_t st = 0; st < n; st += EW_SPAN) bw.push_back({(int)z, (int)st});
		c.nblk = (int)bw.size() - c.blk0;
		cp.push_back(c);
		Cg32Scal q{};
		const int cg = L.cgit[m];
		const double rel = k > 0 && pl.P.cg_rel_coarse > 0 ? pl.P.cg_rel_coarse
			: pl.P.cg_rel;
		q.rel = rel > 0;
		q.eps = q.rel ? rel : (pl.P.cg_eps >= 0 ? (double)(float)pl.P.cg_eps
				: (double)eps);
		q.maxit_from_count = cg == -1;
		q.maxit = cg == -1 ? 0 : std::max(cg, 0);
		const int cap = cg == -1 ? (int)std::min<size_t>(n, 1 << 20) : std::max(cg, 0);
		q.trace0 = C.trace_len;
		q.trace_cap = pl.P.trace ? cap : 0;
		C.trace_len += q.trace_cap;
		C.maxit_host = std::max<long>(C.maxit_host, cg == -1 ? (long)n : cg);
		C.sc0.push_back(q);
	}
	Cg32Batch &B = C.B;
	B = Cg32Batch{};
	if (pl.P.cg_mg) TRY(mg_plan(pl, cp, bilap, C.mgw, B));
	TRY(pl.upload(&B.planes, cp));
	TRY(pl.upload(&B.scal, C.sc0));
	TRY(pl.upload(&B.bw, bw));
	TRY(pl.alloc(&B.part, 2 * bw.size()));
	TRY(pl.alloc(&B.nactive, 1));
	if (C.trace_len) TRY(pl.alloc(&B.trace, C.trace_len));
	B.nplanes = (int)C.ms.size();
	B.nblk = (int)bw.size();
	// iterations in chunks, captured once as a graph
	C.chunk = pl.P.cg_mg ? std::max(1, std::min(pl.P.cg_chunk, 4))
		: std::max(1, pl.P.cg_chunk);
	C.chunk += C.chunk & 1;   // even: the direction buffers swap every iteration
	B.bilap = bilap;
	if (!pl.cap_stream)
		HIPCHK(hipStreamCreateWithFlags(&pl.cap_stream, hipStreamNonBlocking));
	hipGraph_t graph;
	HIPCHK(hipStreamBeginCapture(pl.cap_stream, hipStreamCaptureModeThreadLocal));
	for (int z = 0; z < C.chunk; z++) spk::cg32_iter(B, pl.cap_stream, z & 1);
	HIPCHK(hipStreamEndCapture(pl.cap_stream, &graph));
	const hipError_t ie = hipGraphInstantiate(&C.exec, graph, nullptr, nullptr, 0);
	(void)hipGraphDestroy(graph);
	HIPCHK(ie);
	C.built = true;
	return SP_OK;
}

static int run_cg32(Plan &pl, int k, int bilap, float eps, hipStream_t s)
{
	LevelPlan &L = pl.lv[k];
	auto &C = L.cg;
	if (!C.built) TRY(cg32_build(pl, k, bilap, eps));
	if (C.ms.empty()) return SP_OK;
	Cg32Batch &B = C.B;
	B.mg = C.mgw.data();
	// staging: per-plane scalars, the rr trace, the level's hole counts
	const size_t nsc = C.sc0.size() * sizeof(Cg32Scal);
	const size_t ntr = C.trace_len * sizeof(double);
	const size_t ncn = L.members.size() * sizeof(int);
	// no copy of the staging area is in flight here: every run ends with a
	// stream synchronization after its last copy
	TRY(pl.stage(nsc + ntr + ncn));
	char *stg = (char *)pl.h_stage;
	memcpy(stg, C.sc0.data(), nsc);
	HIPCHK(hipMemcpyAsync(B.scal, stg, nsc, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemsetAsync(B.nactive, 0, sizeof(int), s));
	ProfScope pr("cg32", s);
	// profiling runs launch the iterations directly, with a timer around
	// the timed kernels; other runs replay the captured graph
	const bool prof = prof_enabled();
	double px0 = 0;
	for (int m : C.ms) px0 += (double)L.sp[m].w * L.sp[m].h;
	B.px0 = px0;
	B.prof = prof ? prof_fn : nullptr;
	spk::cg32_setup(B, s);
	long launched = 0;
	// profiling runs go one iteration at a time, so that no launch after
	// convergence (an empty one) enters the per-kernel averages
	// Other runs launch at first as many chunks as the last solve of this
	// plan needed, back to back, before the first look at the stop flag
	// (repeated solves of one shape converge alike), then one at a time.
	int ahead = prof ? 1 : std::max(1, C.last_chunks);
	long chunks = 0;
	for (long done = 0; done < C.maxit_host;) {
		for (int a = 0; a < ahead && done < C.maxit_host; a++) {
			const int step = prof ? 1 : C.chunk;
			if (prof) spk::cg32_iter(B, s, (int)(done & 1));
			else HIPCHK(hipGraphLaunch(C.exec, s));
			done += step;
			launched += step;
			chunks++;
		}
		ahead = 1;
		HIPCHK(hipMemcpyAsync(pl.h_done, B.nactive, sizeof(int), hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (pl.h_done[0] == 0) break;
	}
	if (!prof) C.last_chunks = (int)chunks;
	B.prof = nullptr;
	spk::cg32_finish(B, s);
	HIPCHK(hipGetLastError());
	// per-plane results: iterations and trace (the members' hole counters
	// are consecutive: one copy)
	HIPCHK(hipMemcpyAsync(stg, B.scal, nsc, hipMemcpyDeviceToHost, s));
	if (ntr) HIPCHK(hipMemcpyAsync(stg + nsc, B.trace, ntr, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(stg + nsc + ntr, L.sp[0].cnt, ncn, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	std::vector<Cg32Scal> sc(C.sc0.size());
	memcpy(sc.data(), stg, nsc);
	std::vector<double> trace(C.trace_len);
	if (ntr) memcpy(trace.data(), stg + nsc, ntr);
	std::vector<int> cnts(C.ms.size());
	for (size_t z = 0; z < C.ms.size(); z++)
		memcpy(&cnts[z], stg + nsc + ntr + C.ms[z] * sizeof(int), sizeof(int));
	double unk = 0;
	for (size_t z = 0; z < C.ms.size(); z++) {
		const int p = L.members[C.ms[z]];
		CgTrace &tr = pl.traces[p][k];
		tr.n = cnts[z];
		const int it = sc[z].iters;
		tr.rr.assign(trace.begin() + sc[z].trace0,
				trace.begin() + sc[z].trace0 + std::min(it, sc[z].trace_cap));
		if (k == 0) {
			pl.cg_iters_finest[p] = it;
			pl.status_finest[p] = sc[z].status;
			pl.rr_finest[p] = sc[z].rr;
		}
		unk += (double)it * cnts[z];
	}
	// algorithmic bytes per unknown and iteration: spmv reads p, writes q
	// (8); update reads r q, writes r (12); direction reads z p x (double),
	// writes p x (28): 48 B.  The multigrid V-cycle is not counted.
	pr.add((pl.P.cg_mg ? 0 : 5) * launched, 48.0 * unk);
	return SP_OK;
}

// -------------------------------------------------------------- solve

// CG histories of the last solve in the process, [plane][level]
static std::mutex g_trace_mu;
static std::vector<std::vector<CgTrace>> g_last_traces;

extern "C" int sp_last_cg_trace(int plane, int level, int *n, double *rr, int cap)
{
	std::lock_guard<std::mutex> lk(g_trace_mu);
	if (plane < 0 || plane >= (int)g_last_traces.size() || level < 0 ||
			level >= (int)g_last_traces[plane].size())
		return -1;
	const CgTrace &t = g_last_traces[plane][level];
	if (n) *n = t.n;
	const int k = (int)t.rr.size();
	if (rr) for (int i = 0; i < std::min(k, cap); i++) rr[i] = t.rr[i];
	return k;
}

static int plan_execute(Plan &pl, hipStream_t s, sp_stats *st)
{
	const int nlev = (int)pl.lv.size();
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	HIPCHK(hipEventRecord(e0, s));
	for (auto &v : pl.traces) for (auto &t : v) { t.n = 0; t.rr.clear(); }
	// data term from the source image (plambda "x,l" / "x,lr", nearest rule)
	for (size_t k = 0; k < pl.guide.size(); k++)
		spk::launch_plambda_lap(pl.guide[k].second, pl.guide[k].first,
				pl.guide_wh[k].first, pl.guide_wh[k].second, 1, pl.P.guidance - 1, 1, s);
	// pyramid of data, simpois.c:260-262
	{
		ProfScope pr("pyramid", s);
		LevelPlan &L0 = pl.lv[0];
		spk::launch_prepare(L0.d_sp, L0.d_bw, L0.nbw, L0.d_bcnt, s);
		spk::launch_count(L0.d_sp, (int)L0.members.size(), L0.d_first, L0.d_bcnt, s);
		for (int k = 1; k < nlev; k++) {
			LevelPlan &L = pl.lv[k];
			spk::launch_restrict(L.d_parent, L.d_sp, L.d_bw, L.nbw, pl.P.pmsfac,
					L.d_bcnt, s);
			spk::launch_count(L.d_sp, (int)L.members.size(), L.d_first, L.d_bcnt, s);
		}
		HIPCHK(hipGetLastError());
	}
	// coarse to fine
	const bool pois = pl.P.tstep > 0;
	for (int k = nlev - 1; k >= 0; k--) {
		LevelPlan &L = pl.lv[k];
		spk::launch_init(L.d_sp, L.d_coarse, L.d_bw, L.nbw, s);
		TRY(run_gs(pl, L, s));
		if (L.any_cg) {   // simpois.c:278-284; cgit == 0 is a no-op there
			const float eps = pois ? 1e-6f : 1e-9f;
			if (pl.P.cg_fp32) {
				TRY(run_cg32(pl, k, pois ? 0 : 1, eps, s));
			} else {
				for (size_t m = 0; m < L.members.size(); m++)
					if (L.cgit[m])
						TRY(run_cg64(pl, k, (int)m, L.cgit[m], pois ? 0 : 1, eps, s));
			}
		}
	}
	spk::launch_finalize(pl.lv[0].d_sp, pl.d_outs, pl.lv[0].d_bw, pl.lv[0].nbw, s);
	HIPCHK(hipGetLastError());
	if (pl.gs_flow_used)
		HIPCHK(hipMemcpyAsync(pl.h_gs_err, pl.gs_err, sizeof(int), hipMemcpyDeviceToHost, s));
	HIPCHK(hipEventRecord(e1, s));
	HIPCHK(hipEventSynchronize(e1));
	if (pl.gs_flow_used && *pl.h_gs_err) {
		*pl.h_gs_err = 0;
		(void)hipMemset(pl.gs_err, 0, sizeof(int));
		return set_err(SP_ERR_HIP, "Gauss-Seidel tile dependence wait timed out");
	}
	float ms = 0;
	(void)hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (st) {
		st->levels = nlev;
		int mx = 0;
		for (int v : pl.cg_iters_finest) mx = std::max(mx, v);
		st->cg_iters_finest = mx;
		st->elapsed_ms = ms;
	}
	{
		std::lock_guard<std::mutex> lk(g_trace_mu);
		g_last_traces = pl.traces;
	}
	if (pl.P.trace)   // reference order: plane by plane, coarse to fine
		for (size_t p = 0; p < pl.planes.size(); p++)
			for (int k = pl.depth[p] - 1; k >= 0; k--) {
				const CgTrace &t = pl.traces[p][k];
				for (size_t i = 0; i < t.rr.size(); i++)
					fprintf(stderr, "n=%d iter=%d, rr_new=%g\n", t.n, (int)i, t.rr[i]);
			}
	return SP_OK;
}

static int check_plane_args(const void *out, const void *in, int w, int h, int pd)
{
	if (!out || !in) return set_err(SP_ERR_ARG, "NULL buffer");
	if (w < 1 || h < 1 || pd < 1) return set_err(SP_ERR_ARG, "bad size %dx%dx%d", w, h, pd);
	if ((long)w * h > (1L << 30)) return set_err(SP_ERR_ARG, "image too large");
	if (out == in) return set_err(SP_ERR_ARG, "out must not alias in");
	return SP_OK;
}

// ------------------------------------------------------------ plan cache
//
// Plans are kept by call signature (plane sizes, buffer addresses,
// parameters), most recently used first, within a count and a byte budget;
// repeated solves of the same call skip allocation, tiling and graph
// instantiation.  A call holds the cache's lock while it runs.

static std::mutex g_cache_mu;
static std::list<Plan *> g_cache;
static int g_cache_max = 256;
static size_t g_cache_max_bytes = (size_t)16 << 30;

// drop least recently used plans (never `keep`) until the cache fits
static void cache_trim(const Plan *keep, int max_plans, size_t max_bytes)
{
	size_t bytes = 0;
	for (const Plan *q : g_cache) bytes += q->bytes;
	for (auto it = g_cache.end(); it != g_cache.begin() &&
			((int)g_cache.size() > max_plans || bytes > max_bytes);) {
		--it;
		if (*it == keep) continue;
		bytes -= (*it)->bytes;
		delete *it;
		it = g_cache.erase(it);
	}
}

extern "C" void sp_plan_cache_limit(int max_plans, size_t max_bytes)
{
	std::lock_guard<std::mutex> lk(g_cache_mu);
	g_cache_max = std::max(max_plans, 1);
	g_cache_max_bytes = max_bytes;
	cache_trim(nullptr, g_cache_max, g_cache_max_bytes);
}

extern "C" void sp_plan_cache_clear(void)
{
	std::lock_guard<std::mutex> lk(g_cache_mu);
	cache_trim(nullptr, 0, 0);
}

// the cached plan of a call (created on a miss); the caller holds g_cache_mu
static int cache_get(const std::vector<PlaneIn> &planes, const sp_params &P,
		bool has_f, int transpose, Plan **out)
{
	std::vector<char> key;
	auto put = [&](const void *v, size_t n) {
		key.insert(key.end(), (const char *)v, (const char *)v + n);
	};
	put(&transpose, sizeof transpose);
	for (const PlaneIn &q : planes) {
		put(&q.w, sizeof q.w); put(&q.h, sizeof q.h);
		put(&q.g, sizeof q.g); put(&q.f, sizeof q.f); put(&q.out, sizeof q.out);
	}
	put(&P.tstep, sizeof P.tstep); put(&P.niter, sizeof P.niter);
	put(&P.scale, sizeof P.scale); put(&P.cgit, sizeof P.cgit);
	put(&P.pmsfac, sizeof P.pmsfac); put(&P.ponlit, sizeof P.ponlit);
	put(&P.cg_eps, sizeof P.cg_eps); put(&P.trace, sizeof P.trace);
	put(&P.cg_fp32, sizeof P.cg_fp32); put(&P.cg_rel, sizeof P.cg_rel);
	put(&P.cg_chunk, sizeof P.cg_chunk); put(&P.cg_mg, sizeof P.cg_mg);
	put(&P.autotune, sizeof P.autotune); put(&P.guidance, sizeof P.guidance);
	put(&P.cg_rel_coarse, sizeof P.cg_rel_coarse);
	for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
		if ((*it)->key == key) {
			g_cache.splice(g_cache.begin(), g_cache, it);
			*out = g_cache.front();
			return SP_OK;
		}
	for (int attempt = 0;; attempt++) {
		Plan *pl = new Plan;
		pl->transpose = transpose;
		const int rc = plan_create(*pl, planes, P, has_f);
		if (rc == SP_OK) {
			pl->key = key;
			g_cache.push_front(pl);
			*out = pl;
			return SP_OK;
		}
		delete pl;
		// out of device memory with plans cached: free them, try once more
		if (rc != SP_ERR_NOMEM || attempt || g_cache.empty()) return rc;
		cache_trim(nullptr, 0, 0);
	}
}

// FP64 residual of each plane's output over its holes and the holes
// themselves (finest level; its members are the planes in order)
static int plane_residuals(Plan &pl, hipStream_t s, std::vector<double> &rel,
		std::vector<int> &holes)
{
	LevelPlan &L0 = pl.lv[0];
	const int np = (int)L0.members.size();
	rel.assign(np, -1.0);
	holes.assign(np, 0);
	HIPCHK(hipMemcpyAsync(holes.data(), L0.sp[0].cnt, np * sizeof(int),
			hipMemcpyDeviceToHost, s));
	if (pl.R != 3) {   // (no CG system of the triharmonic operator)
		if (!pl.res_part) {
			TRY(pl.alloc(&pl.res_part, 2 * (size_t)L0.nbw));
			TRY(pl.alloc(&pl.res_out, 3 * (size_t)np));
		}
		spk::launch_residual(L0.d_sp, L0.d_bw, L0.nbw, pl.R == 2, pl.res_part, s);
		spk::launch_residual_reduce(pl.res_part, L0.d_first, np, pl.res_out, s);
		HIPCHK(hipGetLastError());
	}
	std::vector<double> h3(3 * (size_t)np);
	if (pl.R != 3)
		HIPCHK(hipMemcpyAsync(h3.data(), pl.res_out, h3.size() * sizeof(double),
				hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	if (pl.R != 3)
		for (int m = 0; m < np; m++) rel[m] = h3[3 * (size_t)m + 2];
	return SP_OK;
}

// transpose: the CG solves with the transposed biharmonic operator (the
// adjoint of the biharmonic fill; FP32 CG without multigrid or sweeps)
static int solve_batch(int n, const sp_image *imgs, const sp_params *p, void *stream,
		sp_stats *st, sp_plane_stats *ps, int transpose)
{
	if (n < 1 || !imgs || !p) return set_err(SP_ERR_ARG, "empty batch");
	if (transpose && !(p->tstep < 0 && p->cg_fp32 && !p->cg_mg && p->niter == 0))
		return set_err(SP_ERR_ARG, "transposed solves run the biharmonic FP32 CG "
				"without multigrid or sweeps");
	TRY(ensure_device());
	std::vector<PlaneIn> planes;
	bool has_f = false;
	for (int i = 0; i < n; i++) {
		const sp_image &I = imgs[i];
		TRY(check_plane_args(I.out, I.g, I.w, I.h, I.pd));
		const size_t wh = (size_t)I.w * I.h;
		has_f |= I.f != nullptr;
		for (int l = 0; l < I.pd; l++)
			planes.push_back({I.w, I.h, I.g + wh * l, I.f ? I.f + wh * l : nullptr,
				I.out + wh * l});
	}
	hipStream_t s = (hipStream_t)stream;
	std::lock_guard<std::mutex> lk(g_cache_mu);
	Plan *pl = nullptr;
	TRY(cache_get(planes, *p, has_f, transpose, &pl));
	const int rc = plan_execute(*pl, s, st);
	if (rc == SP_OK && ps) {
		std::vector<double> rel;
		std::vector<int> holes;
		TRY(plane_residuals(*pl, s, rel, holes));
		for (size_t q = 0; q < planes.size(); q++) {
			sp_plane_stats &o = ps[q];
			o.status = pl->status_finest[q];
			o.levels = pl->depth[q];
			o.holes = holes[q];
			o.sweeps = pl->lv[0].nsweep[q];
			o.cg_iters = pl->cg_iters_finest[q];
			o.cg_rr = pl->rr_finest[q];
			o.rel_residual = rel[q];
			if (pl->R != 3 && !std::isfinite(rel[q])) o.status = SP_PLANE_NONFINITE;
		}
	}
	cache_trim(pl, g_cache_max, g_cache_max_bytes);
	return rc;
}

extern "C" int sp_solve_batch_device_ex(int n, const sp_image *imgs,
		const sp_params *p, void *stream, sp_stats *st, sp_plane_stats *ps)
{
	return solve_batch(n, imgs, p, stream, st, ps, 0);
}

extern "C" int sp_solve_batch_device(int n, const sp_image *imgs,
		const sp_params *p, void *stream, sp_stats *st)
{
	return sp_solve_batch_device_ex(n, imgs, p, stream, st, nullptr);
}

extern "C" int sp_solve_device(float *d_out, const float *d_in, const float *d_dat,
		int w, int h, int pd, const sp_params *p, void *stream, sp_stats *st)
{
	if (!p) return set_err(SP_ERR_ARG, "NULL params");
	TRY(check_plane_args(d_out, d_in, w, h, pd));
	sp_image I{w, h, pd, d_in, d_dat, d_out};
	return sp_solve_batch_device(1, &I, p, stream, st);
}

extern "C" int sp_solve_host(float *out, const float *in, const float *dat,
		int w, int h, int pd, const sp_params *p, sp_stats *st)
{
	if (!p) return set_err(SP_ERR_ARG, "NULL params");
	TRY(check_plane_args(out, in, w, h, pd));
	TRY(ensure_device());
	const size_t n = (size_t)w * h * pd;
	float *d_in = nullptr, *d_dat = nullptr, *d_out = nullptr;
	int rc = SP_OK;
	auto cleanup = [&]() {
		(void)hipFree(d_in); (void)hipFree(d_dat); (void)hipFree(d_out);
	};
	if (hipMalloc(&d_in, n * 4) != hipSuccess || hipMalloc(&d_out, n * 4) != hipSuccess
			|| (dat && hipMalloc(&d_dat, n * 4) != hipSuccess)) {
		cleanup();
		return set_err(SP_ERR_NOMEM, "device allocation of %zu bytes failed", n * 4);
	}
	if (hipMemcpy(d_in, in, n * 4, hipMemcpyHostToDevice) != hipSuccess
			|| (dat && hipMemcpy(d_dat, dat, n * 4, hipMemcpyHostToDevice) != hipSuccess))
		rc = set_err(SP_ERR_HIP, "upload failed");
	if (!rc) rc = sp_solve_device(d_out, d_in, d_dat, w, h, pd, p, nullptr, st);
	if (!rc && hipMemcpy(out, d_out, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
		rc = set_err(SP_ERR_HIP, "download failed");
	cleanup();
	return rc;
}

// ------------------------------------------------------------- adjoint

// The adjoint system's planes (g2, f2, lambda) live in scratch kept for the
// last batch shape, so that repeated backward passes of one shape solve on
// the same addresses and find their solve plan in the cache.
struct AdjScratch {
	std::vector<int> shape;          // w, h, pd of each image
	Plan mem;                        // owns the scratch and the descriptors
	float *buf = nullptr;
	size_t total = 0;
	AdjPlane *d_ap = nullptr;
	BlockWork *d_bw = nullptr;
	int nbw = 0;
};
static std::mutex g_adj_mu;
static AdjScratch *g_adj = nullptr;

// ------------------------------------------ adjoint of the finite sweeps
//
// With -c 0 the output of simpois (simpois.c:247-285) is n Gauss-Seidel
// sweeps at every pyramid level, an affine map of the known pixels and the
// data term for a fixed mask.  Its gradient, level by level from the finest:
//
// * The level's output is g at its known pixels and u at its holes
//   (cleant_cgpois.c:251-256): dL/dg gets v there, lambda = v at the holes.
// * One sweep, on the holes (known neighbours enter through B g):
//   u <- u + M (b - A u), A = -Laplacian, M = (I/t + L)^-1, b = -f + B g,
//   L / U the couplings before / after a pixel in raster order.  With
//   S = I - M A the adjoint is S^T, and S^T = (I/t + U) S_b (I/t + U)^-1
//   with S_b = I - (I/t + U)^-1 A a backward sweep (A symmetric).  So with
//   nu_n = (I/t + U)^-1 lambda and nu_{j-1} = S_b nu_j:
//     dL/d(initial u) = (I/t + U) nu_0,  dL/db = nu_n + ... + nu_1,
//   dL/df = -dL/db and dL/dg += B^T dL/db at the known pixels.  On the image
//   turned by 180 degrees a backward sweep is a forward one, and
//   (I/t + U)^-1 lambda is one forward sweep from 0 with data -lambda: the
//   library's own Gauss-Seidel kernels run both (boundary 0 at the known
//   pixels, data 0 after the first), one sweep at a time so that each
//   nu_j is added up.
// * The initial value at the holes is the bilinear x2 of the coarser
//   level's output (simpois.c:199-239): its transpose gives the coarser
//   level's v, and the recursion goes on there.
// * The coarser g and f are 2x2 means of the finite samples, f times PMSFAC
//   (simpois.c:172-196, 260-262): their transposes carry the coarser
//   levels' dL/dg and dL/df down to the finest level, coarsest first.
//
// The sweeps are linear with coefficients fixed by the mask, so the
// backward pass reads no forward iterate: nothing is stored during the
// forward solve and nothing is recomputed.  Requires tstep > 0 (the
// symmetric 5-point operator) and finite or NaN data.
struct AsLevel {
	int w = 0, h = 0, nsweep = 0;
	size_t n = 0;               // pixels of one plane at this level
	// pd planes each, plane l at l * n
	unsigned char *hole = nullptr;
	float *v = nullptr, *gg = nullptr, *gf = nullptr, *gi = nullptr;
	float *g_rot = nullptr, *f_rot = nullptr, *u_rot = nullptr, *acc = nullptr;
	Plan *sweep = nullptr;      // one sweep of the turned planes, this size
};

struct AsScratch {
	std::vector<int> key;       // w, h, pd, levels, niter, PONLIT
	Plan mem;                   // owns the level buffers
	std::vector<AsLevel> lv;
	~AsScratch() { for (auto &L : lv) delete L.sweep; }
};
static std::mutex g_as_mu;
static AsScratch *g_as = nullptr;

static int as_scratch(int w, int h, int pd, const sp_params &P, AsScratch **out)
{
	std::vector<std::pair<int, int>> shp;
	for (int cw = w, ch = h, sc = P.scale;;) {
		shp.push_back({cw, ch});
		if (!(sc > 1 && (cw > 1 || ch > 1))) break;
		cw = (int)ceil(cw / 2.0);
		ch = (int)ceil(ch / 2.0);
		sc--;
	}
	std::vector<int> key = {w, h, pd, (int)shp.size(), P.niter, (int)P.ponlit};
	if (g_as && g_as->key == key) { *out = g_as; return SP_OK; }
	delete g_as;
	g_as = nullptr;
	AsScratch *A = new AsScratch;
	A->key = key;
	A->lv.resize(shp.size());
	sp_params P1 = P;
	P1.niter = 1; P1.scale = 1; P1.cgit = 0; P1.ponlit = 0; P1.autotune = 0;
	P1.guidance = 0; P1.trace = 0;
	int rc = SP_OK;
	for (size_t k = 0; k < shp.size() && !rc; k++) {
		AsLevel &L = A->lv[k];
		L.w = shp[k].first;
		L.h = shp[k].second;
		L.n = (size_t)L.w * L.h;
		// PONLIT: only the level of that width sweeps (simpois.c:271-272)
		L.nsweep = P.ponlit && P.ponlit != L.w ? 0 : std::max(P.niter, 0);
		const size_t m = L.n * pd;
		rc = A->mem.alloc(&L.hole, m);
		if (!rc) rc = A->mem.alloc(&L.v, m);
		if (!rc) rc = A->mem.alloc(&L.gg, m);
		if (!rc) rc = A->mem.alloc(&L.gf, m);
		if (!rc) rc = A->mem.alloc(&L.gi, m);
		if (rc || !L.nsweep) continue;
		rc = A->mem.alloc(&L.g_rot, m);
		if (!rc) rc = A->mem.alloc(&L.f_rot, m);
		if (!rc) rc = A->mem.alloc(&L.u_rot, m);
		if (!rc) rc = A->mem.alloc(&L.acc, m);
		if (rc) break;
		std::vector<PlaneIn> planes;
		for (int l = 0; l < pd; l++)
			planes.push_back(PlaneIn{L.w, L.h, L.g_rot + L.n * l, L.f_rot + L.n * l,
				L.u_rot + L.n * l});
		L.sweep = new Plan;
		rc = plan_create(*L.sweep, planes, P1, true);
	}
	if (rc) { delete A; return rc; }
	g_as = A;
	*out = A;
	return SP_OK;
}

// one sweep of the turned planes held by the level's plan: `prepare`
// recomputes the coefficient planes from g_rot / f_rot, `start` sets u = 0
static int as_sweep(AsLevel &L, bool prepare, bool start, hipStream_t s)
{
	Plan &pl = *L.sweep;
	LevelPlan &L0 = pl.lv[0];
	if (prepare) spk::launch_prepare(L0.d_sp, L0.d_bw, L0.nbw, L0.d_bcnt, s);
	if (start) spk::launch_init(L0.d_sp, L0.d_coarse, L0.d_bw, L0.nbw, s);
	return run_gs(pl, L0, s);
}

// the pd planes of one image
static int adjoint_sweeps_image(const sp_adjoint_image &I, const sp_params &P,
		hipStream_t s, int *levels)
{
	AsScratch *A = nullptr;
	TRY(as_scratch(I.w, I.h, I.pd, P, &A));
	const int nl = (int)A->lv.size(), np = I.pd;
	const float t = P.tstep;
	*levels = std::max(*levels, nl);
	// hole masks of every level
	spk::as_mask0(I.g, A->lv[0].n, np, A->lv[0].hole, s);
	for (int k = 1; k < nl; k++)
		spk::as_mask(A->lv[k - 1].hole, A->lv[k - 1].w, A->lv[k - 1].h, A->lv[k].hole,
				A->lv[k].w, A->lv[k].h, np, s);
	HIPCHK(hipMemcpyAsync(A->lv[0].v, I.grad_u, A->lv[0].n * np * sizeof(float),
			hipMemcpyDeviceToDevice, s));
	for (int k = 0; k < nl; k++) {
		AsLevel &L = A->lv[k];
		if (!L.nsweep) {
			spk::as_nosweep(L.hole, L.v, L.gg, L.gi, L.gf, L.w, L.h, np, s);
		} else {
			spk::as_setup(L.hole, L.v, L.gg, L.g_rot, L.f_rot, L.w, L.h, np, s);
			TRY(as_sweep(L, true, true, s));            // nu_n = (I/t + U)^-1 lambda
			spk::as_acc(L.acc, L.u_rot, L.n, np, 1, s);
			spk::as_fill(L.f_rot, L.n, np, 0.0f, s);    // data 0 from here on
			for (int j = 1; j < L.nsweep; j++) {
				TRY(as_sweep(L, j == 1, false, s));
				spk::as_acc(L.acc, L.u_rot, L.n, np, 0, s);
			}
			TRY(as_sweep(L, L.nsweep == 1, false, s)); // nu_0
			spk::as_finish(L.hole, L.u_rot, L.acc, L.w, L.h, np, t, L.gi, L.gf, L.gg, s);
		}
		if (k + 1 < nl)
			spk::as_zoom_in_t(L.gi, L.hole, L.w, L.h, A->lv[k + 1].v, A->lv[k + 1].w,
					A->lv[k + 1].h, np, s);
	}
	// coarser levels' dL/dg and dL/df down to the finest, coarsest first
	for (int k = nl - 2; k >= 0; k--) {
		const AsLevel &C = A->lv[k + 1], &F = A->lv[k];
		spk::as_zoom_out_t(C.gg, C.w, C.h, F.hole, 1, F.w, F.h, np, 1.0f, F.gg, s);
		spk::as_zoom_out_t(C.gf, C.w, C.h, F.hole, 0, F.w, F.h, np, (float)P.pmsfac, F.gf, s);
	}
	HIPCHK(hipGetLastError());
	const size_t bytes = A->lv[0].n * np * sizeof(float);
	HIPCHK(hipMemcpyAsync(I.grad_g, A->lv[0].gg, bytes, hipMemcpyDeviceToDevice, s));
	if (I.grad_f)
		HIPCHK(hipMemcpyAsync(I.grad_f, A->lv[0].gf, bytes, hipMemcpyDeviceToDevice, s));
	return SP_OK;
}

static int adjoint_sweeps(int n, const sp_adjoint_image *imgs, const sp_params &P,
		hipStream_t s, sp_stats *st)
{
	if (P.guidance)
		return set_err(SP_ERR_ARG, "the adjoint takes the data term itself "
				"(guidance = 0); differentiate plambda's Laplacian separately");
	std::lock_guard<std::mutex> lk(g_as_mu);
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	HIPCHK(hipEventRecord(e0, s));
	int levels = 0;
	for (int i = 0; i < n; i++) {
		const int rc = adjoint_sweeps_image(imgs[i], P, s, &levels);
		if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
	}
	HIPCHK(hipEventRecord(e1, s));
	HIPCHK(hipEventSynchronize(e1));
	float ms = 0;
	(void)hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (st) { st->levels = levels; st->cg_iters_finest = 0; st->elapsed_ms = ms; }
	return SP_OK;
}

extern "C" int sp_adjoint_batch_device(int n, const sp_adjoint_image *imgs,
		const sp_params *p, void *stream, sp_stats *st)
{
	if (n < 1 || !imgs || !p) return set_err(SP_ERR_ARG, "empty batch");
	// tstep > 0: the 5-point operator; tstep < 0: the 13-point one, whose
	// clamped taps make it non-symmetric near the border (the adjoint solve
	// then applies A^T)
	const bool bih = p->tstep < 0;
	if (!(p->tstep > 0) && !bih)
		return set_err(SP_ERR_ARG, "tstep must be non-zero");
	if (bih && p->cgit == 0)
		return set_err(SP_ERR_ARG, "the gradient of the finite sweeps covers the "
				"5-point operator (tstep > 0); the biharmonic one needs the "
				"converged fill (cgit != 0)");
	if (p->guidance)
		return set_err(SP_ERR_ARG, "the adjoint takes the data term itself "
				"(guidance = 0); differentiate plambda's Laplacian separately");
	TRY(ensure_device());
	hipStream_t s = (hipStream_t)stream;
	std::vector<int> shape;
	size_t total = 0;
	for (int i = 0; i < n; i++) {
		const sp_adjoint_image &I = imgs[i];
		if (!I.g || !I.grad_u || !I.grad_g || I.w < 1 || I.h < 1 || I.pd < 1)
			return set_err(SP_ERR_ARG, "bad adjoint image %d", i);
		if ((long)I.w * I.h > (1L << 30))
			return set_err(SP_ERR_ARG, "adjoint image %d too large", i);
		total += (size_t)I.w * I.h * I.pd;
		shape.insert(shape.end(), {I.w, I.h, I.pd});
	}
	// -c 0: the gradient of the sweeps' own output
	if (p->cgit == 0) return adjoint_sweeps(n, imgs, *p, s, st);
	std::lock_guard<std::mutex> lk(g_adj_mu);
	if (!g_adj || g_adj->shape != shape) {
		delete g_adj;
		g_adj = new AdjScratch;
		g_adj->shape = shape;
		g_adj->total = total;
		std::vector<BlockWork> bw;
		int id = 0;
		for (int i = 0; i < n; i++)
			for (int l = 0; l < imgs[i].pd; l++, id++)
				for (size_t st2 = 0; st2 < (size_t)imgs[i].w * imgs[i].h; st2 += EW_SPAN)
					bw.push_back({id, (int)st2});
		int rc = g_adj->mem.alloc(&g_adj->buf, 3 * total);
		if (!rc) rc = g_adj->mem.alloc(&g_adj->d_ap, (size_t)id);
		if (!rc) rc = g_adj->mem.upload(&g_adj->d_bw, bw);
		g_adj->nbw = (int)bw.size();
		if (rc) { delete g_adj; g_adj = nullptr; return rc; }
	}
	AdjScratch &A = *g_adj;
	std::vector<AdjPlane> ap;
	std::vector<sp_image> fwd(n);
	size_t off = 0;
	for (int i = 0; i < n; i++) {
		const sp_adjoint_image &I = imgs[i];
		const size_t wh = (size_t)I.w * I.h, m = wh * I.pd;
		float *g2 = A.buf + off, *f2 = A.buf + total + off, *lam = A.buf + 2 * total + off;
		off += m;
		fwd[i] = sp_image{I.w, I.h, I.pd, g2, f2, lam};
		for (int l = 0; l < I.pd; l++)
			ap.push_back({I.g + wh * l, I.grad_u + wh * l, g2 + wh * l, f2 + wh * l,
				lam + wh * l, I.grad_g + wh * l, I.grad_f ? I.grad_f + wh * l : nullptr,
				I.w, I.h});
	}
	// the descriptors are read by kernels of this stream only: copy in
	// stream order (an earlier call's kernels may still read them)
	TRY(A.mem.stage(ap.size() * sizeof(AdjPlane)));
	HIPCHK(hipStreamSynchronize(s));
	memcpy(A.mem.h_stage, ap.data(), ap.size() * sizeof(AdjPlane));
	HIPCHK(hipMemcpyAsync(A.d_ap, A.mem.h_stage, ap.size() * sizeof(AdjPlane),
			hipMemcpyHostToDevice, s));
	spk::launch_adj_prepare(A.d_ap, A.d_bw, A.nbw, s);
	HIPCHK(hipGetLastError());
	if (bih) {
		// A^T lambda = v on the holes: the pyramid of the adjoint data gives
		// each level's CG its start, the FP32 CG (x in FP64) solves with the
		// transposed operator to the caller's tolerance
		sp_params q = *p;
		q.niter = 0;
		q.cg_fp32 = 1;
		q.cg_mg = 0;
		TRY(solve_batch(n, fwd.data(), &q, stream, st, nullptr, 1));
		spk::launch_adj_gather_bilap(A.d_ap, A.d_bw, A.nbw, s);
	} else {
		TRY(sp_solve_batch_device(n, fwd.data(), p, stream, st));
		spk::launch_adj_gather(A.d_ap, A.d_bw, A.nbw, s);
	}
	HIPCHK(hipGetLastError());
	HIPCHK(hipStreamSynchronize(s));
	return SP_OK;
}

// ------------------------------------------------------------------ AMLE
//
// amle (amle.c:203-246) on pd planes: the bounds of each plane, then niter
// raster-order sweeps of tsup (and of tinf when the trace is on) on the
// skewed tiles of sp_amle.h; the output is tsup (amle.c:244).  tinf only
// feeds the diagnostic line (amle.c:226-233), so it is swept only with the
// trace.  The schedules and scratch of the last call are kept for the next
// call of the same size and parameters.

struct AmleSched {
	int T = 0;
	GsTile *d_tiles = nullptr;
	GsInst *d_insts = nullptr;
	std::vector<int> off;
	long px = 0;                      // pixel-sweep updates per run
	// the dependence-ordered launch (k_amle_flow): instance tile grids, one
	// flag per grid cell, resident work-groups, flag value of the last run
	GsFlowInst *d_fi = nullptr;
	int *d_flags = nullptr;
	int nblocks = 0, epoch = 0;
	int ntiles = 0;
};

struct AmlePlan {
	int w = 0, h = 0, pd = 0, nn = 0, niter = 0, trace = 0, R = 1, var = 0;
	Plan mem;                         // owns the device buffers
	float *c = nullptr, *tinf = nullptr, *snap = nullptr;
	float2 *mm = nullptr;
	void *part = nullptr;
	const float *u_for = nullptr;     // output buffer the instances point at
	std::vector<AmleSched> sched;
	// the end of the last call's work on its stream: a call on another
	// stream waits for it before it reuses the scratch
	hipEvent_t done = nullptr;
	hipStream_t last = nullptr;
	bool used = false;
	~AmlePlan() { if (done) (void)hipEventDestroy(done); }
};

static AmleSched *amle_sched(AmlePlan &A, int T)
{
	for (auto &sc : A.sched)
		if (sc.T == T) return &sc;
	return nullptr;
}

static int amle_build(AmlePlan &A, int T)
{
	if (T <= 0 || amle_sched(A, T)) return SP_OK;
	const int ninst = A.trace ? 2 * A.pd : A.pd;
	std::vector<GsTile> tiles;
	AmleSched sc;
	sc.T = T;
	int bt = 0;
	if (!amle_with_shape(A.R, A.var, [&](auto sh) {
			using S = decltype(sh);
			build_tiles<S>(A.w, A.h, T, ninst, tiles, sc.off);
			bt = S::BT;
		}))
		return set_err(SP_ERR_ARG, "no AMLE tile shape %d for radius %d", A.var, A.R);
	sc.px = (long)ninst * ((T + bt - 1) / bt) * A.w * A.h;
	TRY(A.mem.upload(&sc.d_tiles, tiles));
	TRY(A.mem.alloc(&sc.d_insts, ninst));
	sc.ntiles = (int)tiles.size();
	if (g_gs_flow && !tiles.empty()) {
		// every instance has the same tile grid (flags as in gs_flow_build)
		GsFlowInst g{0, 0, 0, 0};
		amle_with_shape(A.R, A.var, [&](auto sh) {
			tile_grid<decltype(sh)>(A.w, A.h, T, g.nI, g.nJ, g.nT);
		});
		const long per = (long)g.nI * g.nJ * g.nT;
		if (per * ninst <= 0x7fffffffL) {
			std::vector<GsFlowInst> fi(ninst, g);
			for (int l = 0; l < ninst; l++) fi[l].base = (int)(per * l);
			std::vector<int> flags((size_t)(per * ninst), 0x7fffffff);
			for (const GsTile &t : tiles)
				flags[(size_t)fi[t.inst].base + ((size_t)t.tt * g.nJ + t.tj) * g.nI + t.ti] = 0;
			TRY(A.mem.upload(&sc.d_fi, fi));
			TRY(A.mem.upload(&sc.d_flags, flags));
			sc.nblocks = std::min(sc.ntiles, spk::amle_flow_blocks(A.R, A.var,
					A.nn == 4 && (A.w > 1 || A.h > 1)));
			if (!A.mem.gs_claim) {
				TRY(A.mem.alloc(&A.mem.gs_claim, 1));
				TRY(A.mem.alloc(&A.mem.gs_err, 1));
				HIPCHK(hipMemset(A.mem.gs_err, 0, sizeof(int)));
				if (hipHostMalloc(&A.mem.h_gs_err, sizeof(int)) != hipSuccess)
					return set_err(SP_ERR_NOMEM, "pinned host allocation of 4 bytes failed");
				*A.mem.h_gs_err = 0;
			}
		}
	}
	A.sched.push_back(sc);
	return SP_OK;
}

// point the instances of every schedule at the current output buffer.  The
// copies go on the call's stream, after the work queued there before (which
// may still read the old instances), from pinned staging that no copy
// still reads (the stream is drained first).
static int amle_bind(AmlePlan &A, float *d_out, hipStream_t s)
{
	if (A.u_for == d_out) return SP_OK;
	const size_t n = (size_t)A.w * A.h;
	std::vector<GsInst> in;
	for (auto &sc : A.sched) {
		for (int l = 0; l < A.pd; l++)
			in.push_back({d_out + n * l, A.c + n * l, A.w, A.h, sc.T, 0.0f, nullptr});
		if (A.trace)
			for (int l = 0; l < A.pd; l++)
				in.push_back({A.tinf + n * l, A.c + n * l, A.w, A.h, sc.T, 0.0f, nullptr});
	}
	TRY(A.mem.stage(in.size() * sizeof(GsInst)));
	HIPCHK(hipStreamSynchronize(s));
	memcpy(A.mem.h_stage, in.data(), in.size() * sizeof(GsInst));
	size_t off = 0;
	for (auto &sc : A.sched) {
		const size_t k = (size_t)(A.trace ? 2 : 1) * A.pd;
		HIPCHK(hipMemcpyAsync(sc.d_insts, (GsInst *)A.mem.h_stage + off,
				k * sizeof(GsInst), hipMemcpyHostToDevice, s));
		off += k;
	}
	A.u_for = d_out;
	return SP_OK;
}

static int amle_sweeps(AmlePlan &A, int T, hipStream_t s)
{
	if (T <= 0) return SP_OK;
	AmleSched *sc = amle_sched(A, T);
	if (!sc) return set_err(SP_ERR_ARG, "AMLE schedule for %d sweeps missing", T);
	ProfScope ps("amle_tile", s);
	const bool fast4 = A.nn == 4 && (A.w > 1 || A.h > 1);
	if (sc->d_flags && sc->epoch >= 0x7ffffffe) sc->d_flags = nullptr;   // flags exhausted
	if (sc->d_flags) {
		spk::launch_amle_flow(A.R, A.var, sc->d_tiles, sc->ntiles, sc->nblocks, sc->d_insts,
				sc->d_fi, A.mem.gs_claim, sc->d_flags, ++sc->epoch, GS_FLOW_SPIN,
				A.mem.gs_err, A.nn, fast4, s);
		A.mem.gs_flow_used = true;
		ps.add(1, 12.0 * (double)sc->px);
	} else {
		for (size_t k = 0; k + 1 < sc->off.size(); k++)
			spk::launch_amle_tiles(A.R, A.var, sc->d_tiles + sc->off[k],
					sc->off[k + 1] - sc->off[k], sc->d_insts, A.nn, fast4, s);
		// algorithmic bytes: u and c read and u written once per pixel per pass
		ps.add((long)sc->off.size() - 1, 12.0 * (double)sc->px);
	}
	HIPCHK(hipGetLastError());
	return SP_OK;
}

static int amle_args(const float *out, const float *in, int w, int h, int pd,
		int nn, int niter)
{
	if (!out || !in || w < 1 || h < 1 || pd < 1)
		return set_err(SP_ERR_ARG, "bad image arguments");
	if ((long)w * h > (1L << 30) || (long)w * h * pd > 0x7fffffffL)
		return set_err(SP_ERR_ARG, "image too large");
	if (nn < 1 || nn > AMLE_NN_MAX)
		return set_err(SP_ERR_ARG, "AMLE neighbourhood size must be 1..32");
	if (niter < 0) return set_err(SP_ERR_ARG, "negative AMLE iteration count");
	return SP_OK;
}

static std::mutex g_amle_mu;
static AmlePlan *g_amle = nullptr;
static int amle_run(AmlePlan &A, float *d_out, const float *d_in, const float *d_init,
		hipStream_t s);

// tile shape of the 4-neighbour sweeps: SIMPOIS_HIP_AMLE_VARIANT (read at
// each call), else 32x32x16 (variant 6) for one plane, whose wavefronts hold
// few tiles, and 32x64x16 (variant 0) for several (measured on 4096^2
// planes); variant 0 for the larger neighbourhoods
static int amle_variant(int R, int pd)
{
	const char *e = getenv("SIMPOIS_HIP_AMLE_VARIANT");
	const int v = e ? atoi(e) : pd == 1 ? 6 : 0;
	return R == 1 && v > 0 && v < amle_shape_count(1) ? v : 0;
}

extern "C" int sp_amle_device(float *d_out, const float *d_in, const float *d_init,
		int w, int h, int pd, int nn, int niter, int trace, void *stream)
{
	TRY(amle_args(d_out, d_in, w, h, pd, nn, niter));
	TRY(ensure_device());
	hipStream_t s = (hipStream_t)stream;
	std::lock_guard<std::mutex> lk(g_amle_mu);
	const int R = amle_radius(nn);
	const int var = amle_variant(R, pd);
	trace = trace != 0;
	AmlePlan *A = g_amle;
	// a dependence wait of an earlier call that timed out (the calls are
	// asynchronous: it is reported by the next call, once that work is done)
	if (A && A->used && A->mem.gs_flow_used && hipEventQuery(A->done) == hipSuccess
			&& *A->mem.h_gs_err) {
		*A->mem.h_gs_err = 0;
		(void)hipMemset(A->mem.gs_err, 0, sizeof(int));
		return set_err(SP_ERR_HIP, "AMLE tile dependence wait timed out in an earlier call");
	}
	if (!A || A->w != w || A->h != h || A->pd != pd || A->nn != nn ||
			A->niter != niter || A->trace != trace || A->var != var) {
		// the old plan's scratch may still be in use on its stream
		if (A && A->used) HIPCHK(hipEventSynchronize(A->done));
		delete g_amle;
		g_amle = A = new AmlePlan;
		A->w = w; A->h = h; A->pd = pd; A->nn = nn; A->niter = niter;
		A->trace = trace; A->R = R; A->var = var;
		const size_t n = (size_t)w * h * pd;
		int rc = A->mem.alloc(&A->c, n);
		if (!rc && trace) rc = A->mem.alloc(&A->tinf, n);
		if (!rc && trace) rc = A->mem.alloc(&A->snap, 2 * n);
		if (!rc) rc = A->mem.alloc(&A->mm, pd);
		if (!rc) rc = A->mem.alloc((char **)&A->part, spk::amle_part_bytes(w * h, pd));
		if (!rc && !trace) rc = amle_build(*A, niter);
		for (int it = 0; !rc && trace && it < niter; it += 10) {
			rc = amle_build(*A, 1);
			if (!rc) rc = amle_build(*A, std::min(niter - it - 1, 9));
		}
		if (rc) { delete g_amle; g_amle = nullptr; return rc; }
	}
	// scratch reuse across streams: wait for the previous call's work
	if (A->used && A->last != s) HIPCHK(hipStreamWaitEvent(s, A->done, 0));
	if (!A->done) HIPCHK(hipEventCreateWithFlags(&A->done, hipEventDisableTiming));
	TRY(amle_bind(*A, d_out, s));
	const int rc = amle_run(*A, d_out, d_in, d_init, s);
	if (A->mem.gs_flow_used)
		HIPCHK(hipMemcpyAsync(A->mem.h_gs_err, A->mem.gs_err, sizeof(int),
				hipMemcpyDeviceToHost, s));
	HIPCHK(hipEventRecord(A->done, s));
	A->last = s;
	A->used = true;
	return rc;
}

static int amle_run(AmlePlan &Ar, float *d_out, const float *d_in, const float *d_init,
		hipStream_t s)
{
	AmlePlan *A = &Ar;
	const int w = A->w, h = A->h, pd = A->pd, niter = A->niter, trace = A->trace;
	const int n = w * h;
	{
		ProfScope ps("amle_prep", s);
		spk::launch_amle_prep(d_in, d_init, d_out, A->c, A->tinf, A->part, A->mm,
				n, pd, s);
		ps.add(3, 4.0 * n * pd * (trace ? 5 : 4));
	}
	HIPCHK(hipGetLastError());
	if (!trace) return amle_sweeps(*A, niter, s);
	// trace: every tenth sweep alone, between snapshots, then the diagnostic
	// line of each plane as amle.c:230-233 prints it
	const size_t np = (size_t)n * pd;
	const int nb = spk::amle_part_blocks(n);
	std::vector<double> hp((size_t)nb * pd * 5);
	for (int it = 0; it < niter; it += 10) {
		HIPCHK(hipMemcpyAsync(A->snap, d_out, np * 4, hipMemcpyDeviceToDevice, s));
		HIPCHK(hipMemcpyAsync(A->snap + np, A->tinf, np * 4, hipMemcpyDeviceToDevice, s));
		TRY(amle_sweeps(*A, 1, s));
		spk::launch_amle_diag(A->c, A->tinf, d_out, A->snap + np, A->snap, n, pd,
				(double *)A->part, s);
		HIPCHK(hipGetLastError());
		HIPCHK(hipMemcpyAsync(hp.data(), A->part, hp.size() * 8, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		for (int l = 0; l < pd; l++) {
			double cnt = 0, e = 0, ea = 0, ai = 0, as = 0;
			for (int b = 0; b < nb; b++) {
				const double *q = &hp[((size_t)l * nb + b) * 5];
				cnt += q[0];
				if (q[1] > e) e = q[1];
				ea += q[2]; ai += q[3]; as += q[4];
			}
			fprintf(stderr, "iter %d, e = {%g %g}, eactus = {%g %g}\n", it,
					(float)e, (float)((float)ea / (float)cnt), (float)ai, (float)as);
		}
		TRY(amle_sweeps(*A, std::min(niter - it - 1, 9), s));
	}
	return SP_OK;
}

extern "C" int sp_amle_host(float *out, const float *in, const float *init,
		int w, int h, int pd, int nn, int niter, int trace)
{
	TRY(amle_args(out, in, w, h, pd, nn, niter));
	TRY(ensure_device());
	const size_t n = (size_t)w * h * pd;
	Plan tmp;   // owns the device buffers
	float *d_in = nullptr, *d_out = nullptr, *d_init = nullptr;
	TRY(tmp.alloc(&d_in, n));
	TRY(tmp.alloc(&d_out, n));
	HIPCHK(hipMemcpy(d_in, in, n * 4, hipMemcpyHostToDevice));
	if (init) {
		TRY(tmp.alloc(&d_init, n));
		HIPCHK(hipMemcpy(d_init, init, n * 4, hipMemcpyHostToDevice));
	}
	TRY(sp_amle_device(d_out, d_in, d_init, w, h, pd, nn, niter, trace, nullptr));
	HIPCHK(hipMemcpy(out, d_out, n * 4, hipMemcpyDeviceToHost));
	return SP_OK;
}

// ------------------------------------------------------------------ heatd
//
// n explicit diffusion steps (heatd.c:58-103), T steps per launch on
// register tiles: isotropic k_heatd_multi (T = SIMPOIS_HIP_HEATD_T, default
// 6, 1..8), with a diffusion-matrix field k_heatd_multi_field (T =
// SIMPOIS_HIP_HEATD_FT, default HEATD_FIELD_T; tile rows
// SIMPOIS_HIP_HEATD_FR, 16, 24, 32 or 48, default HEATD_FIELD_ROWS); a single
// remaining step uses the one-step kernel.  Launches ping-pong between the
// output and a stream-ordered scratch plane set, so that the last one
// lands in the output.

constexpr int HEATD_FIELD_T = 5, HEATD_FIELD_ROWS = 32;   // measured best at 4096x4096x3

static int heatd_args(const float *out, const float *in, int w, int h, int pd, int n)
{
	if (!out || !in || w < 1 || h < 1 || pd < 1)
		return set_err(SP_ERR_ARG, "bad image arguments");
	if ((long)w * h * pd > 0x7fffffffL) return set_err(SP_ERR_ARG, "image too large");
	if (n < 0) return set_err(SP_ERR_ARG, "negative number of diffusion steps");
	if (out == in) return set_err(SP_ERR_ARG, "out must not alias in");
	return SP_OK;
}

extern "C" int sp_heatd_device(float *d_out, const float *d_in, const float *d_A,
		int w, int h, int pd, int n, float t, void *stream)
{
	TRY(heatd_args(d_out, d_in, w, h, pd, n));
	TRY(ensure_device());
	hipStream_t s = (hipStream_t)stream;
	const size_t np = (size_t)w * h * pd;
	if (n == 0) {
		HIPCHK(hipMemcpyAsync(d_out, d_in, np * 4, hipMemcpyDeviceToDevice, s));
		return SP_OK;
	}
	// steps per launch and (with a field) tile rows, read at each call
	const char *tv = getenv(d_A ? "SIMPOIS_HIP_HEATD_FT" : "SIMPOIS_HIP_HEATD_T");
	int T = tv ? atoi(tv) : d_A ? HEATD_FIELD_T : 6;
	T = T < 1 ? 1 : T > spk::heatd_multi_max() ? spk::heatd_multi_max() : T;
	const char *rv = getenv("SIMPOIS_HIP_HEATD_FR");
	const int rv_n = rv ? atoi(rv) : 0;
	const int rows = !d_A ? 48 : (rv_n == 16 || rv_n == 24 || rv_n == 32 || rv_n == 48)
		? rv_n : HEATD_FIELD_ROWS;
	if (d_A && T > (rows + 1) / 3) T = (rows + 1) / 3;   // rows >= 3T - 1
	// T steps per launch; a remainder, or a chunk the planes are too small
	// for (smaller T needs wider and taller planes), as single steps
	std::vector<int> chunks;
	for (int rem = n; rem > 0; rem -= chunks.back()) {
		const int k = std::min(rem, T);
		chunks.push_back(k > 1 && spk::heatd_multi_fits(w, h, k, rows) ? k : 1);
	}
	const int L = (int)chunks.size();
	float *tmp = nullptr;
	if (L > 1) HIPCHK(hipMallocAsync((void **)&tmp, np * 4, s));
	ProfScope ps(T > 1 ? "heatd_multi" : "heatd_step", s);
	const float *src = d_in;
	for (int k = 0; k < L; k++) {
		float *dst = (L - 1 - k) & 1 ? tmp : d_out;
		const bool ok = chunks[k] == 1
			? (spk::launch_heatd_step(dst, src, d_A, w, h, pd, t, s), true)
			: d_A ? spk::launch_heatd_multi_field(dst, src, d_A, w, h, pd, chunks[k], rows,
					t, s)
			: spk::launch_heatd_multi(dst, src, w, h, pd, chunks[k], t, s);
		if (!ok) {
			if (tmp) (void)hipFreeAsync(tmp, s);
			return set_err(SP_ERR_ARG, "heatd: planes too small for %d steps per launch",
					chunks[k]);
		}
		src = dst;
	}
	// algorithmic bytes: the planes read and written once per launch, the
	// field (3 floats per pixel) read once per launch and plane
	ps.add(L, (double)L * np * (8.0 + (d_A ? 12.0 : 0.0)));
	const hipError_t e = hipGetLastError();
	if (tmp) HIPCHK(hipFreeAsync(tmp, s));
	HIPCHK(e);
	return SP_OK;
}

extern "C" int sp_heatd_host(float *out, const float *in, const float *A,
		int w, int h, int pd, int n, float t)
{
	TRY(heatd_args(out, in, w, h, pd, n));
	TRY(ensure_device());
	const size_t np = (size_t)w * h * pd;
	Plan tmp;   // owns the device buffers
	float *d_in = nullptr, *d_out = nullptr, *d_A = nullptr;
	TRY(tmp.alloc(&d_in, np));
	TRY(tmp.alloc(&d_out, np));
	HIPCHK(hipMemcpy(d_in, in, np * 4, hipMemcpyHostToDevice));
	if (A) {
		TRY(tmp.alloc(&d_A, (size_t)w * h * 3));
		HIPCHK(hipMemcpy(d_A, A, (size_t)w * h * 12, hipMemcpyHostToDevice));
	}
	TRY(sp_heatd_device(d_out, d_in, d_A, w, h, pd, n, t, nullptr));
	HIPCHK(hipMemcpy(out, d_out, np * 4, hipMemcpyDeviceToHost));
	return SP_OK;
}

// drop-ins of heatd.c:58-103 (host buffers, synchronous; failures print a
// message and exit as the reference's allocation failures do)
static void heatd_or_die(int rc)
{
	if (rc) {
		fprintf(stderr, "heatd: %s\n", sp_last_error());
		exit(1);
	}
}

extern "C" void linear_diffusion_separable(float *y, float *x, int w, int h, int pd,
		float *A, int n, float t)
{
	if (n <= 0) return;   // the reference writes neither y nor x
	heatd_or_die(sp_heatd_host(y, x, A, w, h, pd, n, t));
	memcpy(x, y, (size_t)w * h * pd * sizeof(float));   // x ends as y (:84-85)
}

extern "C" void linear_diffusion(float *y, float *x, int w, int h, float *A, int n,
		float t)
{
	linear_diffusion_separable(y, x, w, h, 1, A, n, t);
}

extern "C" void linear_diffusion_step(float *y, float *x, int w, int h, float *A,
		float t)
{
	heatd_or_die(sp_heatd_host(y, x, A, w, h, 1, 1, t));
}

// ------------------------------------------------- plambda's Laplacian

static int lap_args(const float *out, const float *in, int w, int h, int pd,
		int variant, int boundary)
{
	if (!out || !in || w < 1 || h < 1 || pd < 1)
		return set_err(SP_ERR_ARG, "bad image arguments");
	if ((long)w * h * pd > 0x7fffffffL)
		return set_err(SP_ERR_ARG, "image too large");
	if (w > 65535 * 1024)   // column chunks of 1024 pixels per grid row
		return set_err(SP_ERR_ARG, "image too wide");
	if (variant != 0 && variant != 1)
		return set_err(SP_ERR_ARG, "variant must be 0 (x,l) or 1 (x,lr)");
	if (boundary < 0 || boundary > 4)
		return set_err(SP_ERR_ARG, "boundary rule must be 0..4");
	if (out == in) return set_err(SP_ERR_ARG, "out must not alias in");
	return SP_OK;
}

extern "C" int sp_plambda_lap_device(float *d_out, const float *d_in, int w, int h,
		int pd, int variant, int boundary, void *stream)
{
	TRY(lap_args(d_out, d_in, w, h, pd, variant, boundary));
	TRY(ensure_device());
	spk::launch_plambda_lap(d_out, d_in, w, h, pd, variant, boundary,
			(hipStream_t)stream);
	HIPCHK(hipGetLastError());
	return SP_OK;
}

extern "C" int sp_plambda_lap_host(float *out, const float *in, int w, int h, int pd,
		int variant, int boundary)
{
	TRY(lap_args(out, in, w, h, pd, variant, boundary));
	TRY(ensure_device());
	const size_t n = (size_t)w * h * pd;
	Plan tmp;   // owns the device buffers
	float *d_in = nullptr, *d_out = nullptr;
	TRY(tmp.alloc(&d_in, n));
	TRY(tmp.alloc(&d_out, n));
	HIPCHK(hipMemcpy(d_in, in, n * 4, hipMemcpyHostToDevice));
	spk::launch_plambda_lap(d_out, d_in, w, h, pd, variant, boundary, nullptr);
	HIPCHK(hipGetLastError());
	HIPCHK(hipMemcpy(out, d_out, n * 4, hipMemcpyDeviceToHost));
	return SP_OK;
}

// ------------------------------------------------------------ stream copy

extern "C" int sp_stream_copy(float *d_dst, const float *d_src, size_t n, void *stream)
{
	if (!d_dst || !d_src || (n & 3) || ((uintptr_t)d_dst & 15) || ((uintptr_t)d_src & 15))
		return set_err(SP_ERR_ARG, "stream copy needs 16-byte aligned buffers "
				"and a multiple of 4 floats");
	TRY(ensure_device());
	spk::launch_stream_copy(d_dst, d_src, n / 4, (hipStream_t)stream);
	HIPCHK(hipGetLastError());
	return SP_OK;
}

// ------------------------------------------------------------ residual

extern "C" int sp_residual_device(const float *d_u, const float *d_g,
		const float *d_f, int w, int h, int pd, int bilap, double *rel, void *stream)
{
	if (!d_u || !d_g || !rel || w < 1 || h < 1 || pd < 1)
		return set_err(SP_ERR_ARG, "bad residual arguments");
	TRY(ensure_device());
	hipStream_t s = (hipStream_t)stream;
	Plan pl;
	const size_t n = (size_t)w * h;
	std::vector<SpPlane> planes(pd);
	for (int p = 0; p < pd; p++)
		planes[p] = {const_cast<float *>(d_g) + n * p,
			d_f ? const_cast<float *>(d_f) + n * p : nullptr,
			const_cast<float *>(d_u) + n * p, nullptr, nullptr, w, h};
	std::vector<BlockWork> bw;
	std::vector<int> first(1, 0);
	for (int p = 0; p < pd; p++) {
		for (size_t st = 0; st < n; st += EW_SPAN) bw.push_back({p, (int)st});
		first.push_back((int)bw.size());
	}
	SpPlane *dp; BlockWork *dbw; int *dfirst; double *part, *out;
	TRY(pl.upload(&dp, planes));
	TRY(pl.upload(&dbw, bw));
	TRY(pl.upload(&dfirst, first));
	TRY(pl.alloc(&part, 2 * bw.size()));
	TRY(pl.alloc(&out, 3 * pd));
	spk::launch_residual(dp, dbw, (int)bw.size(), bilap, part, s);
	spk::launch_residual_reduce(part, dfirst, pd, out, s);
	HIPCHK(hipGetLastError());
	std::vector<double> h3(3 * pd);
	HIPCHK(hipMemcpyAsync(h3.data(), out, 3 * pd * sizeof(double),
			hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	for (int p = 0; p < pd; p++) rel[p] = h3[3 * p + 2];
	return SP_OK;
}

// ------------------------------------------- residual without allocation
//
// Device memory of sp_residual_device_noalloc: one fixed array per GPU,
// held until the process ends.  In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is used.
constexpr size_t SP_NOALLOC_SCRATCH_BYTES = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[SP_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_dev;   // address on each device

static int scratch_address(unsigned char **p)
{
	int dev = 0;
	HIPCHK(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
	if (!g_scratch_dev[dev]) {
		void *q = nullptr;
		const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !q)
			return set_err(SP_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
					dev, hipGetErrorString(e));
		g_scratch_dev[dev] = (unsigned char *)q;
	}
	*p = g_scratch_dev[dev];
	return SP_OK;
}

extern "C" int sp_residual_device_noalloc(const float *d_u, const float *d_g,
		const float *d_f, int w, int h, int pd, int bilap, double *rel, void *stream)
{
	if (!d_u || !d_g || !rel || w < 1 || h < 1 || pd < 1)
		return set_err(SP_ERR_ARG, "bad residual arguments");
	const size_t n = (size_t)w * h;
	if (n > 0x7fffffffUL) return set_err(SP_ERR_ARG, "plane too large");
	// scratch layout: descriptors, work blocks, first block of each plane,
	// per-block partial sums, per-plane (rr, bb, rel); every entry a kernel
	// reads is written earlier in the same call, so nothing is cleared
	const size_t per = (n + EW_SPAN - 1) / EW_SPAN, nb = per * (size_t)pd;
	size_t off = 0;
	auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
	const size_t o_lv = take((size_t)pd * sizeof(SpPlane)),
		o_bw = take(nb * sizeof(BlockWork)), o_first = take(((size_t)pd + 1) * sizeof(int)),
		o_part = take(2 * nb * sizeof(double)), o_out = take(3 * (size_t)pd * sizeof(double));
	if (off > SP_NOALLOC_SCRATCH_BYTES)
		return set_err(SP_ERR_ARG, "residual of %d planes of %dx%d needs %zu bytes of "
				"scratch, more than the %zu the allocation-free form holds", pd, w, h,
				off, SP_NOALLOC_SCRATCH_BYTES);
	TRY(ensure_device());
	unsigned char *base = nullptr;
	TRY(scratch_address(&base));
	hipStream_t s = (hipStream_t)stream;
	SpPlane *lv = (SpPlane *)(base + o_lv);
	BlockWork *bw = (BlockWork *)(base + o_bw);
	int *first = (int *)(base + o_first);
	double *part = (double *)(base + o_part), *out = (double *)(base + o_out);
	spk::launch_residual_plan(lv, bw, first, d_u, d_g, d_f, w, h, pd, (int)per, s);
	spk::launch_residual(lv, bw, (int)nb, bilap, part, s);
	spk::launch_residual_reduce(part, first, pd, out, s);
	HIPCHK(hipGetLastError());
	// rel[p] = out[3p + 2]: one strided copy, no host buffer
	HIPCHK(hipMemcpy2DAsync(rel, sizeof(double), out + 2, 3 * sizeof(double),
			sizeof(double), pd, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	return SP_OK;
}

// --------------------------------------------------------- CG drop-ins

// single-level masked CG from `init`, cleant_cgpois.c:197-264
static int cg_extend_host(float *out, const float *in, const float *dat, int w,
		int h, const float *init, int maxit, float eps, int bilap)
{
	if (!out || !in || !init || w < 1 || h < 1)
		return set_err(SP_ERR_ARG, "bad CG arguments");
	TRY(ensure_device());
	const size_t n = (size_t)w * h;
	sp_params P;
	sp_params_default(&P);
	P.niter = 0; P.scale = 1; P.cgit = maxit;
	P.tstep = bilap ? -1.0f : 1.0f;
	P.ponlit = 0;
	float *d_in, *d_dat = nullptr, *d_u;
	Plan pl;
	TRY(pl.alloc(&d_in, n));
	TRY(pl.alloc(&d_u, n));
	if (dat) TRY(pl.alloc(&d_dat, n));
	HIPCHK(hipMemcpy(d_in, in, n * 4, hipMemcpyHostToDevice));
	if (dat) HIPCHK(hipMemcpy(d_dat, dat, n * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_u, init, n * 4, hipMemcpyHostToDevice));
	TRY(plan_create(pl, {PlaneIn{w, h, d_in, d_dat, d_u}}, P, dat != nullptr));
	hipStream_t s = nullptr;
	LevelPlan &L0 = pl.lv[0];
	spk::launch_prepare(L0.d_sp, L0.d_bw, L0.nbw, L0.d_bcnt, s);
	spk::launch_count(L0.d_sp, 1, L0.d_first, L0.d_bcnt, s);
	TRY(run_cg64(pl, 0, 0, maxit, bilap, eps, s));
	// :251-256 and the nmask == 0 branch :203-207
	int nmask = 0;
	HIPCHK(hipMemcpy(&nmask, pl.d_cnt, sizeof(int), hipMemcpyDeviceToHost));
	std::vector<float> u(n);
	HIPCHK(hipMemcpy(u.data(), d_u, n * 4, hipMemcpyDeviceToHost));
	for (size_t q = 0; q < n; q++) {
		if (nmask == 0) out[q] = std::isfinite(in[q]) ? in[q] : init[q];
		else out[q] = std::isnan(in[q]) ? u[q] : in[q];
	}
	if (pl.P.trace) {
		const CgTrace &t = pl.traces[0][0];
		for (size_t i = 0; i < t.rr.size(); i++)
			fprintf(stderr, "n=%d iter=%d, rr_new=%g\n", t.n, (int)i, t.rr[i]);
	}
	return SP_OK;
}

static void die_on(int rc, const char *where)
{
	if (rc) {
		fprintf(stderr, "%s: %s\n", where, sp_last_error());
		exit(1);
	}
}

extern "C" void poisson_extension_by_cg(float *out, float *in, float *dat,
		int w, int h, float *init, int maxit, float eps)
{
	die_on(cg_extend_host(out, in, dat, w, h, init, maxit, eps, 0),
			"poisson_extension_by_cg");
}

extern "C" void biharmonic_extension_by_cg(float *out, float *in, float *dat,
		int w, int h, float *init, int maxit, float eps)
{
	die_on(cg_extend_host(out, in, dat, w, h, init, maxit, eps, 1),
			"biharmonic_extension_by_cg");
}

extern "C" void poisson_solver_separable(float *out, float *in, float *dat,
		int w, int h, int pd, float tstep, int niter, int scale, float cgrad)
{
	sp_params P;
	sp_params_default(&P);
	P.tstep = tstep; P.niter = niter; P.scale = scale; P.cgit = (int)cgrad;
	die_on(sp_solve_host(out, in, dat, w, h, pd, &P, nullptr),
			"poisson_solver_separable");
}

extern "C" void poisson_rec(float *u, float *g, float *f, int w, int h,
		float tstep, int niter, int scale, int cgit)
{
	sp_params P;
	sp_params_default(&P);
	P.tstep = tstep; P.niter = niter; P.scale = scale; P.cgit = cgit;
	die_on(sp_solve_host(u, g, f, w, h, 1, &P, nullptr), "poisson_rec");
}
