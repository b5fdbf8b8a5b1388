// Host driver of the simpois path: pyramid plan, Gauss-Seidel scheduling,
// CG driver and the C ABI (include/simpois_hip.h).
//
// Call structure restated from the reference (simpois.c:247-302):
//   per plane: restrict g (and f, times PMSFAC) down the pyramid; from the
//   coarsest level up: u = isfinite(g) ? g : prolong(coarse output);
//   niter Gauss-Seidel sweeps; CG when -c is non-zero; the level's output
//   is the input at non-NaN pixels when the level has holes.
// Here all planes of a call advance together, level by level, on one HIP
// stream; results per plane are identical to the reference's
// channel-by-channel loop because planes never interact.
#include "sp_internal.h"
#include "sp_tiles.h"
#include "../../include/simpois_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
	p->tstep = 0.25f;
	p->niter = 10;
	p->scale = 99;
	p->cgit = 0;
	p->pmsfac = env_knob("PMSFAC", 3);
	p->ponlit = env_knob("PONLIT", 0);
	p->cg_eps = env_knob("CG_EPS", -1);
	const char *q = getenv("SIMPOIS_HIP_QUIET");
	p->trace = !(q && *q && strcmp(q, "0"));
}

// ------------------------------------------------------------- device

static int g_dev_checked = 0, g_dev_ok = 0, g_max_lds = 65536;

static int ensure_device()
{
	static std::mutex mu;
	std::lock_guard<std::mutex> lk(mu);
	if (g_dev_checked) return g_dev_ok ? SP_OK : set_err(SP_ERR_NODEV, "%s", g_err.c_str());
	g_dev_checked = 1;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
		return set_err(SP_ERR_NODEV, "no HIP device visible");
	int dev = 0;
	hipGetDevice(&dev);
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
		return set_err(SP_ERR_NODEV, "hipGetDeviceProperties failed");
	if (!strstr(prop.gcnArchName, "gfx950"))
		return set_err(SP_ERR_NODEV, "device %s is not gfx950", prop.gcnArchName);
	int lds = 0;
	if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)
			== hipSuccess && lds > 0)
		g_max_lds = std::min(lds, 65536);
	g_dev_ok = 1;
	return SP_OK;
}

// --------------------------------------------------------------- plan

struct LevelPlan {
	int w = 0, h = 0;
	int nsweep = 0, cgit = 0;
	bool small = false;
	int small_threads = 64, small_lds = 0;
	std::vector<int> wf_off;                // tile wavefront offsets
	GsTile *d_tiles = nullptr;
	GsInst *d_insts = nullptr;
	int *d_ids = nullptr;
	SpPlane *d_planes = nullptr;
	BlockWork *d_bw = nullptr;
	int nbw = 0;
	std::vector<SpPlane> planes;            // host copy of d_planes
};

struct CgTrace { int n; std::vector<double> rr; };

struct Plan {
	int w = 0, h = 0, pd = 0;
	bool has_f = false;
	sp_params P{};
	int R = 1;
	float teff = 0;
	std::vector<LevelPlan> lv;
	std::vector<void *> allocs;
	int *d_cnt = nullptr;          // nlev * pd hole counters
	float **d_outs = nullptr;      // pd output pointers
	float *c0 = nullptr;           // level-0 coefficient planes
	// CG scratch, sized for the finest level
	double *cg_x = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_q = nullptr,
		*cg_part = nullptr, *cg_rr = nullptr;
	int cg_rr_cap = 0;
	int *cg_done = nullptr, *h_done = nullptr;
	std::vector<std::vector<CgTrace>> traces;   // [plane][level]
	const float *bound_in = nullptr, *bound_dat = nullptr;
	float *bound_out = nullptr;

	~Plan()
	{
		for (void *p : allocs) hipFree(p);
		if (h_done) hipHostFree(h_done);
	}
	template <typename T> int alloc(T **p, size_t n)
	{
		void *q = nullptr;
		HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
		allocs.push_back(q);
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

static std::vector<BlockWork> block_list(int w, int h, int pd)
{
	std::vector<BlockWork> bw;
	size_t n = (size_t)w * h, span = 256 * 4;
	for (int p = 0; p < pd; p++)
		for (size_t s = 0; s < n; s += span)
			bw.push_back({p, (int)s});
	return bw;
}

static int plan_create(Plan &pl, int w, int h, int pd, bool has_f,
		const sp_params &P)
{
	pl.w = w; pl.h = h; pl.pd = pd; pl.has_f = has_f; pl.P = P;
	pl.R = op_radius(P.tstep, &pl.teff);
	// pyramid shapes, simpois.c:253-256
	{
		int cw = w, ch = h, sc = P.scale;
		for (;;) {
			LevelPlan L;
			L.w = cw; L.h = ch;
			pl.lv.push_back(L);
			if (!(sc > 1 && (cw > 1 || ch > 1))) break;
			cw = (int)ceil(cw / 2.0);
			ch = (int)ceil(ch / 2.0);
			sc--;
		}
	}
	const int nlev = (int)pl.lv.size();
	TRY(pl.alloc(&pl.d_cnt, (size_t)nlev * pd));
	TRY(pl.alloc(&pl.d_outs, pd));
	TRY(pl.alloc(&pl.c0, (size_t)w * h * pd));
	HIPCHK(hipHostMalloc((void **)&pl.h_done, 4 * sizeof(int)));
	pl.traces.assign(pd, std::vector<CgTrace>(nlev));
	for (int k = 0; k < nlev; k++) {
		LevelPlan &L = pl.lv[k];
		const size_t n = (size_t)L.w * L.h;
		float *g = nullptr, *f = nullptr, *u = nullptr, *c = nullptr;
		if (k > 0) {
			TRY(pl.alloc(&g, n * pd));
			if (has_f) TRY(pl.alloc(&f, n * pd));
			TRY(pl.alloc(&u, n * pd));
			TRY(pl.alloc(&c, n * pd));
		} else {
			c = pl.c0;
		}
		L.planes.resize(pd);
		for (int p = 0; p < pd; p++) {
			SpPlane &S = L.planes[p];
			S.g = g ? g + n * p : nullptr;
			S.f = f ? f + n * p : nullptr;
			S.u = u ? u + n * p : nullptr;
			S.c = c + n * p;
			S.cnt = pl.d_cnt + (size_t)k * pd + p;
			S.w = L.w; S.h = L.h;
		}
		TRY(pl.upload(&L.d_planes, L.planes));
		auto bw = block_list(L.w, L.h, pd);
		L.nbw = (int)bw.size();
		TRY(pl.upload(&L.d_bw, bw));

		// PONLIT restricts the work to the level of that width (:272-273)
		L.nsweep = std::max(P.niter, 0);
		L.cgit = P.cgit;
		if (P.ponlit != 0 && P.ponlit != (double)L.w) L.nsweep = L.cgit = 0;

		std::vector<GsInst> gi(pd);
		for (int p = 0; p < pd; p++)
			gi[p] = {L.planes[p].u, L.planes[p].c, L.w, L.h, L.nsweep, pl.teff, nullptr};
		TRY(pl.upload(&L.d_insts, gi));
		std::vector<int> ids(pd);
		for (int p = 0; p < pd; p++) ids[p] = p;
		TRY(pl.upload(&L.d_ids, ids));
		if (L.nsweep > 0) {
			L.small_lds = spk::small_lds_bytes(L.w, L.h);
			L.small = (long)8 * L.w * L.h <= g_max_lds;
			if (L.small) {
				const int ct = pl.R * pl.R + 1;
				long slots = (long)L.h * ((L.w - 1) / ct + 1);
				long th = (slots + 63) / 64 * 64;
				L.small_threads = (int)std::min<long>(1024, std::max<long>(64, th));
			} else {
				std::vector<GsTile> tiles;
				if (pl.R == 1) build_tiles<1>(L.w, L.h, L.nsweep, pd, tiles, L.wf_off);
				else if (pl.R == 2) build_tiles<2>(L.w, L.h, L.nsweep, pd, tiles, L.wf_off);
				else build_tiles<3>(L.w, L.h, L.nsweep, pd, tiles, L.wf_off);
				TRY(pl.upload(&L.d_tiles, tiles));
			}
		}
	}
	return SP_OK;
}

// point level 0 at the caller's planes (device pointers)
static int plan_bind(Plan &pl, float *d_out, const float *d_in, const float *d_dat,
		hipStream_t s)
{
	if (pl.bound_in == d_in && pl.bound_dat == d_dat && pl.bound_out == d_out)
		return SP_OK;
	LevelPlan &L = pl.lv[0];
	const size_t n = (size_t)pl.w * pl.h;
	std::vector<float *> outs(pl.pd);
	for (int p = 0; p < pl.pd; p++) {
		SpPlane &S = L.planes[p];
		S.g = const_cast<float *>(d_in) + n * p;
		S.f = d_dat ? const_cast<float *>(d_dat) + n * p : nullptr;
		S.u = d_out + n * p;
		outs[p] = d_out + n * p;
	}
	HIPCHK(hipMemcpyAsync(L.d_planes, L.planes.data(), pl.pd * sizeof(SpPlane),
			hipMemcpyHostToDevice, s));
	std::vector<GsInst> gi(pl.pd);
	for (int p = 0; p < pl.pd; p++)
		gi[p] = {L.planes[p].u, L.planes[p].c, L.w, L.h, L.nsweep, pl.teff, nullptr};
	HIPCHK(hipMemcpyAsync(L.d_insts, gi.data(), pl.pd * sizeof(GsInst),
			hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(pl.d_outs, outs.data(), pl.pd * sizeof(float *),
			hipMemcpyHostToDevice, s));
	HIPCHK(hipStreamSynchronize(s));   // host vectors die here
	pl.bound_in = d_in; pl.bound_dat = d_dat; pl.bound_out = d_out;
	return SP_OK;
}

static int run_gs(Plan &pl, LevelPlan &L, hipStream_t s)
{
	if (L.nsweep <= 0) return SP_OK;
	if (L.small) {
		spk::launch_gs_small(pl.R, L.d_insts, L.d_ids, pl.pd, L.small_lds,
				L.small_threads, s);
	} else {
		for (size_t k = 0; k + 1 < L.wf_off.size(); k++)
			spk::launch_gs_tiles(pl.R, L.d_tiles + L.wf_off[k],
					L.wf_off[k + 1] - L.wf_off[k], L.d_insts, s);
	}
	HIPCHK(hipGetLastError());
	return SP_OK;
}

static int cg_ensure(Plan &pl, size_t n, int nblk, int maxit)
{
	if (!pl.cg_x) {
		const size_t n0 = (size_t)pl.w * pl.h;
		TRY(pl.alloc(&pl.cg_x, n0));
		TRY(pl.alloc(&pl.cg_r, n0));
		TRY(pl.alloc(&pl.cg_p, n0));
		TRY(pl.alloc(&pl.cg_q, n0));
		TRY(pl.alloc(&pl.cg_part, 2 * 1024));
		TRY(pl.alloc(&pl.cg_done, 4));
	}
	(void)n; (void)nblk;
	if (maxit + 2 > pl.cg_rr_cap) {
		int cap = std::max(maxit + 2, 1024);
		TRY(pl.alloc(&pl.cg_rr, cap));
		pl.cg_rr_cap = cap;
	}
	return SP_OK;
}

// masked CG on one plane of one level (cleant_cgpois.c:197-264), bilap
// selects the biharmonic operator; eps is the built-in threshold
static int run_cg_plane(Plan &pl, int k, int p, int cgit, int bilap, float eps,
		hipStream_t s)
{
	LevelPlan &L = pl.lv[k];
	SpPlane &S = L.planes[p];
	int nmask = 0;
	HIPCHK(hipMemcpyAsync(&nmask, S.cnt, sizeof(int), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	if (nmask == 0) return SP_OK;   // :203-207, the level rule covers it
	float e = pl.P.cg_eps >= 0 ? (float)pl.P.cg_eps : eps;
	float fm = cgit == -1 ? (float)nmask : (float)cgit;
	int maxit = (int)fm;
	if (maxit <= 0) return SP_OK;   // x = x0: u is unchanged
	const size_t n = (size_t)L.w * L.h;
	int nblk = (int)std::min<size_t>(1024, std::max<size_t>(1, (n + 2047) / 2048));
	TRY(cg_ensure(pl, n, nblk, maxit));
	spk::CgDev d;
	d.g = S.g; d.f = S.f; d.u = S.u;
	d.x = pl.cg_x; d.r = pl.cg_r; d.p = pl.cg_p; d.q = pl.cg_q;
	d.part = pl.cg_part; d.rr = pl.cg_rr; d.done = pl.cg_done;
	d.w = L.w; d.h = L.h; d.bilap = bilap; d.nblk = nblk;
	d.stop = (double)e;
	spk::launch_cg_setup(d, s);
	int it = 0, chunk = 16, ran = maxit;
	while (it < maxit) {
		int m = std::min(chunk, maxit - it);
		for (int c = 0; c < m; c++) spk::launch_cg_iter(d, it + c, s);
		it += m;
		HIPCHK(hipGetLastError());
		HIPCHK(hipMemcpyAsync(pl.h_done, pl.cg_done, 2 * sizeof(int),
				hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (pl.h_done[0]) { ran = pl.h_done[1]; break; }
		chunk = std::min(chunk * 2, 256);
	}
	spk::launch_cg_finish(d, s);
	HIPCHK(hipGetLastError());
	CgTrace &tr = pl.traces[p][k];
	tr.n = nmask;
	tr.rr.resize(ran);
	HIPCHK(hipMemcpyAsync(tr.rr.data(), pl.cg_rr + 1, ran * sizeof(double),
			hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	return SP_OK;
}

static int plan_execute(Plan &pl, hipStream_t s, sp_stats *st)
{
	const int nlev = (int)pl.lv.size();
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	HIPCHK(hipEventRecord(e0, s));
	for (auto &v : pl.traces) for (auto &t : v) { t.n = 0; t.rr.clear(); }
	HIPCHK(hipMemsetAsync(pl.d_cnt, 0, sizeof(int) * nlev * pl.pd, s));
	// pyramid of data, simpois.c:260-262
	spk::launch_prepare(pl.lv[0].d_planes, pl.lv[0].d_bw, pl.lv[0].nbw, s);
	for (int k = 1; k < nlev; k++)
		spk::launch_restrict(pl.lv[k - 1].d_planes, pl.lv[k].d_planes,
				pl.lv[k].d_bw, pl.lv[k].nbw, pl.P.pmsfac, s);
	HIPCHK(hipGetLastError());
	// coarse to fine
	for (int k = nlev - 1; k >= 0; k--) {
		LevelPlan &L = pl.lv[k];
		spk::launch_init(L.d_planes, k + 1 < nlev ? pl.lv[k + 1].d_planes : nullptr,
				L.d_bw, L.nbw, s);
		TRY(run_gs(pl, L, s));
		if (L.cgit) {   // simpois.c:278-284; cgit == 0 is a no-op there
			const bool pois = pl.P.tstep > 0;
			for (int p = 0; p < pl.pd; p++)
				TRY(run_cg_plane(pl, k, p, L.cgit, pois ? 0 : 1,
						pois ? 1e-6f : 1e-9f, s));
		}
	}
	spk::launch_finalize(pl.lv[0].d_planes, pl.d_outs, pl.lv[0].d_bw,
			pl.lv[0].nbw, s);
	HIPCHK(hipGetLastError());
	HIPCHK(hipEventRecord(e1, s));
	HIPCHK(hipEventSynchronize(e1));
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	if (st) {
		st->levels = nlev;
		st->cg_iters_finest = (int)pl.traces[0][0].rr.size();
		st->elapsed_ms = ms;
	}
	if (pl.P.trace)   // reference order: channel by channel, coarse to fine
		for (int p = 0; p < pl.pd; p++)
			for (int k = nlev - 1; k >= 0; k--) {
				const CgTrace &t = pl.traces[p][k];
				for (size_t i = 0; i < t.rr.size(); i++)
					fprintf(stderr, "n=%d iter=%d, rr_new=%g\n", t.n, (int)i, t.rr[i]);
			}
	return SP_OK;
}

static int check_args(const void *out, const void *in, int w, int h, int pd,
		const sp_params *p)
{
	if (!out || !in || !p) return set_err(SP_ERR_ARG, "NULL buffer or params");
	if (w < 1 || h < 1 || pd < 1) return set_err(SP_ERR_ARG, "bad size %dx%dx%d", w, h, pd);
	if ((long)w * h > (1L << 30)) return set_err(SP_ERR_ARG, "image too large");
	if (out == in) return set_err(SP_ERR_ARG, "out must not alias in");
	return SP_OK;
}

extern "C" int sp_solve_device(float *d_out, const float *d_in, const float *d_dat,
		int w, int h, int pd, const sp_params *p, void *stream, sp_stats *st)
{
	TRY(check_args(d_out, d_in, w, h, pd, p));
	TRY(ensure_device());
	hipStream_t s = (hipStream_t)stream;
	Plan pl;
	TRY(plan_create(pl, w, h, pd, d_dat != nullptr, *p));
	TRY(plan_bind(pl, d_out, d_in, d_dat, s));
	return plan_execute(pl, s, st);
}

extern "C" int sp_solve_host(float *out, const float *in, const float *dat,
		int w, int h, int pd, const sp_params *p, sp_stats *st)
{
	TRY(check_args(out, in, w, h, pd, p));
	TRY(ensure_device());
	const size_t n = (size_t)w * h * pd;
	float *d_in = nullptr, *d_dat = nullptr, *d_out = nullptr;
	int rc = SP_OK;
	auto cleanup = [&]() { hipFree(d_in); hipFree(d_dat); hipFree(d_out); };
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
	auto bw = block_list(w, h, pd);
	std::vector<int> first(pd + 1, 0);
	for (size_t k = 0; k < bw.size(); k++) first[bw[k].plane + 1] = (int)k + 1;
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

// --------------------------------------------------------- CG drop-ins

// single-level masked CG from `init`, cleant_cgpois.c:197-264
static int cg_extend_host(float *out, const float *in, const float *dat, int w,
		int h, const float *init, int maxit, float eps, int bilap)
{
	if (!out || !in || !init || w < 1 || h < 1)
		return set_err(SP_ERR_ARG, "bad CG arguments");
	TRY(ensure_device());
	sp_params P;
	sp_params_default(&P);
	P.niter = 0; P.scale = 1; P.cgit = maxit;
	P.tstep = bilap ? -1.0f : 1.0f;
	P.ponlit = 0;
	Plan pl;
	TRY(plan_create(pl, w, h, 1, dat != nullptr, P));
	const size_t n = (size_t)w * h;
	float *d_in, *d_dat = nullptr, *d_u;
	TRY(pl.alloc(&d_in, n));
	TRY(pl.alloc(&d_u, n));
	if (dat) TRY(pl.alloc(&d_dat, n));
	HIPCHK(hipMemcpy(d_in, in, n * 4, hipMemcpyHostToDevice));
	if (dat) HIPCHK(hipMemcpy(d_dat, dat, n * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_u, init, n * 4, hipMemcpyHostToDevice));
	hipStream_t s = nullptr;
	TRY(plan_bind(pl, d_u, d_in, d_dat, s));
	HIPCHK(hipMemsetAsync(pl.d_cnt, 0, sizeof(int), s));
	spk::launch_prepare(pl.lv[0].d_planes, pl.lv[0].d_bw, pl.lv[0].nbw, s);
	TRY(run_cg_plane(pl, 0, 0, maxit, bilap, eps, s));
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
