// CPU emulation of the Gauss-Seidel kernels, used by the test-suite.
//
// Runs the device kernels' own phase functions (imscript_amd/csrc/sp_gs.h)
// on the CPU, thread by thread, with a barrier between phases, and compares
// the result with the oracle's raster-order sweeps (oracle/simpois_oracle.c)
// bit for bit.  Every tile of a wavefront loads before any tile of that
// wavefront computes or stores, which is the worst interleaving a launch
// allows, and a pixel stored by two tiles of one wavefront is reported.
// Built with AddressSanitizer by the tests, so any LDS or image access out
// of bounds aborts the run.
//
// usage: gs_emulate W H T tstep seed holefrac withf mode [variant]
//   mode: tiles | flow | small | cells; variant: tile shape variant of the radius (sp_gs.h)
//        gs_emulate amle W H T nn seed holefrac [variant [flow]]
//   the AMLE sweeps (sp_amle.h) against the oracle's raster-order sweeps
//   (oracle/amle_oracle.c); variant: AMLE tile shape of radius amle_radius(nn)
// prints "OK <n>" (n = holes) or "MISMATCH ..." / "CONFLICT ..." and exits 1
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "../../imscript_amd/csrc/sp_gs.h"
#include "../../imscript_amd/csrc/sp_tiles.h"
#include "../../imscript_amd/csrc/sp_amle.h"

extern "C" void orc_gs_sweeps(float *u, const float *g, const float *f, int w,
		int h, float tstep, int niter);
extern "C" void orc_amle_sweeps(float *x, const unsigned char *hole, int w, int h,
		int nn, int nsweep);

static unsigned rng_state;
static double urand()
{
	rng_state = rng_state * 1664525u + 1013904223u;
	return (rng_state >> 8) * (1.0 / 16777216.0);
}

template <class S, class Step>
static int run_tiles(float *u, const float *c, int W, int H, int T, float t,
		Step step)
{
	using G = Geo<S>;
	std::vector<GsTile> tiles;
	std::vector<int> off;
	build_tiles<S>(W, H, T, 1, tiles, off);
	GsInst in{u, c, W, H, T, t, nullptr};
	std::vector<int> owner((size_t)W * H, -1);
	for (size_t w = 0; w + 1 < off.size(); w++) {
		const int n = off[w + 1] - off[w];
		std::vector<float *> lds(n);
		std::vector<TileCtx<S>> ctx(n);
		for (int k = 0; k < n; k++) {
			lds[k] = new float[G::LDS_FLOATS];
			for (int e = 0; e < G::LDS_FLOATS; e++) lds[k][e] = -12345.0f;
			ctx[k] = tile_ctx<S>(tiles[off[w] + k], in, lds[k]);
		}
		for (int k = 0; k < n; k++)
			for (int tid = 0; tid < G::NT; tid++) tile_load<S>(ctx[k], in, tid);
		for (int k = 0; k < n; k++)
			for (int sg = 0; sg < G::NSTEP; sg++)
				for (int tid = 0; tid < G::NT; tid++) step(ctx[k], tid, sg);
		// write-write conflicts between tiles of one wavefront
		for (int k = 0; k < n; k++) {
			const TileCtx<S> &cc = ctx[k];
			for (int e = 0; e < G::LW * G::LH; e++) {
				int row = e / G::LW, col = e - row * G::LW;
				int j = row + cc.jlo, i = col + cc.K - G::SH * j;
				if (i < 0 || i >= W || j < 0 || j >= H || sp_known(cc.C[e])) continue;
				if (!tile_owns<S>(cc, i, j)) continue;
				size_t q = (size_t)j * W + i;
				if (owner[q] == (int)w * 1000000 + k) continue;
				if (owner[q] >= (int)w * 1000000) {
					printf("CONFLICT wavefront %zu pixel %d,%d\n", w, i, j);
					return 1;
				}
				owner[q] = (int)w * 1000000 + k;
			}
		}
		for (int k = 0; k < n; k++)
			for (int tid = 0; tid < G::NT; tid++) tile_store<S>(ctx[k], in, tid);
		for (int k = 0; k < n; k++) delete[] lds[k];
	}
	return 0;
}

template <int R>
static int run_small(float *u, const float *c, int W, int H, int T, float t)
{
	const int cells = small_cells(R, W, H);
	float *lds = new float[2 * cells];
	for (int e = 0; e < 2 * cells; e++) lds[e] = -12345.0f;
	float *U = lds, *C = lds + cells;
	GsInst in{u, c, W, H, T, t, nullptr};
	const long slots = small_slot_count(R, W, H);
	int nt = (int)std::min<long>(1024, std::max<long>(64, (slots + 63) / 64 * 64));
	if (slots > (long)nt * SMALL_MAXSLOT) {
		printf("TOO MANY SLOTS %ld\n", slots);
		return 1;
	}
	std::vector<SmallSlots> sl(nt);
	for (int tid = 0; tid < nt; tid++) {
		small_load<R>(U, C, in, tid, nt);
		small_slots<R>(W, H, tid, nt, sl[tid]);
	}
	const long ns = small_nsteps(R, W, H, T);
	for (long s = 0; s < ns; s++)
		for (int tid = 0; tid < nt; tid++)
			small_step<R>(U, C, W, H, T, t, (int)s, sl[tid]);
	for (int tid = 0; tid < nt; tid++) small_store<R>(U, C, in, tid, nt);
	delete[] lds;
	return 0;
}

// The dependence-ordered launch (k_gs_flow): tiles start in a random order
// in which every existing predecessor (I'-a, J'-b, T'-c), a, b, c in
// {0, 1}, has stored before a tile loads; up to 64 tiles are in flight
// between their load and their store, finished in random order.
template <class S, class Step>
static int run_flow(float *u, const float *c, int W, int H, int T, float t, unsigned seed,
		Step step)
{
	using G = Geo<S>;
	std::vector<GsTile> tiles;
	std::vector<int> off;
	build_tiles<S>(W, H, T, 1, tiles, off);
	int nI, nJ, nT;
	tile_grid<S>(W, H, T, nI, nJ, nT);
	auto cell = [&](int ti, int tj, int tt) { return ((size_t)tt * nJ + tj) * nI + ti; };
	// grid state: -1 no tile, 0 waiting, 1 in flight, 2 stored
	std::vector<signed char> st((size_t)nI * nJ * nT, -1);
	for (const GsTile &x : tiles) st[cell(x.ti, x.tj, x.tt)] = 0;
	auto ready = [&](const GsTile &x) {
		for (int q = 1; q < 8; q++) {
			const int a = q & 1, b = q >> 1 & 1, cc = q >> 2;
			if (x.ti < a || x.tj < b || x.tt < cc) continue;
			const signed char v = st[cell(x.ti - a, x.tj - b, x.tt - cc)];
			if (v == 0 || v == 1) return false;
		}
		return true;
	};
	GsInst in{u, c, W, H, T, t, nullptr};
	struct Fly { TileCtx<S> ctx; std::vector<float> lds; GsTile tl; };
	std::vector<Fly *> fly;
	std::vector<char> started(tiles.size(), 0);
	size_t done = 0;
	unsigned r = seed * 2654435761u + 3;
	auto rnd = [&](unsigned n) { r = r * 1664525u + 1013904223u; return (r >> 8) % n; };
	while (done < tiles.size()) {
		std::vector<size_t> rd;
		for (size_t k = 0; k < tiles.size(); k++)
			if (!started[k] && ready(tiles[k])) rd.push_back(k);
		const bool start = !rd.empty() && (fly.empty() || (fly.size() < 64 && rnd(2)));
		if (start) {
			const size_t k = rd[rnd((unsigned)rd.size())];
			Fly *f = new Fly;
			f->lds.assign(G::LDS_FLOATS, -12345.0f);
			f->tl = tiles[k];
			f->ctx = tile_ctx<S>(f->tl, in, f->lds.data());
			for (int tid = 0; tid < G::NT; tid++) tile_load<S>(f->ctx, in, tid);
			started[k] = 1;
			st[cell(f->tl.ti, f->tl.tj, f->tl.tt)] = 1;
			fly.push_back(f);
			continue;
		}
		if (fly.empty()) {
			printf("STUCK with %zu of %zu tiles stored\n", done, tiles.size());
			return 1;
		}
		const size_t z = rnd((unsigned)fly.size());
		Fly *f = fly[z];
		fly.erase(fly.begin() + z);
		for (int sg = 0; sg < G::NSTEP; sg++)
			for (int tid = 0; tid < G::NT; tid++) step(f->ctx, tid, sg);
		for (int tid = 0; tid < G::NT; tid++) tile_store<S>(f->ctx, in, tid);
		st[cell(f->tl.ti, f->tl.tj, f->tl.tt)] = 2;
		delete f;
		done++;
	}
	return 0;
}

// the whole-level kernel by unknown pixels (R = 1): phases thread by
// thread, the exclusive scan of the packed counts done here as the device
// does it with a work-group scan
template <int MAXS>
static int run_cells(float *u, const float *c, int W, int H, int T, float t)
{
	if ((long)W * H > GSC_MAXCELLS) {
		printf("TOO MANY CELLS %ld\n", (long)W * H);
		return 1;
	}
	const int nt = gsc_threads(W, H);
	if (gsc_slots(W, H, nt) > MAXS) {
		printf("TOO MANY SLOTS %d\n", gsc_slots(W, H, nt));
		return 1;
	}
	const long words = gsc_lds_bytes(W, H) / 4;
	std::vector<float> lds(words, -12345.0f);
	float *U = lds.data();
	int *L = (int *)(U + (W + 2) * (H + 2));
	for (long e = (W + 2) * (H + 2); e < words; e++) L[e - (W + 2) * (H + 2)] = -1;
	GsInst in{u, c, W, H, T, t, nullptr};
	for (int tid = 0; tid < nt; tid++) gsc_load(U, in, tid, nt);
	std::vector<int> cnt(nt), pre(nt);
	int tot = 0;
	for (int tid = 0; tid < nt; tid++) {
		cnt[tid] = gsc_count(in, tid, nt);
		pre[tid] = tot;
		tot += cnt[tid];
	}
	for (int tid = 0; tid < nt; tid++) gsc_fill(L, in, tid, nt, pre[tid], tot);
	std::vector<GscThread<MAXS>> th(nt);
	for (int tid = 0; tid < nt; tid++) gsc_take<MAXS>(th[tid], U, L, in, tid, nt, tot);
	// the device's loop, left at the first step after which the values are
	// final (a sweep that changed no pixel)
	const long ns = small_nsteps(1, W, H, T);
	long s = 0;
	for (int last = -1; s < ns; s++) {
		bool ch = false;
		for (int tid = 0; tid < nt; tid++) ch |= gsc_step<MAXS>(th[tid], U, W, T, t, (int)s);
		if (ch) last = (int)s;
		if (gsc_final(last, W, H, (int)s)) break;
	}
	fprintf(stderr, "CELLS LEFT AT STEP %ld OF %ld\n", s < ns ? s + 1 : ns, ns);
	for (int tid = 0; tid < nt; tid++) gsc_store<MAXS>(th[tid], in);
	return 0;
}

// AMLE: random values with holes, a random start at the holes (as the
// bounds or an initial image would give), nn neighbours; seeds from 100
// put +-Inf among the values, whose NaN offspring the scans must order as
// the reference does
static int main_amle(int argc, char **argv)
{
	if (argc < 8 || argc > 10) {
		fprintf(stderr, "usage: gs_emulate amle W H T nn seed holefrac [variant [flow]]\n");
		return 2;
	}
	const bool flow = argc == 10 && !strcmp(argv[9], "flow");
	const int W = atoi(argv[2]), H = atoi(argv[3]), T = atoi(argv[4]);
	const int nn = atoi(argv[5]);
	rng_state = (unsigned)atoi(argv[6]) * 2654435761u + 11;
	const double frac = atof(argv[7]);
	const int var = argc >= 9 ? atoi(argv[8]) : 0;
	const size_t n = (size_t)W * H;
	std::vector<float> u(n), uref(n), c(n);
	std::vector<unsigned char> hole(n);
	int holes = 0;
	const bool infs = atoi(argv[6]) >= 100;   // seeds from 100: some +-Inf values
	for (size_t q = 0; q < n; q++) {
		u[q] = (float)(urand() * 100 - 50);
		if (urand() < 0.2) u[q] = (float)(int)u[q];     // ties
		if (infs && urand() < 0.01) u[q] = urand() < 0.5 ? INFINITY : -INFINITY;
		hole[q] = urand() < frac;
		holes += hole[q];
		c[q] = hole[q] ? 0.0f : sp_float(SP_KNOWN_BITS);
		uref[q] = u[q];
	}
	orc_amle_sweeps(uref.data(), hole.data(), W, H, nn, T);
	int rc = 0;
	const bool fast4 = nn == 4 && (W > 1 || H > 1);   // the device's dispatch
	const unsigned fseed = (unsigned)atoi(argv[6]);
	if (!amle_with_shape(amle_radius(nn), var, [&](auto sh) {
			using S = decltype(sh);
			auto go = [&](auto step) {
				rc = flow ? run_flow<S>(u.data(), c.data(), W, H, T, 0.0f, fseed, step)
					: run_tiles<S>(u.data(), c.data(), W, H, T, 0.0f, step);
			};
			if constexpr (S::R == 1) {
				if (fast4) {
					// per-thread state of every tile, made at step 0
					std::vector<std::vector<Amle4Thread>> st;
					go([&st](const TileCtx<S> &cc, int tid, int sg) {
						if (sg == 0 && tid == 0) st.emplace_back(Geo<S>::NT);
						auto &v = st.back();
						if (sg == 0) v[tid] = amle4_thread<S>(cc, tid);
						amle4_step<S>(cc, v[tid], sg); });
					return;
				}
			}
			go([nn](const TileCtx<S> &cc, int tid, int sg) {
				amle_tile_step<S>(cc, tid, sg, nn); }); })) {
		printf("NO SHAPE %d for nn %d\n", var, nn);
		return 2;
	}
	if (rc) return 1;
	for (size_t q = 0; q < n; q++)
		if (memcmp(&u[q], &uref[q], 4)) {
			printf("MISMATCH at %zu,%zu: %.9g vs %.9g\n", q % W, q / W, u[q], uref[q]);
			return 1;
		}
	printf("OK %d\n", holes);
	return 0;
}

int main(int argc, char **argv)
{
	if (argc > 1 && !strcmp(argv[1], "amle")) return main_amle(argc, argv);
	if (argc != 9 && argc != 10) {
		fprintf(stderr, "usage: gs_emulate W H T tstep seed holefrac withf "
				"tiles|flow|small|cells [variant]\n");
		return 2;
	}
	const int var = argc == 10 ? atoi(argv[9]) : 0;
	const int W = atoi(argv[1]), H = atoi(argv[2]), T = atoi(argv[3]);
	const float tstep = (float)atof(argv[4]);
	rng_state = (unsigned)atoi(argv[5]) * 2654435761u + 7;
	const double frac = atof(argv[6]);
	const bool withf = atoi(argv[7]) != 0;
	const bool small = !strcmp(argv[8], "small"), cells = !strcmp(argv[8], "cells");
	const bool flow = !strcmp(argv[8], "flow");
	const size_t n = (size_t)W * H;
	float *g = new float[n], *f = withf ? new float[n] : nullptr;
	float *u = new float[n], *uref = new float[n], *c = new float[n];
	int holes = 0;
	// seeds from 100: a few known pixels of +-Inf (neighbours the update
	// skips) or +-3e37 (no overflow at t <= 0.25)
	const bool wild = atoi(argv[5]) >= 100;
	for (size_t q = 0; q < n; q++) {
		g[q] = (float)(urand() * 100);
		if (wild && urand() < 0.0002)
			g[q] = (urand() < 0.5 ? INFINITY : 3e37f) * (urand() < 0.5 ? 1 : -1);
		if (urand() < frac) { g[q] = NAN; holes++; }
		if (f) f[q] = (float)(urand() * 4 - 2);
		u[q] = !isnan(g[q]) ? g[q] : (float)(urand() * 100);
		uref[q] = u[q];
	}
	for (size_t q = 0; q < n; q++) c[q] = sp_coef(g[q], f ? f[q] : 0.0f);
	orc_gs_sweeps(uref, g, f, W, H, tstep, T);
	float teff = tstep;
	int R = 1;
	if (tstep < 0) {
		R = 2;
		if (tstep < -1000) { R = 3; teff = tstep; teff += 1000; teff *= -1; }
	}
	int rc;
	if (cells) {
		if (R != 1) {
			printf("CELLS NEED R = 1\n");
			return 2;
		}
		rc = run_cells<8>(u, c, W, H, T, teff);
	} else if (small) rc = R == 1 ? run_small<1>(u, c, W, H, T, teff)
		: R == 2 ? run_small<2>(u, c, W, H, T, teff) : run_small<3>(u, c, W, H, T, teff);
	else if (!sp_with_shape(R, var, [&](auto sh) {
			using S = decltype(sh);
			const unsigned fseed = (unsigned)atoi(argv[5]);
			auto go = [&](auto step) {
				rc = flow ? run_flow<S>(u, c, W, H, T, teff, fseed, step)
					: run_tiles<S>(u, c, W, H, T, teff, step);
			};
			if constexpr (S::R == 1) {
				// the device's R = 1 loop: per-thread state made at step 0
				// (a tile's steps run together, tile after tile)
				std::vector<std::vector<Gs1Thread>> st;
				go([&st](const TileCtx<S> &cc, int tid, int sg) {
					if (sg == 0 && tid == 0) st.emplace_back(Geo<S>::NT);
					auto &v = st.back();
					if (sg == 0) v[tid] = gs1_thread<S>(cc, tid);
					gs1_step<S>(cc, v[tid], sg); });
				return;
			}
			go([](const TileCtx<S> &cc, int tid, int sg) { tile_step<S>(cc, tid, sg); }); })) {
		printf("NO SHAPE %d for radius %d\n", var, R);
		return 2;
	}
	if (rc) return 1;
	for (size_t q = 0; q < n; q++)
		if (memcmp(&u[q], &uref[q], 4)) {
			printf("MISMATCH at %zu,%zu: %.9g vs %.9g\n", q % W, q / W, u[q], uref[q]);
			return 1;
		}
	printf("OK %d\n", holes);
	delete[] g; delete[] f; delete[] u; delete[] uref; delete[] c;
	return 0;
}
