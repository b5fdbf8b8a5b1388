// Host-side construction of the Gauss-Seidel tile schedule (sp_gs.h):
// the skewed tiles of one level, grouped by wavefront I'+J'+T'.  Shared by
// the host driver and the CPU emulator of the tests.
#pragma once
#include <algorithm>
#include <vector>
#include "sp_gs.h"

// extents of a level's skewed tile grid (tiles outside the image included)
template <class S>
inline void tile_grid(int W, int H, int T, int &nI, int &nJ, int &nT)
{
	constexpr int R = S::R;
	const int SH = R - 1;
	long ispan, jspan;
	if (R == 1) { ispan = (long)W + T - 1; jspan = (long)H + T - 1; }
	else {
		ispan = (long)(W - 1) + (long)SH * (H - 1) + (long)R * SH * (T - 1) + 1;
		jspan = (long)(H - 1) + (long)R * (T - 1) + 1;
	}
	nI = (int)((ispan + S::BI - 1) / S::BI);
	nJ = (int)((jspan + S::BJ - 1) / S::BJ);
	nT = (T + S::BT - 1) / S::BT;
}

// skewed tiles of one level, grouped by wavefront I'+J'+T'
template <class S>
inline void build_tiles(int W, int H, int T, int pd, std::vector<GsTile> &out,
		std::vector<int> &off)
{
	constexpr int R = S::R;
	const int SH = R - 1;
	int nI, nJ, nT;
	tile_grid<S>(W, H, T, nI, nJ, nT);
	std::vector<std::vector<GsTile>> wf(nI + nJ + nT);
	for (int tt = 0; tt < nT; tt++) {
		const long k0 = (long)tt * S::BT;
		const long bte = std::min<long>(S::BT, T - k0);
		for (int tj = 0; tj < nJ; tj++) {
			long jl, jh;
			if (R == 1) { jl = (long)tj * S::BJ - k0 - (bte - 1); jh = (long)tj * S::BJ + S::BJ - 1 - k0; }
			else { jl = (long)tj * S::BJ - R * (k0 + bte - 1); jh = (long)tj * S::BJ + S::BJ - 1 - R * k0; }
			if (jh < 0 || jl >= H) continue;
			for (int ti = 0; ti < nI; ti++) {
				long il, ih;
				if (R == 1) { il = (long)ti * S::BI - k0 - (bte - 1); ih = (long)ti * S::BI + S::BI - 1 - k0; }
				else {
					il = (long)ti * S::BI - (long)SH * ((long)tj * S::BJ + S::BJ - 1);
					ih = (long)ti * S::BI + S::BI - 1 - (long)SH * tj * S::BJ;
				}
				if (ih < 0 || il >= W) continue;
				for (int p = 0; p < pd; p++)
					wf[ti + tj + tt].push_back({p, ti, tj, tt});
			}
		}
	}
	out.clear();
	off.assign(1, 0);
	for (auto &v : wf) {
		if (v.empty()) continue;
		out.insert(out.end(), v.begin(), v.end());
		off.push_back((int)out.size());
	}
}


// tile_grid for variant v of radius R
inline void tile_grid_var(int R, int v, int W, int H, int T, int &nI, int &nJ, int &nT)
{
	nI = nJ = nT = 0;
	sp_with_shape(R, v, [&](auto sh) { tile_grid<decltype(sh)>(W, H, T, nI, nJ, nT); });
}

// the same for variant v of radius R; returns the variant's sweeps per tile
inline int build_tiles_var(int R, int v, int W, int H, int T, int pd,
		std::vector<GsTile> &out, std::vector<int> &off)
{
	int bt = 0;
	sp_with_shape(R, v, [&](auto sh) {
		using S = decltype(sh);
		build_tiles<S>(W, H, T, pd, out, off);
		bt = S::BT;
	});
	return bt;
}
