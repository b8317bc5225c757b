// rc_kernels.hip -- CDNA4 (gfx950) kernels of the RheoCell per-step hot path.
//
// k_box_gather   WetPolarModel::computeGlobalSums box part
//                (src/Interactions/WetPolarModel.cpp:236-260) as an
//                atomic-free gather: one thread per box site, one workgroup
//                per 32x8 box tile, covering patches visited in cell-index
//                order so the sum order equals the reference scatter order.
// k_stage<MODE>  one workgroup per cell, patch staged in LDS:
//                BaseInteraction::derivX/derivY/Laplacian
//                (src/Interactions/BaseInteraction.h:196-263),
//                WetPolarModel::f_interaction / calc_internal_forces
//                (WetPolarModel.cpp:262-339), the dry velocity v = f/xi
//                (WetPolarModel.cpp:194-198), FD_CPUBackend::first_step
//                (src/Backends/FD_CPUBackend.cpp:44-120) including the
//                per-cell reductions, MultiPhaseField::set_positions
//                recentering (src/Fields/MultiPhaseField.cpp:223-247) and
//                WetPolarModel::updateDirectedActiveForces (:342-414).
// k_cg_*         matrix-free Jacobi-PCG for the wet velocity solve
//                (WetPolarModel.cpp:200-233); restates Eigen's
//                ConjugateGradient iteration with both components at once.
//
// All arithmetic is FP64 and is compiled with -ffp-contract=off so that
// every per-site expression is the IEEE evaluation of the reference source.
#include "rc_kernels.h"
#include "rc_conv.h"

#include <hip/hip_ext.h>

#include <float.h>

namespace rc {

// Optional per-launch timing: when set, the next launches record these events
// at kernel start/end (hipExtLaunchKernel), i.e. pure kernel durations.
static thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;
void set_launch_events(hipEvent_t start, hipEvent_t stop) {
	t_ev_start = start;
	t_ev_stop = stop;
}
#define RC_LAUNCH(K, G, B, SH, S, ...)                                                         \
	do {                                                                                       \
		if(t_ev_start) hipExtLaunchKernelGGL(K, G, B, SH, S, t_ev_start, t_ev_stop, 0, __VA_ARGS__); \
		else hipLaunchKernelGGL(K, G, B, SH, S, __VA_ARGS__);                                  \
	} while(0)

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline int box_x(int side, int site, int distX) {  // SquareBox.cpp:31-37
	int x = (site - ((site / side) * side)) + distX;
	while(x < 0) x += side;
	while(x >= side) x -= side;
	return x;
}
__device__ inline int box_y(int side, int site, int distY) {  // SquareBox.cpp:39-45
	int y = (site / side) + distY;
	while(y < 0) y += side;
	while(y >= side) y -= side;
	return y;
}
__device__ inline int box_elem(int side, int site, int dX, int dY) {
	return box_x(side, site, dX) + box_y(side, site, dY) * side;
}

// Wave reductions on the DPP crossbar (no LDS traffic): within each 16-lane
// row, quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and row_mirror leave
// the row total in every lane; the four row totals are then read with
// v_readlane.  Every lane of the wave must be active.
template <int CTRL>
__device__ inline int dpp_i32(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true); }
template <int CTRL>
__device__ inline double dpp_f64(double v) {
	const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
	const unsigned lo = (unsigned) dpp_i32<CTRL>((int) (unsigned) u);
	const unsigned hi = (unsigned) dpp_i32<CTRL>((int) (unsigned) (u >> 32));
	return __builtin_bit_cast(double, ((unsigned long long) hi << 32) | lo);
}
__device__ inline double readlane_f64(double v, int lane) {
	const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
	const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) u, lane);
	const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (u >> 32), lane);
	return __builtin_bit_cast(double, ((unsigned long long) hi << 32) | lo);
}
constexpr int DPP_QUAD_1032 = 0xB1, DPP_QUAD_2301 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;
constexpr int DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;   // rotate right within a row of 16 lanes
// sum over the 64 lanes, the same value in every lane
__device__ inline double wave_sum(double v) {
	v += dpp_f64<DPP_QUAD_1032>(v);
	v += dpp_f64<DPP_QUAD_2301>(v);
	v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
	v += dpp_f64<DPP_ROW_MIRROR>(v);
	return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ inline int wave_min(int v) {
	v = min(v, dpp_i32<DPP_QUAD_1032>(v));
	v = min(v, dpp_i32<DPP_QUAD_2301>(v));
	v = min(v, dpp_i32<DPP_ROW_HALF_MIRROR>(v));
	v = min(v, dpp_i32<DPP_ROW_MIRROR>(v));
	return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
			min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Sum NV values over a block of NW waves (256 threads by default); the result
// is valid in thread 0.  s_red holds NW * NV doubles.
template <int NV, int NW = 4>
__device__ inline void block_sum(double (&v)[NV], double *s_red) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for(int j = 0; j < NV; ++j) v[j] = wave_sum(v[j]);
	if(lane == 0) {
#pragma unroll
		for(int j = 0; j < NV; ++j) s_red[wave * NV + j] = v[j];
	}
	__syncthreads();
	if(threadIdx.x == 0) {
#pragma unroll
		for(int j = 0; j < NV; ++j) {
			double t = s_red[j];
#pragma unroll
			for(int w = 1; w < NW; ++w) t += s_red[w * NV + j];
			v[j] = t;
		}
	}
	__syncthreads();
}

// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8).  The
// item (cell, tile, block of box sites) of workgroup b is chosen so that each
// XCD takes one contiguous range of the n items: neighbouring cells and tiles
// share box data, which then stays in one XCD's L2.
constexpr int NXCD = 8;
__device__ inline int xcd_item(int b, int n) {
	const int q = n / NXCD, r = n % NXCD, x = b % NXCD, j = b / NXCD;
	return x * q + min(x, r) + j;
}

__device__ inline bool circ_overlap(int a0, int alen, int b0, int blen, int L) {
	// do [a0, a0+alen) and [b0, b0+blen) intersect on the circle of length L?
	int d1 = b0 - a0;
	if(d1 < 0) d1 += L;
	int d2 = a0 - b0;
	if(d2 < 0) d2 += L;
	return d1 < alen || d2 < blen;
}

// ---------------------------------------------------------------------------
// Tile lists: cells whose patch intersects a 32x8 box tile, in cell order.
// Built from the compact per-cell geometry (geo) by one pass over the
// box's cells; the corrector's gather reuses the predictor's lists (the
// patches only move at the corrector's recentering).
// ---------------------------------------------------------------------------

// A tile's list entry, relative to the tile's origin (x0, y0):
// {lx0 | ly0 << 16, LX | LY << 16, offx | offy << 16, row start} with
// lx0 = (x0 - cx) mod L, ly0 = (y0 - cy) mod L, so that box site
// (x0 + sx, y0 + sy) lies at column (lx0 + sx) mod L, row (ly0 + sy) mod L of
// the cell's patch.  Lists are padded to a multiple of 16 entries with cells
// that cover nothing (LX = LY = 0): the site loops read whole batches.
__device__ inline int4 tile_entry(const int4 g, int L, int x0, int y0) {
	int lx0 = x0 - g.x % L;
	if(lx0 < 0) lx0 += L;
	int ly0 = y0 - g.x / L;
	if(ly0 < 0) ly0 += L;
	return make_int4(lx0 | (ly0 << 16), g.y, g.z, g.w);
}

// row of the patch site covering tile site (sx, sy), or -1 (sx < 32 <= L or sx < L)
__device__ inline int entry_row(const int4 e, int sx, int sy, int L) {
	int lx = (e.x & 0xFFFF) + sx;
	if(lx >= L) lx -= L;
	int ly = (e.x >> 16) + sy;
	if(ly >= L) ly -= L;
	const int LXj = e.y & 0xFFFF, LYj = e.y >> 16;
	int qx = lx - (e.z & 0xFFFF);
	if(qx < 0) qx += LXj;
	int qy = ly - (e.z >> 16);
	if(qy < 0) qy += LYj;
	// branch-free, so that a batch's entries are all read before the first is used
	return lx < LXj && ly < LYj ? e.w + qy * LXj + qx : -1;
}

__device__ inline void pad_list(int n, int4 *s_list) {
	const int np = (n + 15) & ~15;
	for(int j = n + (int) threadIdx.x; j < np; j += blockDim.x) s_list[j] = make_int4(0, 0, 0, 0);
}

// A prebuilt list is read through the scalar cache: the count and the
// entries are the same for every lane, so they arrive in SGPRs and no LDS
// staging or barrier is needed.  (The lists are written by an earlier kernel;
// the padding to a multiple of 16 is written with them.)
using cint_p = const int __attribute__((address_space(4))) *;
__device__ inline int list_count(const int *tile_count, int tile) { return ((cint_p) tile_count)[tile]; }
__device__ inline int4 list_entry(const int4 *tile_list, int tile, int j) {
	const cint_p e = (cint_p) (tile_list + tile * MAXT + j);
	return make_int4(e[0], e[1], e[2], e[3]);
}

// Builds the tile's list from the compact cell geometry (cell order) into LDS
// and the global list; returns its length.  Thread t tests cells 4t .. 4t+3
// of each chunk of 1024 (their geometry requested at once); a ballot prefix
// over the wave and the wave totals place the hits in cell order.
__device__ int tile_list_build(const GatherArgs &a, int tile, const DevBox &B, int x0, int y0, int tw, int th,
		int4 *s_list, int *s_wcnt, int *s_cells = nullptr) {
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, L = B.L;
	const unsigned long long lt = (1ull << lane) - 1ull;
	int n = 0;
	for(int base = B.cell_begin; base < B.cell_end; base += 1024) {
		int4 g[4];
		bool hit[4];
#pragma unroll
		for(int u = 0; u < 4; ++u) {
			const int c = base + 4 * tid + u;
			g[u] = c < B.cell_end ? a.geo[c] : make_int4(0, 0, 0, 0);
		}
		int pre = 0, tot = 0;
#pragma unroll
		for(int u = 0; u < 4; ++u) {
			const int c = base + 4 * tid + u;
			hit[u] = c < B.cell_end && circ_overlap(g[u].x % L, g[u].y & 0xFFFF, x0, tw, L) &&
					circ_overlap(g[u].x / L, g[u].y >> 16, y0, th, L);
			const unsigned long long m = __ballot(hit[u]);
			pre += __popcll(m & lt);
			tot += __popcll(m);
		}
		if(lane == 0) s_wcnt[wave] = tot;
		__syncthreads();
		int pos = n + pre, all = 0;
		for(int w = 0; w < 4; ++w) {
			if(w < wave) pos += s_wcnt[w];
			all += s_wcnt[w];
		}
#pragma unroll
		for(int u = 0; u < 4; ++u)
			if(hit[u]) {
				if(pos < MAXT) {
					const int4 e = tile_entry(g[u], L, x0, y0);
					s_list[pos] = e;
					if(s_cells) s_cells[pos] = base + 4 * tid + u;
					a.tile_cells[tile * MAXT + pos] = base + 4 * tid + u;
					a.tile_geo[tile * MAXT + pos] = e;
				}
				++pos;
			}
		n += all;
		__syncthreads();   // s_wcnt is rewritten by the next chunk
	}
	if(tid == 0) a.tile_count[tile] = min(n, MAXT);
	pad_list(min(n, MAXT), s_list);
	for(int j = min(n, MAXT) + tid; j < ((min(n, MAXT) + 15) & ~15); j += blockDim.x)
		a.tile_geo[tile * MAXT + j] = make_int4(0, 0, 0, 0);
	__syncthreads();
	return n;
}

constexpr int GCHUNK = 8;   // covering patches examined per batch of independent loads (CG gathers)
constexpr int GCHUNK_PHI = 24;   // the same for the sum phi^2 gather (one batch at typical overlaps; entries past the count are masked)
constexpr int GCHUNK_LIST = 32;  // list entries a reusing gather requests before it knows the count

// ---------------------------------------------------------------------------
// k_box_gather: phi2[k] = sum over covering patches (cell order) of phi^2,
// cover[k] = number of covering patches (the reference's overlap list size)
// ---------------------------------------------------------------------------
// The tile's box: box 0 from the launch arguments (ONE), else the tables.
template <bool ONE>
__device__ inline DevBox gather_box(const GatherArgs &a, int tile) {
	if(ONE) {
		DevBox B{};
		B.L = a.one_L;
		B.ntx = a.one_ntx;
		B.cell_end = a.one_cells;
		return B;
	}
	return a.boxes[a.tile_box[tile]];
}
// NEM: the nematic model's box sums of phi Q00 and phi Q01 in the same pass
// (WetModel::computeGlobalSums, WetModel.cpp:423-431: phi Q added in cell
// order, as k_box_field's which = 2), from each listed cell's Q staged in LDS.
template <bool ONE, bool NEM>
__global__ __launch_bounds__(256) void k_box_gather(GatherArgs a) {
	__shared__ int4 s_list[MAXT];
	__shared__ int s_wcnt[4];
	__shared__ int s_ci[NEM ? MAXT : 1];
	__shared__ double2 s_q[NEM ? MAXT : 1];
	const int tile = xcd_item(blockIdx.x, gridDim.x), tid = threadIdx.x;
	const DevBox B = gather_box<ONE>(a, tile);
	const int t = tile - B.tile_off, L = B.L;
	const int x0 = (t % B.ntx) * TW, y0 = (t / B.ntx) * TH;
	const int tw = min(TW, L - x0), th = min(TH, L - y0);
	const int sx = tid % TW, sy = tid / TW;
	double acc = 0.0, a00 = 0.0, a01 = 0.0;
	int cnt = 0;
	auto add = [&](const int (&row)[GCHUNK_PHI], int base) {
		double v[GCHUNK_PHI];
#pragma unroll
		for(int k = 0; k < GCHUNK_PHI; ++k) v[k] = row[k] >= 0 ? a.phi[row[k]] : 0.;
#pragma unroll
		for(int k = 0; k < GCHUNK_PHI; ++k)
			if(row[k] >= 0) {
				acc += v[k] * v[k];
				++cnt;
				if(NEM) {
					const double2 q = s_q[base + k];
					a00 += q.x * v[k];
					a01 += q.y * v[k];
				}
			}
	};
	int n;
	if(a.build_lists) {
		if(a.st->flags) return;
		n = tile_list_build(a, tile, B, x0, y0, tw, th, s_list, s_wcnt, NEM ? s_ci : nullptr);
	}
	else {
		// the status word, the count and the list's first GCHUNK_LIST entries
		// requested at once (one memory round trip; a list is padded with
		// empty entries to a multiple of 16, and only entries below the padded
		// count are read)
		const unsigned flags = a.st->flags;
		n = list_count(a.tile_count, tile);
		const int4 *src = a.tile_geo + (size_t) tile * MAXT;
		const int *srcc = a.tile_cells + (size_t) tile * MAXT;
		const int4 e = src[min(tid, GCHUNK_LIST - 1)];
		const int ec = NEM ? srcc[min(tid, GCHUNK_LIST - 1)] : 0;
		// all in hand here: the compiler issues the loads together
		asm volatile("" :: "v"(e.x), "v"(e.y), "v"(e.z), "v"(e.w), "v"(ec), "s"(flags), "s"(n));
		if(tid < GCHUNK_LIST) {
			s_list[tid] = e;
			if(NEM) s_ci[tid] = ec;
		}
		if(flags) return;
		const int np = min((n + 15) & ~15, MAXT);
		for(int j = GCHUNK_LIST + tid; j < np; j += blockDim.x) {
			s_list[j] = src[j];
			if(NEM) s_ci[j] = srcc[j];
		}
		__syncthreads();
	}
	if(n > MAXT) {
		if(tid == 0) atomicOr(&a.st->flags, ST_TILE);
		return;
	}
	if(NEM) {
		for(int j = tid; j < n; j += blockDim.x) {
			const DevCell &C = a.cells[s_ci[j]];
			s_q[j] = make_double2(C.Q00, C.Q01);
		}
		__syncthreads();
	}
	if(sx >= tw || sy >= th) return;
	for(int base = 0; base < n; base += GCHUNK_PHI) {
		int row[GCHUNK_PHI];
#pragma unroll
		for(int k = 0; k < GCHUNK_PHI; ++k) {
			const int j = base + k;
			row[k] = j < n ? entry_row(s_list[min(j, MAXT - 1)], sx, sy, L) : -1;
		}
		add(row, base);
	}
	const int k = B.box_off + (y0 + sy) * L + x0 + sx;
	a.phi2[k] = acc;
	a.cover[k] = cnt;
	if(NEM) {
		a.sq00[k] = a00;
		a.sq01[k] = a01;
	}
	if(cnt > STORE_MAX) {
		if(atomicOr(&a.st->flags, ST_OVERLAP) == 0u) a.st->site = k;
	}
}

hipError_t launch_gather(const GatherArgs &a, int ntiles, hipStream_t s) {
	if(a.nboxes == 1) {
		if(a.nem) RC_LAUNCH((k_box_gather<true, true>), dim3(ntiles), dim3(256), 0, s, a);
		else RC_LAUNCH((k_box_gather<true, false>), dim3(ntiles), dim3(256), 0, s, a);
	}
	else {
		if(a.nem) RC_LAUNCH((k_box_gather<false, true>), dim3(ntiles), dim3(256), 0, s, a);
		else RC_LAUNCH((k_box_gather<false, false>), dim3(ntiles), dim3(256), 0, s, a);
	}
	return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Observable box fields (VelocityField::calc_field, VelocityField.cpp:69-78;
// ShapeField::calc_field, ShapeField.cpp:69-78): f(k) += w * phi over the
// fields in index order, so one sum per site in cell order reproduces the
// reference's additions exactly.  With which = 2 the same gather gives the
// nematic model's box sums sumQ00 / sumQ01 (WetModel::computeGlobalSums,
// WetModel.cpp:423-431: phi Q added in cell order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_box_field(FieldArgs a) {
	const int tile = a.tile0 + xcd_item(blockIdx.x, gridDim.x), tid = threadIdx.x;
	const DevBox B = a.boxes[a.tile_box[tile]];
	const int t = tile - B.tile_off, L = B.L;
	const int x0 = (t % B.ntx) * TW, y0 = (t / B.ntx) * TH;
	const int tw = min(TW, L - x0), th = min(TH, L - y0);
	const int n = min(list_count(a.tile_count, tile), MAXT);
	const int sx = tid % TW, sy = tid / TW;
	if(sx >= tw || sy >= th) return;
	double fx = 0., fy = 0.;
	// GCHUNK covering cells' rows requested at once, then added in cell order
	for(int base = 0; base < n; base += GCHUNK) {
		int row[GCHUNK];
		double2 w[GCHUNK];
		double ph[GCHUNK];
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) row[q] = entry_row(list_entry(a.tile_geo, tile, base + q), sx, sy, L);
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) {
			w[q] = make_double2(0., 0.);
			ph[q] = 0.;
			if(row[q] < 0) continue;
			if(a.which == 0) w[q] = a.v[row[q]];
			else {
				const int ci = a.tile_cells[tile * MAXT + base + q];
				w[q] = a.which == 1 ? make_double2(a.cells[ci].S00, a.cells[ci].S01)
						: make_double2(a.cells[ci].Q00, a.cells[ci].Q01);
			}
			ph[q] = a.phi[row[q]];
		}
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) {
			if(row[q] < 0) continue;
			fx += w[q].x * ph[q];
			fy += w[q].y * ph[q];
		}
	}
	const int k = (a.global ? B.box_off : 0) + (y0 + sy) * L + x0 + sx;
	a.fx[k] = fx;
	a.fy[k] = fy;
}

hipError_t launch_box_field(const FieldArgs &a, int ntiles, hipStream_t s) {
	if(ntiles <= 0) return hipSuccess;
	RC_LAUNCH(k_box_field, dim3(ntiles), dim3(256), 0, s, a);
	return hipGetLastError();
}

// the coarse field of VelocityField::f_field / ShapeField::f_field
// (VelocityField.cpp:106-121): site (j, i) adds f(ss) / (g g) over the g x g
// sites from (j - g/2, i - g/2), rows k outer, columns l inner, periodic.
// Up to COARSE_LDS_G a 16 x 16 block of sites stages its (15 + g)^2 window
// of both fields in LDS; larger windows read straight from the fields.
constexpr int COARSE_B = 16;
constexpr int COARSE_LDS_G = 33;
__global__ __launch_bounds__(256) void k_box_coarse_lds(const double *__restrict__ fx, const double *__restrict__ fy,
		double *__restrict__ cx, double *__restrict__ cy, int L, int g) {
	extern __shared__ double s_w[];
	const int W = COARSE_B + g - 1, h = g / 2;
	const int bx0 = blockIdx.x * COARSE_B, by0 = blockIdx.y * COARSE_B;
	for(int e = threadIdx.x; e < W * W; e += 256) {
		const int wy = e / W, wx = e - wy * W;
		int xx = (bx0 - h + wx) % L, yy = (by0 - h + wy) % L;
		xx += xx < 0 ? L : 0;
		yy += yy < 0 ? L : 0;
		s_w[e] = fx[yy * L + xx];
		s_w[W * W + e] = fy[yy * L + xx];
	}
	__syncthreads();
	const int tx = threadIdx.x % COARSE_B, ty = threadIdx.x / COARSE_B;
	const int j = bx0 + tx, i = by0 + ty;
	if(j >= L || i >= L) return;
	const double gg = (double) (g * g);
	double sx = 0., sy = 0.;
	for(int k = 0; k < g; ++k) {
		const double *rx = s_w + (ty + k) * W + tx, *ry = rx + W * W;
		for(int l = 0; l < g; ++l) {
			sx += rx[l] / gg;
			sy += ry[l] / gg;
		}
	}
	cx[i * L + j] = sx;
	cy[i * L + j] = sy;
}

__global__ __launch_bounds__(256) void k_box_coarse(const double *__restrict__ fx, const double *__restrict__ fy,
		double *__restrict__ cx, double *__restrict__ cy, int L, int g) {
	const int site = blockIdx.x * 256 + threadIdx.x;
	if(site >= L * L) return;
	const int i = site / L, j = site - i * L;
	const double gg = (double) (g * g);
	const int xs = (j - g / 2 + L) % L, ys = (i - g / 2 + L) % L;
	double sx = 0., sy = 0.;
	for(int k = 0; k < g; ++k) {
		const int yy = (ys + k) % L;
		for(int l = 0; l < g; ++l) {
			const int ss = (xs + l) % L + yy * L;
			sx += fx[ss] / gg;
			sy += fy[ss] / gg;
		}
	}
	cx[site] = sx;
	cy[site] = sy;
}

hipError_t launch_box_coarse(const double *fx, const double *fy, double *cx, double *cy, int L, int size_grid,
		hipStream_t s) {
	if(L <= 0 || size_grid < 1 || size_grid > L) return hipErrorInvalidValue;
	if(size_grid <= COARSE_LDS_G) {
		const int W = COARSE_B + size_grid - 1, nb = (L + COARSE_B - 1) / COARSE_B;
		RC_LAUNCH(k_box_coarse_lds, dim3(nb, nb), dim3(256), sizeof(double) * 2 * W * W, s, fx, fy, cx, cy, L,
				size_grid);
	}
	else RC_LAUNCH(k_box_coarse, dim3((L * L + 255) / 256), dim3(256), 0, s, fx, fy, cx, cy, L, size_grid);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_stage<MODE>: force evaluation (+ field update) of one cell
// ---------------------------------------------------------------------------
// Column strips: a cell's LX x LY patch (ring-buffer coordinates, storage
// q = qy * LX + qx) is covered by nb = CT / LX row blocks of R = ceil(LY / nb)
// rows, CT the threads per cell; thread t takes column qx = t % LX of block
// t / LX.  Every thread of a used block evaluates R rows: the last block's rows
// past LY are the periodic images of rows 0, 1, ... (read from their wrapped
// storage rows, evaluated like any other row and then discarded), so no row of
// the unrolled loops carries a per-thread guard.  Threads of unused blocks
// evaluate rows 0 .. R-1 of their column and write nothing.
__device__ __host__ inline int strip_rows(int LX, int LY, int CT) {
	const int nb = CT / LX;
	return nb > 0 ? (LY + nb - 1) / nb : 1 << 20;
}

// x / y; with RCP (y a power of two, rcp its exact reciprocal) the product
// x * rcp, which rounds the same exact quotient
template <bool RCP>
__device__ inline double div_by(double x, double y, double rcp) { return RCP ? x * rcp : x / y; }

// Sum of 8 per-lane values over a workgroup of NW waves.  A transposing
// butterfly inside each row of 16 lanes (xor 1, xor 2, then rotations by 4
// and 8; each exchange hands over the half of the values the partner keeps)
// leaves in lane l the row's sum of value 4*b0 + 2*b1 + b2 (b = bits of l);
// lanes 0-7 of every row store the row sums, and threads 0..7 add the 4*NW
// row sums of their value in a fixed order.  Returns the total of value
// threadIdx.x in threads 0..7.  Every lane must be active.
template <int NW>
__device__ inline double block_sum8(const double (&v)[8], double *s_part) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
	double a4[4], a2[2];
#pragma unroll
	for(int i = 0; i < 4; ++i) {
		const double send = b0 ? v[i] : v[i + 4], keep = b0 ? v[i + 4] : v[i];
		a4[i] = keep + dpp_f64<DPP_QUAD_1032>(send);
	}
#pragma unroll
	for(int i = 0; i < 2; ++i) {
		const double send = b1 ? a4[i] : a4[i + 2], keep = b1 ? a4[i + 2] : a4[i];
		a2[i] = keep + dpp_f64<DPP_QUAD_2301>(send);
	}
	double s;
	{
		const double send = b2 ? a2[0] : a2[1], keep = b2 ? a2[1] : a2[0];
		s = keep + dpp_f64<DPP_ROW_ROR4>(send);
	}
	s += dpp_f64<DPP_ROW_ROR8>(s);
	if((lane & 8) == 0) s_part[(wave * 4 + (lane >> 4)) * 8 + (b0 ? 4 : 0) + (b1 ? 2 : 0) + (b2 ? 1 : 0)] = s;
	__syncthreads();
	double t = 0.;
	if(threadIdx.x < 8) {
		t = s_part[threadIdx.x];
#pragma unroll
		for(int r = 1; r < 4 * NW; ++r) t += s_part[r * 8 + threadIdx.x];
	}
	return t;
}

// BaseInteraction::Laplacian with walls (BaseInteraction.h:218-262): bits of m
// = open E 1, W 2, N 4, S 8; a closed side drops out and the opposite one
// counts twice, in the reference's expressions
__device__ inline double lap_masked(unsigned m, double pN, double pS, double pE, double pW, double ph) {
	const bool E = m & 1, W = m & 2, N = m & 4, S = m & 8;
	if(N && S) {
		if(E && W) return pN + pS + pE + pW - 4 * ph;
		else if(E) return pN + pS + 2 * pE - 4 * ph;
		else if(W) return pN + pS + 2 * pW - 4 * ph;
		return pN + pS - 2 * ph;
	}
	else if(N) {
		if(E && W) return 2 * pN + pE + pW - 4 * ph;
		else if(E) return 2 * pN + 2 * pE - 4 * ph;
		else if(W) return 2 * pN + 2 * pW - 4 * ph;
		return 2 * pN - 2 * ph;
	}
	else if(S) {
		if(E && W) return 2 * pS + pE + pW - 4 * ph;
		else if(E) return 2 * pS + 2 * pE - 4 * ph;
		else if(W) return 2 * pS + 2 * pW - 4 * ph;
		return 2 * pS - 2 * ph;
	}
	return 0.;
}

// BaseField::set_F_ext / set_f_density_ext (BaseField.h:54-79): the external
// forces' free energy and its density at a site of wall field w
// (the loops run over the fixed maximum with a guard: constant indices keep
// the parameter copy in scalar registers instead of private memory)
__device__ inline double f_ext(const DevParams &P, double ph, double w) {
	double F_ext = 0.;
#pragma unroll
	for(int f = 0; f < 4; ++f)
		if(f < P.nforce) F_ext += P.fF[f] * ph * w * w;
	return F_ext;
}
__device__ inline double f_density_ext(const DevParams &P, double ph, double w) {
	double f_density_ext = 0.;
#pragma unroll
	for(int f = 0; f < 4; ++f)
		if(f < P.nforce) f_density_ext += P.fD[f] * ph * ph * w * w;
	return f_density_ext;
}

// BaseField::set_V_ext (BaseField.h:107-115): the x velocity of the last
// force's profile at box row y (ShearFlowChannel.cpp:41-45,
// PoiseuilleFlowChannel.cpp) or of the cell index (Stretching.cpp); the y
// velocity is 0
__device__ inline double v_ext(const DevParams &P, int y, int L, int index) {
	const double lambda = P.vext_lambda, shear_rate = P.vext_rate;
	const int ly = L;
	if(P.vext_kind == 1) {
		if(y > lambda && y < ly - lambda)
			return ((y - (lambda + 1) + 0.5) - 0.5 * (ly - 2 * (lambda + 1))) * shear_rate / (ly - 2 * (lambda + 1));
	}
	else if(P.vext_kind == 2) {
		if(y > lambda && y < ly - lambda)
			return 4 * (y - (lambda + 1)) * shear_rate * (1 - (y - (lambda + 1)) / (ly - 2 * (lambda + 1))) / (ly - 2 * (lambda + 1));
	}
	else if(P.vext_kind == 3) {
		const int yy = index / 10, x = index - yy * 10;
		if(x == 0) return -shear_rate;
		if(x == 10 - 1) return shear_rate;
	}
	return 0.;
}

// CT threads per cell; NPT >= the strip rows of every patch (compile time), so
// the per-site registers are fully unrolled.  RCP: lambda and the friction of
// every parameter set are powers of two (divisions become exact multiplications).
// ST: the predictor (first_step(true)) when set, else the corrector (a.store
// equals ST; MODE_CF takes ST = 0).  WL: walls and external forces (a.wmask,
// a.walls): masked stencils, the forces' free energy in mu, no force on wall
// sites and the forces' advection in the update.  LDS layout: stage_lds_bytes
// (rc_kernels.h).
// The cell records, the workgroup's cell list, the status word and the rows
// and box sums of the first loads lead the arguments (preloaded in SGPRs).
template <int MODE, int NPT, bool RCP, int CT, bool ST, bool WL = false>
__global__ __launch_bounds__(CT, NPT <= 7 ? 4 : 1) void k_stage(const DevCell *cells_in, const int *cell_list,
		const DevStatus *st_in, const double *phi_c_in, const double *phi2_in, StageArgs a) {
	constexpr bool store = ST;
	// (the status word is requested here and tested once the first loads
	// are under way, before anything is written)
	const unsigned st_flags = st_in->flags;
	constexpr int NW = CT / 64;
	extern __shared__ double2 smem2[];
	__shared__ double s_red[NW * 4 * 8];
	__shared__ int s_redi[2 * NW];
	__shared__ double s_bc;
	const int slot = xcd_item(blockIdx.x, gridDim.x);
	const int ci = cell_list ? cell_list[slot] : slot, tid = threadIdx.x;
	const DevCell C = cells_in[ci];
	const DevBox B = a.boxes[C.box];
	const DevParams P = a.params[2 * C.box + a.cf_set];
	const int LX = C.LX, LY = C.LY, L = B.L, PW = LX + 2, WX = LX + 2, WY = LY + 2;
	const int TR = LY + NPT;                 // tile rows: LY storage rows and their periodic images
	double2 *s_pm = smem2;                   // {phi, mu}: row qy at s_pm + qy * PW, column qx at + qx + 1
	double2 *s_trow = s_pm + TR * PW;        // {cos y, sin y} of each tile row's box row
	int4 *s_rowi = (int4 *) (s_trow + TR);   // {by * L, (ly + 1) * WX, ly, 0}
	double *s_rep = (double *) (s_rowi + TR);               // Repulsion tile (jq), as s_pm
	const bool do_rep = a.jq && MODE != MODE_FS && !a.pure;
	const bool fs = MODE != MODE_CF;
	const bool stage_win = a.win && P.omega != 0;
	// box window of sum phi^2 (omega != 0); without it, an in-range stand-in
	// for the discarded window load below
	double *s_win = stage_win ? s_rep + (a.jq ? ((TR * PW + 1) & ~1) : 0) : (double *) smem2;

	// this thread's strip: column qx, rows qy0 .. qy0 + R - 1, the first nr of them its own
	const int R = strip_rows(LX, LY, CT);
	const int qx = tid % LX, blk = tid / LX;
	int qy0 = blk * R;
	const bool idle = qy0 >= LY || blk >= CT / LX;
	if(idle) qy0 = 0;
	const int nr = idle ? 0 : min(R, LY - qy0);
	int lx = qx + C.offx;                 // box-window column (ring offset applied)
	if(lx >= LX) lx -= LX;
	const int row0 = C.start + qy0 * LX + qx;
	const int cx = C.corner % L, cy = C.corner / L;
	double2 *pm = s_pm + qy0 * PW + qx + 1;   // this thread's site of strip row 0

	// the strip's column of phi with one row above and below (periodic), read
	// straight from the rows; pold / dold (corrector) and v (MODE_FS) likewise
	const double *phic = phi_c_in + C.start + qx;
	double col[NPT + 2];
#pragma unroll
	for(int k = 0; k < NPT + 2; ++k) {
		col[k] = 0.;
		if(k <= R + 1) {
			int qy = qy0 + k - 1;
			if(qy < 0) qy += LY;
			if(qy >= LY) qy -= LY;
			col[k] = phic[qy * LX];
		}
	}
	// the box sums of phi^2 at the strip's rows, requested with the strip (rows
	// past the strip read valid box sites and are discarded)
	const int bcol = B.box_off + (cx + lx >= L ? cx + lx - L : cx + lx);
	const double *phi2b = phi2_in + bcol;
	double p2g[NPT];
	unsigned wm[NPT];    // WL: wall mask bits and wall field of the strip's sites
	double wv[NPT];
#pragma unroll
	for(int k = 0; k < NPT; ++k) {
		int qk = qy0 + k;
		if(qk >= LY) qk -= LY;
		if(qk >= LY) qk = 0;
		int ly = qk + C.offy;
		if(ly >= LY) ly -= LY;
		int by = cy + ly;
		if(by >= L) by -= L;
		p2g[k] = phi2b[by * L];
		wm[k] = 31u;
		wv[k] = 0.;
		if constexpr(WL) {
			wm[k] = a.wmask[bcol + by * L];
			wv[k] = a.walls[bcol + by * L];
		}
	}
	// Any status flag halts the device, except that a growth request raised
	// by a corrector of this same step leaves the rest of that corrector stage
	// to run (cells of other stage groups are launched after it); the host
	// serves the growth once the stage is complete.  Correctors of later steps,
	// already queued, halt.  (The loads above stay above: an ordering point.)
	asm volatile("" ::: "memory");
	const int step = a.step + (a.step_base ? *a.step_base : 0);
	if(st_flags) {
		const bool same_step = MODE != MODE_CF && !store && (st_flags & ST_KIND_MASK) == ST_GROWTH &&
				(st_flags >> ST_STEP_SHIFT) == (unsigned) (step + 1);
		if(!same_step) return;
	}
	// row tables, the box window (adhesion only) and the phi tile with its
	// periodic columns
	for(int t = tid; t < TR; t += CT) {
		const int q = t % LY;
		int ly = q + C.offy;
		if(ly >= LY) ly -= LY;
		int by = cy + ly;
		if(by >= L) by -= L;
		s_rowi[t] = make_int4(by * L, (ly + 1) * WX, ly, 0);
		s_trow[t] = make_double2(a.trig[C.trig_y + ly], a.trig[C.trig_y + LY + ly]);
	}
	if(stage_win) {
		int wy = tid / WX, wx = tid - wy * WX;
		const int dy = CT / WX, dx = CT - dy * WX;
		for(int w = tid; w < WX * WY; w += CT) {
			int bx = cx + wx - 1;
			if(bx < 0) bx += L;
			else if(bx >= L) bx -= L;
			int by = cy + wy - 1;
			if(by < 0) by += L;
			else if(by >= L) by -= L;
			s_win[w] = a.phi2[B.box_off + by * L + bx];
			wx += dx;
			wy += dy;
			if(wx >= WX) {
				wx -= WX;
				++wy;
			}
		}
	}
	if(!idle) {
#pragma unroll
		for(int k = 0; k < NPT; ++k) {
			if(k >= R) break;
			pm[k * PW].x = col[k + 1];
			if(qx == 0) pm[k * PW + LX].x = col[k + 1];
			if(qx == LX - 1) pm[k * PW - LX].x = col[k + 1];
		}
	}
	__syncthreads();

	// phase 1: perimeter = sum |grad phi|^2 (computeGlobalSums, WetPolarModel.cpp:243-246),
	// needed before mu only by the perimeter term (beta != 0)
	double perimeter = 0.;
	if(P.beta != 0) {
		double per[1] = { 0.0 };
#pragma unroll
		for(int k = 0; k < NPT; ++k) {
			if(k >= R) break;
			double dx = .5 * (pm[k * PW + 1].x - pm[k * PW - 1].x);
			double dy = .5 * (col[k + 2] - col[k]);
			if(WL && (wm[k] & 3u) != 3u) dx = 0.;
			if(WL && (wm[k] & 12u) != 12u) dy = 0.;
			if(k < nr) per[0] += dx * dx + dy * dy;
		}
		block_sum<1, NW>(per, s_red);
		if(tid == 0) s_bc = per[0];
		__syncthreads();
		perimeter = s_bc;
	}

	// phase 2: chemical potential mu = dF/dphi (f_interaction, WetPolarModel.cpp:262-301).
	// With omega == 0 (beta == 0) the adhesion (perimeter) term is exactly +-0 and is not evaluated.
	// Loop-invariant factors are formed once, in the reference's evaluation order
	// (mu / a0, 4 (kappa / lambda) and 4 (beta / p0) come formed from the host).
	const double Acoef = -4 * P.mu_a0 * (1 - C.area / P.a0);
	const double Rcoef = P.rcoef;
	const double lam2 = 2 * P.lambda;
	const double Pcoef = P.beta != 0 ? P.pcoef * (1 - perimeter / P.p0) : 0.;
	double mu[NPT];
	double per2[1] = { 0.0 };
#pragma unroll
	for(int k = 0; k < NPT; ++k) {
		mu[k] = 0.;
		if(k >= R) break;
		const double ph = col[k + 1];
		const double pE = pm[k * PW + 1].x, pW = pm[k * PW - 1].x, pN = col[k + 2], pS = col[k];
		double lap = pN + pS + pE + pW - 4 * ph;
		double dx = .5 * (pE - pW);
		double dy = .5 * (pN - pS);
		if constexpr(WL) {
			lap = lap_masked(wm[k], pN, pS, pE, pW, ph);
			if((wm[k] & 3u) != 3u) dx = 0.;
			if((wm[k] & 12u) != 12u) dy = 0.;
		}
		const bool valid = k < nr;
		if(MODE == MODE_CF && P.beta == 0 && valid) per2[0] += dx * dx + dy * dy;
		// Both loads always issue and the value is selected: a select of the
		// two pointers would become one generic-address (flat) load.  The
		// window slot read without staging is in range and discarded.
		const int w = stage_win ? s_rowi[qy0 + k].y + lx + 1 : 0;
		const double p2w = s_win[w];
		const double p2 = stage_win ? p2w : p2g[k];
		const double CH = P.gamma * (div_by<RCP>(8 * ph * (1 - ph) * (1 - 2 * ph), P.lambda, P.lambda_rcp) - lam2 * lap);
		const double A = Acoef * ph;
		const double Rep = Rcoef * ph * (p2 - ph * ph);
		double Adh = 0.;
		if(P.omega != 0) {
			double laplacianSquare = s_win[w + 1] + s_win[w + WX] + s_win[w - 1] + s_win[w - WX] - 4. * p2;
			if(WL && !(wm[k] & 16u)) laplacianSquare = s_win[w + 1] + s_win[w - 1] - 2. * p2;   // a wall site
			const double lsquare = 2 * ph * lap + 2 * (dx * dx + dy * dy);
			const double dd = laplacianSquare - lsquare;
			const double suppress = dd / sqrt(1 + dd * dd);
			Adh = -4 * P.lambda * P.omega * suppress * ph;
		}
		double Pp = 0.;
		if(P.beta != 0) Pp = Pcoef * lap;
		double V = CH + A + Rep + Adh + Pp;
		if constexpr(WL) V = f_ext(P, ph, wv[k]) + V;   // freeEnergy starts at F_ext
		mu[k] = V;
		if(!idle) {
			pm[k * PW].y = V;
			if(qx == 0) pm[k * PW + LX].y = V;
			if(qx == LX - 1) pm[k * PW - LX].y = V;
		}
		if(valid) {
			const int row = row0 + k * LX;
			if(MODE == MODE_CF && a.out_mu) {
				a.out_mu[row] = V;
				a.out_dx[row] = dx;
				a.out_dy[row] = dy;
				a.out_lap[row] = lap;
			}
			if(do_rep) {
				const double rn = a.rep[row] + Rep;
				a.rep[row] = rn;
				double *rp = s_rep + (qy0 + k) * PW + qx + 1;
				rp[0] = rn;
				if(qx == 0) rp[LX] = rn;
				if(qx == LX - 1) rp[-LX] = rn;
			}
		}
	}
	// the rows only the update reads (pold / dold of the corrector, v of
	// MODE_FS) are requested after the box loads above, so that their
	// transfer overlaps the evaluation of mu instead of preceding it
	double pold[NPT], dold[NPT];
	double2 vv[NPT];
#pragma unroll
	for(int k = 0; k < NPT; ++k) {
		pold[k] = dold[k] = 0.;
		vv[k] = make_double2(0., 0.);
		if(k < nr) {
			if(fs && !store) {
				pold[k] = a.phi_old[row0 + k * LX];
				dold[k] = a.dphi_old[row0 + k * LX];
			}
			if(MODE == MODE_FS) vv[k] = a.vrows[row0 + k * LX];
		}
	}
	if(MODE == MODE_CF && P.beta == 0) {
		block_sum<1, NW>(per2, s_red);
		if(tid == 0) s_bc = per2[0];
		__syncthreads();
		perimeter = s_bc;
	}
	__syncthreads();   // mu of every site (and the Repulsion tile) is in LDS

	// phase 3: forces (calc_internal_forces) and, unless MODE_CF, the Heun update.
	// The per-cell sums (area, sumF, CoM trig sums, shape) are accumulated with
	// FMA and the x trig sums as cos/sin(x) times the strip's sum of phi: their
	// order already differs from the reference's sequential sum.
	const DevParams Q = a.params[2 * C.box + a.fs_set];
	double red[10];
#pragma unroll
	for(int j = 0; j < 10; ++j) red[j] = 0.0;
	int ymin = LY;
	bool edge = false;
	const bool dry_cf = P.friction_cell == 0;
	double muN = 0., muS = 0.;   // mu one row above / below the strip (storage rows, periodic)
	if(MODE != MODE_FS) {
		int qa = qy0 + R;
		if(qa >= LY) qa -= LY;
		int qb = qy0 - 1;
		if(qb < 0) qb += LY;
		muN = s_pm[qa * PW + qx + 1].y;
		muS = s_pm[qb * PW + qx + 1].y;
	}
	// every row load has landed (they were requested before mu was evaluated):
	// one wait here, not a wait behind the previous row's stores in each row
	// (the memory counter also counts stores)
	__builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
#pragma unroll
	for(int k = 0; k < NPT; ++k) {
		if(k >= R) break;
		const bool valid = k < nr;
		const int row = row0 + k * LX;
		const double ph = col[k + 1];
		const double2 eE = pm[k * PW + 1], eW = pm[k * PW - 1];
		double dx = .5 * (eE.x - eW.x);
		double dy = .5 * (col[k + 2] - col[k]);
		if(WL && (wm[k] & 3u) != 3u) dx = 0.;
		if(WL && (wm[k] & 12u) != 12u) dy = 0.;
		double vx, vy;
		if(MODE == MODE_FS) {
			vx = vv[k].x;
			vy = vv[k].y;
		}
		else {
			const double mN = k + 1 < R ? mu[k + 1] : muN;
			const double mS = k > 0 ? mu[k - 1] : muS;
			const double f_passive_x = (-1) * 0.5 * (eE.y - eW.y);
			const double f_passive_y = (-1) * 0.5 * (mN - mS);
			if(do_rep && valid) {
				int qn = qy0 + k + 1, qs = qy0 + k - 1;
				if(qn == LY) qn = 0;
				if(qs < 0) qs = LY - 1;
				const double *rp = s_rep + (qy0 + k) * PW + qx + 1;
				red[8] += (-1) * 0.5 * (rp[1] - rp[-1]) * ph;
				red[9] += (-1) * 0.5 * (s_rep[qn * PW + qx + 1] - s_rep[qs * PW + qx + 1]) * ph;
			}
			double Theta = (-dx * C.Q00) < 0 ? 0.0 : 1.0;
			const double fQx = P.lambda * (-dx * C.Q00) * Theta * C.Q00;
			Theta = (-dy * C.Q01) < 0 ? 0.0 : 1.0;
			const double fQy = P.lambda * (-dy * C.Q01) * Theta * C.Q01;
			double fx = f_passive_x * P.passive_alpha + fQx * P.zetaQ_self;
			double fy = f_passive_y * P.passive_alpha + fQy * P.zetaQ_self;
			if(WL && !(wm[k] & 16u)) fx = fy = 0.;   // no force on a wall site
			vx = div_by<RCP>(fx, P.friction, P.friction_rcp);
			vy = div_by<RCP>(fy, P.friction, P.friction_rcp);
			if(MODE == MODE_CF) {
				if(valid) {
					if(dry_cf) a.vrows[row] = make_double2(vx, vy);
					else a.frows[row] = make_double2(fx, fy);
				}
				continue;
			}
			if(a.write_v && valid) a.vrows[row] = make_double2(vx, vy);
		}
		// FD_CPUBackend::first_step body (FD_CPUBackend.cpp:66-101)
		double dphi = -Q.J0 * mu[k] - vx * dx - vy * dy;
		if constexpr(WL) {   // external advection (the y velocity is 0)
			if(Q.vext_kind) dphi = dphi - v_ext(Q, s_rowi[qy0 + k].x / L, L, ci - B.cell_begin) * dx;
		}
		double phn;
		if(store) {
			if(valid) a.dphi_old[row] = dphi;
			phn = ph + Q.dt * .5 * (dphi + dphi);
		}
		else {
			phn = pold[k] + Q.dt * .5 * (dphi + dold[k]);
		}
		if(valid) {
			const double2 tr = s_trow[qy0 + k];
			a.phi_out[row] = phn;
			red[0] = fma(phn, phn, red[0]);
			red[1] += phn;
			red[4] = fma(phn, tr.x, red[4]);
			red[5] = fma(phn, tr.y, red[5]);
			if(!store && phn < 0.09 && phn > 0.01) {
				edge = true;
				ymin = min(ymin, s_rowi[qy0 + k].z);
			}
			red[6] += -0.5 * (dx * dx - dy * dy);
			red[7] += -dx * dy;
		}
	}
	int xmin = edge ? lx : LX;
	if(nr > 0 && MODE != MODE_CF) {
		red[2] = red[1] * a.trig[C.trig_x + lx];
		red[3] = red[1] * a.trig[C.trig_x + LX + lx];
	}

	if(MODE == MODE_CF) {
		if(a.pure) {
			// the nematic model keeps no perimeter (WetModel::computeGlobalSums)
			if(tid == 0 && !P.nematic) a.cells[ci].perimeter = perimeter;
			return;
		}
		double tr[2] = { red[8], red[9] };
		block_sum<2, NW>(tr, s_red);
		if(tid == 0) {
			DevCell &G = a.cells[ci];
			G.perimeter = perimeter;
			G.trx = tr[0];
			G.try_ = tr[1];
		}
		return;
	}

	// per-cell sums: area, sumF, the four CoM trig sums, S00, S01 (thread j
	// holds total j), then the CIL sums when they are accumulated
	double trxy[2] = { 0., 0. };
	if(do_rep) {
		trxy[0] = red[8];
		trxy[1] = red[9];
		block_sum<2, NW>(trxy, s_red);
		__syncthreads();
	}
	{   // the interface bounding box: per-wave minima, read after the barrier of block_sum8
		const int lane = tid & 63, wave = tid >> 6;
		const int mx = wave_min(xmin), my = wave_min(ymin);
		if(lane == 0) {
			s_redi[wave] = mx;
			s_redi[NW + wave] = my;
		}
	}
	const double r8[8] = { red[0], red[1], red[2], red[3], red[4], red[5], red[6], red[7] };
	const double tot = block_sum8<NW>(r8, s_red);
	// the totals (threads 0..7) and the CIL sums (block_sum: thread 0) to LDS
	__shared__ double s_tail[16];
	if(tid < 8) s_tail[tid] = tot;
	if(tid == 0) {
		s_tail[8] = trxy[0];
		s_tail[9] = trxy[1];
	}
	__syncthreads();
	const double area = s_tail[0], sumF = s_tail[1];
	const double sxc = s_tail[2], sxs = s_tail[3];
	const double syc = s_tail[4], sys = s_tail[5];
	const double S00 = s_tail[6], S01 = s_tail[7];
	const double trx0 = s_tail[8], try0 = s_tail[9];

	// ---- per-cell tail: lane 0 of waves 0 / 1 takes the two CoM atan2, of
	// waves 2 / 3 the polarity angle and its cos / sin (one function per wave,
	// so the four run side by side on four SIMDs instead of one after another
	// in one wave); thread 0 writes the cell ----
	const double trx = MODE == MODE_FUSED ? (do_rep ? trx0 : 0.) : C.trx;
	const double try_ = MODE == MODE_FUSED ? (do_rep ? try0 : 0.) : C.try_;
	const int twave = tid >> 6;
	const bool nematic = MODE == MODE_FS && Q.nematic;
	if((tid & 63) == 0 && twave < 4) {
		double val = 0.;
		double th_old = C.thetaQ_old, th = 0.;
		if(twave == 0) val = LX * (atan2(-sxs / sumF, -sxc / sumF) + PI_F) / (2 * PI_F);
		else if(twave == 1) val = LY * (atan2(-sys / sumF, -syc / sumF) + PI_F) / (2 * PI_F);
		else if(nematic) {
			// WetModel::updateDirectedActiveForces (WetModel.cpp:657-670): the unit
			// vector of the shape tensor's long axis, or the anchoring direction
			// (update_anchoring, :703-716), which then replaces the director
			double arg = 0.5 * atan2(S01, S00);
			if(Q.anchoring) {
				const double walls_length = 8;
				const double com_old_y = store ? C.CoMy : C.CoM_oldy;
				const double dist1 = (com_old_y - walls_length);
				const double dist2 = ((double) L - walls_length) - com_old_y;
				if(dist1 < dist2) arg = (PI_F / 2) - (PI_F / 2) * dist1 / (((double) L - 2 * walls_length) / 2);
				else arg = (PI_F / 2) + ((PI_F / 2) * dist2 / (((double) L - 2 * walls_length) / 2));
			}
			val = twave == 2 ? cos(arg) : sin(arg);
		}
		else {
			// WetPolarModel::updateDirectedActiveForces (WetPolarModel.cpp:365-396)
			if(store) {
				const double g = a.gauss ? a.gauss[ci] : 0.0;
				th_old = C.thetaQ + Q.sqrt_dt * Q.zetaQ_inter * g;
			}
			th = th_old - Q.dt * Q.J_Q * (atan2(trx * C.Q01 - try_ * C.Q00, trx * C.Q00 + try_ * C.Q01));
			double arg = th;
			if(Q.anchoring) {  // update_anchoring (WetPolarModel.cpp:417-430)
				const double walls_length = 8;
				const double com_old_y = store ? C.CoMy : C.CoM_oldy;
				const double dist1 = (com_old_y - walls_length);
				const double dist2 = ((double) L - walls_length) - com_old_y;
				if(dist1 < dist2) arg = (PI_F / 2) - (PI_F / 2) * dist1 / (((double) L - 2 * walls_length) / 2);
				else arg = (PI_F / 2) + ((PI_F / 2) * dist2 / (((double) L - 2 * walls_length) / 2));
			}
			val = twave == 2 ? cos(arg) : sin(arg);
		}
		s_tail[10 + twave] = val;
		if(twave == 2) {
			s_tail[14] = th;
			s_tail[15] = th_old;
		}
	}
	__syncthreads();
	double CoMx = s_tail[10], CoMy = s_tail[11];
	const double c2 = s_tail[12], s3 = s_tail[13];
	double th = s_tail[14], th_old = s_tail[15];
	if(tid != 0) return;
	if(a.pure) return;   // timing run (shape tuner): no cell state is written
	if(!(area == area) || !(sumF == sumF)) {
		// a NaN field (no counterpart: the reference carries it on): halt
		// before it reaches the cell's geometry; the first cell is named
		if(!(atomicOr(&a.st->flags, ST_NAN) & ST_NAN)) a.st->first_cell = ci;
		return;
	}
	DevCell &G = a.cells[ci];
	if(store) {
		G.CoM_oldx = C.CoMx;
		G.CoM_oldy = C.CoMy;
		G.area_old = C.area;
		G.sumF_old = C.sumF;
	}
	G.area = area;
	G.sumF = sumF;
	G.S00 = S00;
	G.S01 = S01;
	G.perimeter = 0.;
	int xsl = LX, ysb = LY;
	if(!store) {
		for(int w = 0; w < NW; ++w) {
			xsl = min(xsl, s_redi[w]);
			ysb = min(ysb, s_redi[NW + w]);
		}
	}
	G.x_sub_left = xsl;
	G.y_sub_bottom = ysb;
	if(!store) {
		// MultiPhaseField::set_positions (MultiPhaseField.cpp:223-247)
		if(xsl >= BORDER && ysb >= BORDER) {
			const int site = box_elem(L, C.corner, (int) CoMx, (int) CoMy);
			CoMy = (int) (site / L) + (CoMy - (int) CoMy);
			CoMx = (int) (site - (int) (site / L) * L) + (CoMx - (int) CoMx);
			const double co0 = box_x(L, C.corner, 0), co1 = box_y(L, C.corner, 0);
			const int corner = box_elem(L, site, (int) -LX / 2, (int) -LY / 2);
			const double cn0 = box_x(L, corner, 0), cn1 = box_y(L, corner, 0);
			const double d0 = cn0 - co0 - rint((cn0 - co0) / L) * L;
			const double d1 = cn1 - co1 - rint((cn1 - co1) / L) * L;
			const int offx = (int) (C.offx + LX - d0) % LX;
			const int offy = (int) (C.offy + LY - d1) % LY;
			G.corner_old = C.corner;
			G.corner = corner;
			G.unrapx = C.unrapx + d0;
			G.unrapy = C.unrapy + d1;
			G.offx = offx;
			G.offy = offy;
			a.geo[ci] = make_int4(corner, LX | (LY << 16), offx | (offy << 16), C.start);
		}
		else {
			// patch growth (MultiPhaseField.cpp:248-310) is done by the host:
			// flag it, with the step, and halt
			if(atomicOr(&a.st->flags, ST_GROWTH | ((unsigned) (step + 1) << ST_STEP_SHIFT)) == 0u) {
				a.st->first_cell = ci;
				a.st->step = step;
			}
		}
	}
	G.CoMx = CoMx;
	G.CoMy = CoMy;
	if(nematic) {
		const double n0_old = store ? C.nem0 : C.nem_old0, n1_old = store ? C.nem1 : C.nem_old1;
		double n0 = C.nem0, n1 = C.nem1, q00 = C.Q00, q01 = C.Q01;
		if(Q.anchoring) {
			n0 = c2;
			n1 = s3;
			q00 = 0.5 * (n0 * n0 - n1 * n1);
			q01 = n0 * n1;
		}
		else if(sqrt(S00 * S00 + S01 * S01) > 0.00000001) {
			const double sgn = (c2 * n0 + s3 * n1 > 0.0) ? 1.0 : -1.0;
			n0 = n0_old + Q.dt * Q.J_Q * (sgn * c2 - n0);
			n1 = n1_old + Q.dt * Q.J_Q * (sgn * s3 - n1);
			q00 = 0.5 * (n0 * n0 - n1 * n1);
			q01 = n0 * n1;
		}
		G.nem_old0 = n0_old;
		G.nem_old1 = n1_old;
		G.nem0 = n0;
		G.nem1 = n1;
		G.Q00 = q00;
		G.Q01 = q01;
		return;
	}
	G.thetaQ_old = th_old;
	G.thetaQ = th;
	if(Q.anchoring) {
		G.Q00 = 0.5 * (c2 * c2 - s3 * s3);
		G.Q01 = c2 * s3;
	}
	else {
		G.Q00 = c2;
		G.Q01 = s3;
	}
	if(MODE == MODE_FUSED) {
		G.trx = trx;
		G.try_ = try_;
	}
}

// stage kernel shapes (threads per cell, rows per thread strip)
static const int STAGE_SHAPES[][2] = { { 256, 4 }, { 256, 7 }, { 512, 4 }, { 512, 7 }, { 512, 8 }, { 512, 16 } };
constexpr int NSTAGE_SHAPES = 6;

template <int MODE, bool RCP, bool ST, bool WL>
static hipError_t set_lds_r(size_t bytes) {
	hipError_t e;
	const int b = (int) bytes;
	if((e = hipFuncSetAttribute((const void *) k_stage<MODE, 4, RCP, 256, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b)) != hipSuccess) return e;
	if((e = hipFuncSetAttribute((const void *) k_stage<MODE, 7, RCP, 256, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b)) != hipSuccess) return e;
	if((e = hipFuncSetAttribute((const void *) k_stage<MODE, 4, RCP, 512, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b)) != hipSuccess) return e;
	if((e = hipFuncSetAttribute((const void *) k_stage<MODE, 7, RCP, 512, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b)) != hipSuccess) return e;
	if((e = hipFuncSetAttribute((const void *) k_stage<MODE, 8, RCP, 512, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b)) != hipSuccess) return e;
	return hipFuncSetAttribute((const void *) k_stage<MODE, 16, RCP, 512, ST, WL>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
}
template <int MODE, bool ST>
static hipError_t set_lds(size_t bytes) {
	hipError_t e;
	if((e = set_lds_r<MODE, false, ST, false>(bytes)) != hipSuccess) return e;
	if((e = set_lds_r<MODE, true, ST, false>(bytes)) != hipSuccess) return e;
	if((e = set_lds_r<MODE, false, ST, true>(bytes)) != hipSuccess) return e;
	return set_lds_r<MODE, true, ST, true>(bytes);
}

hipError_t stage_set_lds_limit(size_t bytes) {
	hipError_t e;
	if((e = set_lds<MODE_FUSED, true>(bytes)) != hipSuccess) return e;
	if((e = set_lds<MODE_FUSED, false>(bytes)) != hipSuccess) return e;
	if((e = set_lds<MODE_CF, false>(bytes)) != hipSuccess) return e;
	if((e = set_lds<MODE_FS, true>(bytes)) != hipSuccess) return e;
	return set_lds<MODE_FS, false>(bytes);
}

int stage_shapes(int (*shapes)[2]) {
	for(int i = 0; i < NSTAGE_SHAPES; ++i) {
		shapes[i][0] = STAGE_SHAPES[i][0];
		shapes[i][1] = STAGE_SHAPES[i][1];
	}
	return NSTAGE_SHAPES;
}
int stage_strip_rows(int LX, int LY, int threads) { return strip_rows(LX, LY, threads); }

// WL: walls and external forces (a.wmask set)
template <int MODE, bool RCP, bool ST, bool WL>
static void launch_mode_r(const StageArgs &a, int ncells, size_t lds, hipStream_t s) {
	const dim3 g(ncells);
	if(a.threads == 256) {
		if(a.npt == 4) RC_LAUNCH((k_stage<MODE, 4, RCP, 256, ST, WL>), g, dim3(256), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a);
		else RC_LAUNCH((k_stage<MODE, 7, RCP, 256, ST, WL>), g, dim3(256), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a);
		return;
	}
	switch(a.npt) {
	case 4: RC_LAUNCH((k_stage<MODE, 4, RCP, 512, ST, WL>), g, dim3(512), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a); break;
	case 7: RC_LAUNCH((k_stage<MODE, 7, RCP, 512, ST, WL>), g, dim3(512), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a); break;
	case 8: RC_LAUNCH((k_stage<MODE, 8, RCP, 512, ST, WL>), g, dim3(512), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a); break;
	default: RC_LAUNCH((k_stage<MODE, 16, RCP, 512, ST, WL>), g, dim3(512), lds, s, a.cells, a.cell_list, a.st, a.phi_c, a.phi2, a); break;
	}
}
template <int MODE, bool ST>
static void launch_mode_s(const StageArgs &a, int ncells, size_t lds, hipStream_t s) {
	if(a.wmask) {
		if(a.rcp) launch_mode_r<MODE, true, ST, true>(a, ncells, lds, s);
		else launch_mode_r<MODE, false, ST, true>(a, ncells, lds, s);
	}
	else if(a.rcp) launch_mode_r<MODE, true, ST, false>(a, ncells, lds, s);
	else launch_mode_r<MODE, false, ST, false>(a, ncells, lds, s);
}
template <int MODE>
static void launch_mode(const StageArgs &a, int ncells, size_t lds, hipStream_t s) {
	if constexpr(MODE != MODE_CF) {
		if(a.store) {
			launch_mode_s<MODE, true>(a, ncells, lds, s);
			return;
		}
	}
	launch_mode_s<MODE, false>(a, ncells, lds, s);
}

hipError_t launch_stage(int mode, const StageArgs &a, int ncells, size_t lds, hipStream_t s) {
	if(a.threads != 256 && a.threads != 512) return hipErrorInvalidValue;
	if(mode == MODE_FUSED) launch_mode<MODE_FUSED>(a, ncells, lds, s);
	else if(mode == MODE_CF) launch_mode<MODE_CF>(a, ncells, lds, s);
	else launch_mode<MODE_FS>(a, ncells, lds, s);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_stage_p<ST, RCP>: the dry fused stage (force evaluation and Heun update of
// k_stage<MODE_FUSED>, for parameter sets without the Repulsion/CIL sums,
// adhesion or perimeter terms, in periodic boxes) as one persistent launch
// over every cell of every replica.  One 1024-thread workgroup per CU walks a
// contiguous range of cells (the ranges of one XCD are contiguous, as with
// xcd_item), software-pipelined over its cells:
//   * the next cell's strips (phi column with the rows above and below, the
//     box sums of phi^2, phi_old and dphi_old, the trig values of its rows
//     and of the thread's column) are loaded into a second register set while
//     the current cell is evaluated, and the geometry of the cell after that
//     one a step earlier still;
//   * the updated field (and dphi) is kept in LDS and written to the rows at
//     the start of the next cell, coalesced, after that cell's loads have been
//     issued, so that no wait for the stores stands before them;
//   * the per-cell tails (CoM atan2, polarity, recentering or the growth
//     request) of all the workgroup's cells run at the end, four lanes per
//     cell, from the per-cell sums kept in LDS.
// The only wait on memory in a cell's evaluation is the one for its own
// prefetched values, at the top.  Per-site expressions, operand order and the
// per-cell sums are those of k_stage<MODE_FUSED>; the parameters of every box
// are staged in LDS once.  Strips as in k_stage with CT = 1024 threads per
// cell and at most PNPT rows per thread (p_fits, checked by the host).
// ---------------------------------------------------------------------------
constexpr int PCT = P_THREADS;
constexpr int PPAR = 12;   // parameters staged per box (PPar order)
enum PPar { PP_GAMMA, PP_LAMBDA, PP_LAMBDA_RCP, PP_MU_A0, PP_A0, PP_RCOEF, PP_ALPHA, PP_ZETA_SELF, PP_FRICTION,
	PP_FRICTION_RCP, PP_J0, PP_DT };

struct PGeoRaw {   // a cell's geometry and the values its evaluation reads, as loaded (vector registers)
	int4 g0;       // corner, LX | LY << 16, offx | offy << 16, start
	int4 g1;       // box, trig_x, trig_y, pad
	double area, Q00, Q01;
};
struct PGeo {      // the same, uniform
	int LX, LY, start, offx, offy, corner, box, trig_x, trig_y, L, box_off;
	double area, Q00, Q01;
};

// vector loads of one address: kept out of the scalar cache, whose counter is
// shared with LDS and would make the next LDS wait also wait for these loads
__device__ inline PGeoRaw p_load_geo(const StageArgs &a, int ci) {
	int vi = ci;
	asm volatile("" : "+v"(vi));
	PGeoRaw r;
	r.g0 = a.geo[vi];
	r.g1 = *(const int4 *) ((const int *) (a.cells + vi) + offsetof(DevCell, box) / 4);
	const double *dp = (const double *) (a.cells + vi);
	r.area = dp[offsetof(DevCell, area) / 8];
	r.Q00 = dp[offsetof(DevCell, Q00) / 8];
	r.Q01 = dp[offsetof(DevCell, Q01) / 8];
	return r;
}
__device__ inline double rfl_f64(double v) {
	const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
	const unsigned lo = (unsigned) __builtin_amdgcn_readfirstlane((int) (unsigned) u);
	const unsigned hi = (unsigned) __builtin_amdgcn_readfirstlane((int) (unsigned) (u >> 32));
	return __builtin_bit_cast(double, ((unsigned long long) hi << 32) | lo);
}
__device__ inline PGeo p_geo(const PGeoRaw &r, const int2 *s_box) {
	PGeo g;
	g.corner = __builtin_amdgcn_readfirstlane(r.g0.x);
	const int lxy = __builtin_amdgcn_readfirstlane(r.g0.y), off = __builtin_amdgcn_readfirstlane(r.g0.z);
	g.LX = lxy & 0xFFFF;
	g.LY = lxy >> 16;
	g.offx = off & 0xFFFF;
	g.offy = off >> 16;
	g.start = __builtin_amdgcn_readfirstlane(r.g0.w);
	g.box = __builtin_amdgcn_readfirstlane(r.g1.x);
	g.trig_x = __builtin_amdgcn_readfirstlane(r.g1.y);
	g.trig_y = __builtin_amdgcn_readfirstlane(r.g1.z);
	const int2 bb = s_box[g.box];
	g.L = __builtin_amdgcn_readfirstlane(bb.x);
	g.box_off = __builtin_amdgcn_readfirstlane(bb.y);
	g.area = rfl_f64(r.area);
	g.Q00 = rfl_f64(r.Q00);
	g.Q01 = rfl_f64(r.Q01);
	return g;
}

// this thread's strip of a cell (k_stage's column strips, CT = PCT)
struct PStrip {
	int R, qx, qy0, nr, lx;
	bool idle;
};
// n / d for 0 <= n < 2048 and 2 <= d <= 1024 with a multiply-high: m =
// floor((2^32 - 1) / d) + 1 exceeds 2^32 / d by e in (0, 1], so n m / 2^32 =
// n / d + n e / 2^32 with n e / 2^32 < 2^-21 < 1 / d: the floor is exact
__device__ inline unsigned p_magic(unsigned d) { return 0xFFFFFFFFu / d + 1u; }
__device__ inline int p_div(int n, unsigned m) { return (int) __umulhi((unsigned) n, m); }

__device__ inline PStrip p_strip(int LX, int LY, int offx) {
	PStrip t;
	const int tid = threadIdx.x;
	const unsigned mx = p_magic((unsigned) LX);   // uniform
	t.R = strip_rows(LX, LY, PCT);
	const int blk = p_div(tid, mx);
	t.qx = tid - blk * LX;
	t.qy0 = blk * t.R;
	t.idle = t.qy0 >= LY || blk >= PCT / LX;
	if(t.idle) t.qy0 = 0;
	t.nr = t.idle ? 0 : min(t.R, LY - t.qy0);
	t.lx = t.qx + offx;
	if(t.lx >= LX) t.lx -= LX;
	return t;
}

// the values a cell's evaluation reads, per thread
template <int PNPT>
struct PData {
	double col[PNPT + 2], p2[PNPT], pold[PNPT], dold[PNPT];
	double2 trow;     // {cos y, sin y} of tile row threadIdx.x (threads < LY + PNPT)
	double2 trx;      // {cos x, sin x} of this thread's column
};

template <bool ST, int PNPT>
__device__ inline void p_load(const StageArgs &a, const PGeo &g, PData<PNPT> &d) {
	const int tid = threadIdx.x, LX = g.LX, LY = g.LY, L = g.L;
	const PStrip t = p_strip(LX, LY, g.offx);
	const int cx = g.corner % L, cy = g.corner / L;
	const double *phic = a.phi_c + g.start + t.qx;
#pragma unroll
	for(int k = 0; k < PNPT + 2; ++k) {
		int qy = t.qy0 + k - 1;
		if(qy < 0) qy += LY;
		if(qy >= LY) qy -= LY;
		if(qy >= LY) qy -= LY;
		d.col[k] = phic[qy * LX];
	}
	const int bcol = g.box_off + (cx + t.lx >= L ? cx + t.lx - L : cx + t.lx);
	const double *phi2b = a.phi2 + bcol;
#pragma unroll
	for(int k = 0; k < PNPT; ++k) {
		int qk = t.qy0 + k;
		if(qk >= LY) qk -= LY;
		if(qk >= LY) qk = 0;
		int ly = qk + g.offy;
		if(ly >= LY) ly -= LY;
		int by = cy + ly;
		if(by >= L) by -= L;
		d.p2[k] = phi2b[by * L];
	}
	if(!ST) {
		// rows past the strip's own (k >= nr) read row qy0 instead and are discarded
		const int row0 = g.start + t.qy0 * LX + t.qx;
#pragma unroll
		for(int k = 0; k < PNPT; ++k) {
			const int rk = k < t.nr ? row0 + k * LX : row0;
			d.pold[k] = a.phi_old[rk];
			d.dold[k] = a.dphi_old[rk];
		}
	}
	else {
#pragma unroll
		for(int k = 0; k < PNPT; ++k) d.pold[k] = d.dold[k] = 0.;
	}
	int ly = tid - p_div(tid, p_magic((unsigned) LY)) * LY + g.offy;
	if(ly >= LY) ly -= LY;
	d.trow = make_double2(a.trig[g.trig_y + ly], a.trig[g.trig_y + LY + ly]);
	d.trx = make_double2(a.trig[g.trig_x + t.lx], a.trig[g.trig_x + LX + t.lx]);
}

// LDS regions of k_stage_p
struct PLds {
	double2 *pm, *trow;
	double *phn, *dph, *par, *sum, *red;
	int2 *box;
	int *min, *redi;
};

// one cell of k_stage_p: evaluates cell ci (geometry g, values cur) and, once
// its tile is in LDS, requests the values of cell cn (into gn / nxt, from the
// geometry rn loaded one cell earlier) and the geometry of cell cn2 (into rn),
// then writes the update staged by the previous cell and stages its own
template <bool ST, bool RCP, int PNPT>
__device__ __forceinline__ void p_cell(const StageArgs &a, const PLds &L_, int ci, int j, int cn, int cn2,
		const PGeo &g, const PData<PNPT> &cur, PGeo &gn, PData<PNPT> &nxt, PGeoRaw &rn, int &prev_start, int &prev_S) {
	constexpr bool store = ST;
	constexpr int NW = PCT / 64;
	const int tid = threadIdx.x;
	const int LX = g.LX, LY = g.LY, PW = LX + 2;
	const PStrip t = p_strip(LX, LY, g.offx);
	const int R = t.R, qx = t.qx, qy0 = t.qy0, nr = t.nr, lx = t.lx;
	const bool idle = t.idle;
	double2 *const s_pm = L_.pm;
	// the tile: rows 0 .. LY - 1 and, past them, images of rows 0 .. PNPT - 1
	double2 *pm = s_pm + qy0 * PW + qx + 1;
	if(tid < LY + PNPT) L_.trow[tid] = cur.trow;
	if(!idle) {
#pragma unroll
		for(int k = 0; k < PNPT; ++k) {
			if(k >= R) break;
			const double v = cur.col[k + 1];
			double2 *e = pm + k * PW;
			e->x = v;
			if(qx == 0) e[LX].x = v;
			if(qx == LX - 1) e[-LX].x = v;
			if(qy0 + k < PNPT) {
				double2 *im = e + LY * PW;
				im->x = v;
				if(qx == 0) im[LX].x = v;
				if(qx == LX - 1) im[-LX].x = v;
			}
		}
	}
	__syncthreads();
	// the next cell's loads, then the previous cell's staged update
	gn = p_geo(rn, L_.box);
	rn = p_load_geo(a, cn2);
	p_load<ST, PNPT>(a, gn, nxt);
	for(int qq = tid; qq < prev_S; qq += PCT) {
		a.phi_out[prev_start + qq] = L_.phn[qq];
		if(ST) a.dphi_old[prev_start + qq] = L_.dph[qq];
	}
	(void) cn;

	// chemical potential (f_interaction, WetPolarModel.cpp:262-301) without
	// the adhesion and perimeter terms (omega = beta = 0: both exactly +0)
	const double *P = L_.par + PPAR * g.box;
	const double Acoef = -4 * P[PP_MU_A0] * (1 - g.area / P[PP_A0]);
	const double Rcoef = P[PP_RCOEF];
	const double lambda = P[PP_LAMBDA];
	const double lam2 = 2 * lambda;
	double mu[PNPT];
#pragma unroll
	for(int k = 0; k < PNPT; ++k) {
		mu[k] = 0.;
		if(k >= R) break;
		const double2 *e = pm + k * PW;
		const double ph = e->x;
		const double pE = e[1].x, pW = e[-1].x, pN = e[PW].x;
		const double pS = qy0 + k == 0 ? s_pm[(LY - 1) * PW + qx + 1].x : e[-PW].x;
		const double lap = pN + pS + pE + pW - 4 * ph;
		const double p2 = cur.p2[k];
		const double CH = P[PP_GAMMA] * (div_by<RCP>(8 * ph * (1 - ph) * (1 - 2 * ph), lambda, P[PP_LAMBDA_RCP]) - lam2 * lap);
		const double A = Acoef * ph;
		const double Rep = Rcoef * ph * (p2 - ph * ph);
		const double Adh = 0., Pp = 0.;
		const double V = CH + A + Rep + Adh + Pp;
		mu[k] = V;
		if(!idle) {
			double2 *w = pm + k * PW;
			w->y = V;
			if(qx == 0) w[LX].y = V;
			if(qx == LX - 1) w[-LX].y = V;
			if(qy0 + k < PNPT) {
				double2 *im = w + LY * PW;
				im->y = V;
				if(qx == 0) im[LX].y = V;
				if(qx == LX - 1) im[-LX].y = V;
			}
		}
	}
	__syncthreads();

	// forces (calc_internal_forces) and the Heun update (first_step); the
	// update is staged in LDS (storage order) and written with the next cell
	const double alpha = P[PP_ALPHA], zself = P[PP_ZETA_SELF], fric = P[PP_FRICTION], frcp = P[PP_FRICTION_RCP];
	const double J0 = P[PP_J0], dt = P[PP_DT];
	double red[8];
#pragma unroll
	for(int jj = 0; jj < 8; ++jj) red[jj] = 0.0;
	int ymin = LY;
	bool edge = false;
	double muN, muS;
	{
		int qa = qy0 + R;
		if(qa >= LY) qa -= LY;
		if(qa >= LY) qa -= LY;
		int qb = qy0 - 1;
		if(qb < 0) qb += LY;
		muN = s_pm[qa * PW + qx + 1].y;
		muS = s_pm[qb * PW + qx + 1].y;
	}
#pragma unroll
	for(int k = 0; k < PNPT; ++k) {
		if(k >= R) break;
		const bool valid = k < nr;
		const double2 *e = pm + k * PW;
		const double ph = e->x;
		const double pN = e[PW].x;
		const double pS = qy0 + k == 0 ? s_pm[(LY - 1) * PW + qx + 1].x : e[-PW].x;
		const double2 eE = e[1], eW = e[-1];
		const double dx = .5 * (eE.x - eW.x);
		const double dy = .5 * (pN - pS);
		const double mN = k + 1 < R ? mu[k + 1] : muN;
		const double mS = k > 0 ? mu[k - 1] : muS;
		const double f_passive_x = (-1) * 0.5 * (eE.y - eW.y);
		const double f_passive_y = (-1) * 0.5 * (mN - mS);
		double Theta = (-dx * g.Q00) < 0 ? 0.0 : 1.0;
		const double fQx = lambda * (-dx * g.Q00) * Theta * g.Q00;
		Theta = (-dy * g.Q01) < 0 ? 0.0 : 1.0;
		const double fQy = lambda * (-dy * g.Q01) * Theta * g.Q01;
		const double fx = f_passive_x * alpha + fQx * zself;
		const double fy = f_passive_y * alpha + fQy * zself;
		const double vx = div_by<RCP>(fx, fric, frcp);
		const double vy = div_by<RCP>(fy, fric, frcp);
		const double dphi = -J0 * mu[k] - vx * dx - vy * dy;
		double phn;
		if(store) phn = ph + dt * .5 * (dphi + dphi);
		else phn = cur.pold[k] + dt * .5 * (dphi + cur.dold[k]);
		if(valid) {
			const int qs = (qy0 + k) * LX + qx;
			L_.phn[qs] = phn;
			if(ST) L_.dph[qs] = dphi;
			const double2 tr = L_.trow[qy0 + k];
			red[0] = fma(phn, phn, red[0]);
			red[1] += phn;
			red[4] = fma(phn, tr.x, red[4]);
			red[5] = fma(phn, tr.y, red[5]);
			if(!store && phn < 0.09 && phn > 0.01) {
				edge = true;
				int ly = qy0 + k + g.offy;
				if(ly >= LY) ly -= LY;
				ymin = min(ymin, ly);
			}
			red[6] += -0.5 * (dx * dx - dy * dy);
			red[7] += -dx * dy;
		}
	}
	const int xmin = edge ? lx : LX;
	if(nr > 0) {
		red[2] = red[1] * cur.trx.x;
		red[3] = red[1] * cur.trx.y;
	}
	{
		const int lane = tid & 63, wave = tid >> 6;
		const int mx = wave_min(xmin), my = wave_min(ymin);
		if(lane == 0) {
			L_.redi[wave] = mx;
			L_.redi[NW + wave] = my;
		}
	}
	const double tot = block_sum8<NW>(red, L_.red);
	if(tid < 8) L_.sum[8 * j + tid] = tot;
	if(tid == 0) {
		int xsl = LX, ysb = LY;
		for(int w = 0; w < NW; ++w) {
			xsl = min(xsl, L_.redi[w]);
			ysb = min(ysb, L_.redi[NW + w]);
		}
		L_.min[2 * j] = xsl;
		L_.min[2 * j + 1] = ysb;
	}
	prev_start = g.start;
	prev_S = LX * LY;
	(void) ci;
}

template <bool ST, bool RCP, int PNPT>
__global__ __launch_bounds__(PCT, 1) void k_stage_p(StageArgs a, int ncells) {
	constexpr bool store = ST;
	const int step = a.step + (a.step_base ? *a.step_base : 0);
	{   // as k_stage: a growth request of this same corrector does not halt it
		const unsigned f = a.st->flags;
		if(f) {
			const bool same_step = !store && (f & ST_KIND_MASK) == ST_GROWTH && (f >> ST_STEP_SHIFT) == (unsigned) (step + 1);
			if(!same_step) return;
		}
	}
	constexpr int NW = PCT / 64;
	extern __shared__ double2 smem2[];
	__shared__ double s_red[NW * 4 * 8];
	__shared__ int s_redi[2 * NW];
	const int tid = threadIdx.x;
	const int q = xcd_item(blockIdx.x, gridDim.x);
	const int c0 = (int) ((long long) q * ncells / gridDim.x), c1 = (int) ((long long) (q + 1) * ncells / gridDim.x);
	if(c0 >= c1) return;
	// LDS: the {phi, mu} tile (a.p_tile double2), the row table (a.p_rows),
	// the staged field update and dphi (a.p_sites each), the boxes'
	// parameters and {L, box_off}, then per cell of the range 8 sums and the
	// interface box minima
	PLds L_;
	L_.pm = smem2;
	L_.trow = L_.pm + a.p_tile;
	L_.phn = (double *) (L_.trow + a.p_rows);
	L_.dph = L_.phn + a.p_sites;
	L_.par = L_.dph + (ST ? a.p_sites : 0);
	L_.box = (int2 *) (L_.par + PPAR * 2 * a.p_nbox);
	L_.sum = (double *) (L_.box + a.p_nbox + (a.p_nbox & 1));
	L_.min = (int *) (L_.sum + 8 * (c1 - c0));
	L_.red = s_red;
	L_.redi = s_redi;
	double *s_sum = L_.sum;
	int *s_min = L_.min;
	for(int i = tid; i < a.p_nbox; i += PCT) {
		const DevBox B = a.boxes[i];
		L_.box[i] = make_int2(B.L, B.box_off);
		const DevParams &P = a.params[2 * i + a.cf_set];
		const DevParams &Q = a.params[2 * i + a.fs_set];
		double *w = L_.par + PPAR * i;
		w[PP_GAMMA] = P.gamma;
		w[PP_LAMBDA] = P.lambda;
		w[PP_LAMBDA_RCP] = P.lambda_rcp;
		w[PP_MU_A0] = P.mu_a0;
		w[PP_A0] = P.a0;
		w[PP_RCOEF] = P.rcoef;
		w[PP_ALPHA] = P.passive_alpha;
		w[PP_ZETA_SELF] = P.zetaQ_self;
		w[PP_FRICTION] = P.friction;
		w[PP_FRICTION_RCP] = P.friction_rcp;
		w[PP_J0] = Q.J0;
		w[PP_DT] = Q.dt;
	}
	__syncthreads();

	// two register sets, A and B, alternate between the cell being evaluated
	// and the next one being loaded; cells past the range reload the last one
	const int last = c1 - 1;
	PGeoRaw rn = p_load_geo(a, c0);
	PGeo gA = p_geo(rn, L_.box), gB;
	rn = p_load_geo(a, min(c0 + 1, last));
	PData<PNPT> dA, dB;
	p_load<ST, PNPT>(a, gA, dA);
	int prev_start = 0, prev_S = 0;   // the cell whose update is staged in LDS
	for(int ci = c0; ci < c1; ci += 2) {
		p_cell<ST, RCP, PNPT>(a, L_, ci, ci - c0, min(ci + 1, last), min(ci + 2, last), gA, dA, gB, dB, rn, prev_start, prev_S);
		if(ci + 1 >= c1) break;
		p_cell<ST, RCP, PNPT>(a, L_, ci + 1, ci + 1 - c0, min(ci + 2, last), min(ci + 3, last), gB, dB, gA, dA, rn, prev_start,
				prev_S);
	}
	__syncthreads();
	for(int qq = tid; qq < prev_S; qq += PCT) {
		a.phi_out[prev_start + qq] = L_.phn[qq];
		if(ST) a.dphi_old[prev_start + qq] = L_.dph[qq];
	}
	if(a.pure) return;   // timing run (shape tuner): no cell state is written

	// ---- per-cell tails, four lanes per cell (k_stage's tail) ----
	const int sub = tid & 3, base = (tid & 63) & ~3;
	for(int j0 = 0; j0 < c1 - c0; j0 += PCT / 4) {
		const int j = j0 + (tid >> 2);
		const bool on = j < c1 - c0;
		const int ci = on ? c0 + j : c0;
		const DevCell C = a.cells[ci];
		const DevBox B = a.boxes[C.box];
		const DevParams &Q = a.params[2 * C.box + a.fs_set];
		const int LX = C.LX, LY = C.LY, L = B.L;
		const double *sm = s_sum + 8 * (ci - c0);
		const double area = sm[0], sumF = sm[1], sxc = sm[2], sxs = sm[3];
		const double syc = sm[4], sys = sm[5], S00 = sm[6], S01 = sm[7];
		const double trx = 0., try_ = 0.;
		double val = 0.;
		double th_old = C.thetaQ_old, th = 0.;
		if(sub == 0) val = LX * (atan2(-sxs / sumF, -sxc / sumF) + PI_F) / (2 * PI_F);
		else if(sub == 1) val = LY * (atan2(-sys / sumF, -syc / sumF) + PI_F) / (2 * PI_F);
		else {
			// WetPolarModel::updateDirectedActiveForces (WetPolarModel.cpp:365-396)
			if(store) {
				const double gs = a.gauss ? a.gauss[ci] : 0.0;
				th_old = C.thetaQ + Q.sqrt_dt * Q.zetaQ_inter * gs;
			}
			th = th_old - Q.dt * Q.J_Q * (atan2(trx * C.Q01 - try_ * C.Q00, trx * C.Q00 + try_ * C.Q01));
			double arg = th;
			if(Q.anchoring) {  // update_anchoring (WetPolarModel.cpp:417-430)
				const double walls_length = 8;
				const double com_old_y = store ? C.CoMy : C.CoM_oldy;
				const double dist1 = (com_old_y - walls_length);
				const double dist2 = ((double) L - walls_length) - com_old_y;
				if(dist1 < dist2) arg = (PI_F / 2) - (PI_F / 2) * dist1 / (((double) L - 2 * walls_length) / 2);
				else arg = (PI_F / 2) + ((PI_F / 2) * dist2 / (((double) L - 2 * walls_length) / 2));
			}
			val = sub == 2 ? cos(arg) : sin(arg);
		}
		double CoMx = __shfl(val, base + 0, 64), CoMy = __shfl(val, base + 1, 64);
		const double c2 = __shfl(val, base + 2, 64), s3 = __shfl(val, base + 3, 64);
		th = __shfl(th, base + 2, 64);
		th_old = __shfl(th_old, base + 2, 64);
		if(!on || sub != 0) continue;
		if(!(area == area) || !(sumF == sumF)) {
			// a NaN field (no counterpart: the reference carries it on): halt
			// before it reaches the cell's geometry; the first cell is named
			if(!(atomicOr(&a.st->flags, ST_NAN) & ST_NAN)) a.st->first_cell = ci;
			continue;
		}
		DevCell &G = a.cells[ci];
		if(store) {
			G.CoM_oldx = C.CoMx;
			G.CoM_oldy = C.CoMy;
			G.area_old = C.area;
			G.sumF_old = C.sumF;
		}
		G.area = area;
		G.sumF = sumF;
		G.S00 = S00;
		G.S01 = S01;
		G.perimeter = 0.;
		const int xsl = store ? LX : s_min[2 * j], ysb = store ? LY : s_min[2 * j + 1];
		G.x_sub_left = xsl;
		G.y_sub_bottom = ysb;
		if(!store) {
			// MultiPhaseField::set_positions (MultiPhaseField.cpp:223-247)
			if(xsl >= BORDER && ysb >= BORDER) {
				const int site = box_elem(L, C.corner, (int) CoMx, (int) CoMy);
				CoMy = (int) (site / L) + (CoMy - (int) CoMy);
				CoMx = (int) (site - (int) (site / L) * L) + (CoMx - (int) CoMx);
				const double co0 = box_x(L, C.corner, 0), co1 = box_y(L, C.corner, 0);
				const int corner = box_elem(L, site, (int) -LX / 2, (int) -LY / 2);
				const double cn0 = box_x(L, corner, 0), cn1 = box_y(L, corner, 0);
				const double d0 = cn0 - co0 - rint((cn0 - co0) / L) * L;
				const double d1 = cn1 - co1 - rint((cn1 - co1) / L) * L;
				const int offx = (int) (C.offx + LX - d0) % LX;
				const int offy = (int) (C.offy + LY - d1) % LY;
				G.corner_old = C.corner;
				G.corner = corner;
				G.unrapx = C.unrapx + d0;
				G.unrapy = C.unrapy + d1;
				G.offx = offx;
				G.offy = offy;
				a.geo[ci] = make_int4(corner, LX | (LY << 16), offx | (offy << 16), C.start);
			}
			else {
				// patch growth (MultiPhaseField.cpp:248-310) is done by the host:
				// flag it, with the step, and halt
				if(atomicOr(&a.st->flags, ST_GROWTH | ((unsigned) (step + 1) << ST_STEP_SHIFT)) == 0u) {
					a.st->first_cell = ci;
					a.st->step = step;
				}
			}
		}
		G.CoMx = CoMx;
		G.CoMy = CoMy;
		G.thetaQ_old = th_old;
		G.thetaQ = th;
		if(Q.anchoring) {
			G.Q00 = 0.5 * (c2 * c2 - s3 * s3);
			G.Q01 = c2 * s3;
		}
		else {
			G.Q00 = c2;
			G.Q01 = s3;
		}
		G.trx = trx;
		G.try_ = try_;
	}
}

template <int PNPT>
static hipError_t launch_stage_p_n(const StageArgs &a0, int ncells, int grid, size_t lds, hipStream_t s) {
	static size_t lds_set[4] = { 0, 0, 0, 0 };
	const int v = (a0.store ? 1 : 0) + (a0.rcp ? 2 : 0);
	const void *fn = v == 0 ? (const void *) k_stage_p<false, false, PNPT> : v == 1 ? (const void *) k_stage_p<true, false, PNPT>
			: v == 2 ? (const void *) k_stage_p<false, true, PNPT> : (const void *) k_stage_p<true, true, PNPT>;
	if(lds > lds_set[v]) {
		hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
		if(e != hipSuccess) return e;
		lds_set[v] = lds;
	}
	switch(v) {
	case 0: RC_LAUNCH((k_stage_p<false, false, PNPT>), dim3(grid), dim3(PCT), lds, s, a0, ncells); break;
	case 1: RC_LAUNCH((k_stage_p<true, false, PNPT>), dim3(grid), dim3(PCT), lds, s, a0, ncells); break;
	case 2: RC_LAUNCH((k_stage_p<false, true, PNPT>), dim3(grid), dim3(PCT), lds, s, a0, ncells); break;
	default: RC_LAUNCH((k_stage_p<true, true, PNPT>), dim3(grid), dim3(PCT), lds, s, a0, ncells); break;
	}
	return hipGetLastError();
}

hipError_t launch_stage_p(const StageArgs &a0, int ncells, int grid, hipStream_t s) {
	// a0.npt: strip rows of the launch (2 or 3, at least every cell's)
	if(ncells <= 0) return hipSuccess;
	if(a0.npt != 2 && a0.npt != 3) return hipErrorInvalidValue;
	if(grid > ncells) grid = ncells;
	const int per = (ncells + grid - 1) / grid;
	const size_t lds = stage_p_lds_bytes(a0, per);
	return a0.npt == 2 ? launch_stage_p_n<2>(a0, ncells, grid, lds, s) : launch_stage_p_n<3>(a0, ncells, grid, lds, s);
}

// ---------------------------------------------------------------------------
// k_nem_force: the nematic model's force evaluation of one cell per
// workgroup (WetModel::f_interaction and calc_internal_forces,
// WetModel.cpp:488-641).  Pass 1 stages phi and evaluates per site the
// chemical potential mu = CH + A + Rep + Adh, the free-energy density
// f = CH_d + A_d + Rep_d and the central differences; it keeps f - mu phi,
// dx, dy in LDS.  Pass 2 takes minus the discrete divergence of the stress
// (f - mu phi) I + grad phi (x) 2 gamma lambda grad phi over the patch's
// periodic neighbours, adds the cell's own nematic force and the inter-cell
// one from the box sums of phi Q, and writes f (wet) or f / friction (dry).
// Expressions keep the reference's operand order.
// ---------------------------------------------------------------------------
// The cell records, the status word, phi and the cell count lead the
// arguments (preloaded in SGPRs): the cell's record is requested at the first
// instruction, and the status word with it (tested before anything is
// written).
__global__ __launch_bounds__(256) void k_nem_force(const DevCell *cells_in, const DevStatus *st_in, const double *phi_in,
		int ncells, NemArgs a) {
	const unsigned flags = st_in->flags;
	extern __shared__ double s_nem[];
	const int ci = xcd_item(blockIdx.x, ncells), tid = threadIdx.x;
	const DevCell C = cells_in[ci];
	const DevBox B = a.boxes[C.box];
	const DevParams P = a.params[2 * C.box + a.set];
	const int LX = C.LX, LY = C.LY, S = C.S, L = B.L;
	const int cx = C.corner % L, cy = C.corner / L;
	// LDS: phi and f - mu phi of the patch; in a wall box also the masked
	// central differences (elsewhere they are formed again from phi where the
	// force needs them: half the LDS, twice the workgroups per CU)
	const bool walls = a.wmask != nullptr;
	double *s_phi = s_nem, *s_s = s_nem + S, *s_dx = s_nem + 2 * S, *s_dy = s_nem + 3 * S;
	// the patch's phi: every load of the thread requested before the first
	// LDS write (one memory round trip, not one per row of 256 sites)
	constexpr int NST = 10;   // 2560 sites in one batch (41^2 = 1681; grown patches up to 49^2 = 2401)
	double phs[NST];
#pragma unroll
	for(int u = 0; u < NST; ++u) phs[u] = phi_in[C.start + min(tid + 256 * u, S - 1)];
	for(int q = tid + 256 * NST; q < S; q += 256) s_phi[q] = phi_in[C.start + q];
#pragma unroll
	for(int u = 0; u < NST; ++u)
		if(tid + 256 * u < S) s_phi[tid + 256 * u] = phs[u];
	if(flags) return;
	__syncthreads();
	const double *phi2 = a.phi2 + B.box_off;
	// box site of storage site (qx, qy) and its four box neighbours
	auto box_site = [&](int qx, int qy, int &kE, int &kW, int &kN, int &kS) {
		int lx = qx + C.offx;
		if(lx >= LX) lx -= LX;
		int ly = qy + C.offy;
		if(ly >= LY) ly -= LY;
		int bx = cx + lx;
		if(bx >= L) bx -= L;
		int by = cy + ly;
		if(by >= L) by -= L;
		const int bxe = bx + 1 == L ? 0 : bx + 1, bxw = bx == 0 ? L - 1 : bx - 1;
		const int byn = by + 1 == L ? 0 : by + 1, bys = by == 0 ? L - 1 : by - 1;
		kE = by * L + bxe;
		kW = by * L + bxw;
		kN = byn * L + bx;
		kS = bys * L + bx;
		return by * L + bx;
	};
	// sites in batches of NB per thread: the batch's box loads are issued
	// together, then evaluated (q = qy LX + qx by a multiply-high division,
	// exact for q, LX < 2^16)
	constexpr int NB = 4;
	const unsigned mlx = 0xFFFFFFFFu / (unsigned) LX + 1u;
	const unsigned char *wmb = a.wmask ? a.wmask + B.box_off : nullptr;
	const double *wlb = a.wmask ? a.walls + B.box_off : nullptr;
	for(int q0 = tid; q0 < S; q0 += NB * 256) {
		int kk[NB], kE[NB], kW[NB], kN[NB], kS[NB];
		double p2[NB], w[NB];
		unsigned m[NB];
#pragma unroll
		for(int u = 0; u < NB; ++u) {
			const int q = min(q0 + u * 256, S - 1);
			const int qy = (int) __umulhi((unsigned) q, mlx), qx = q - qy * LX;
			kk[u] = box_site(qx, qy, kE[u], kW[u], kN[u], kS[u]);
		}
#pragma unroll
		for(int u = 0; u < NB; ++u) {
			p2[u] = phi2[kk[u]];
			m[u] = wmb ? wmb[kk[u]] : 31u;
			w[u] = wlb ? wlb[kk[u]] : 0.;
		}
#pragma unroll
		for(int u = 0; u < NB; ++u) {
			const int q = q0 + u * 256;
			if(q >= S) break;
			const int qy = (int) __umulhi((unsigned) q, mlx), qx = q - qy * LX;
			const int qE = qx + 1 == LX ? q - qx : q + 1, qW = qx == 0 ? q + LX - 1 : q - 1;
			const int qN = qy + 1 == LY ? qx : q + LX, qS = qy == 0 ? q + (LY - 1) * LX : q - LX;
			const double ph = s_phi[q], pE = s_phi[qE], pW = s_phi[qW], pN = s_phi[qN], pS = s_phi[qS];
			const unsigned mm = m[u];
			const double lap = a.wmask ? lap_masked(mm, pN, pS, pE, pW, ph) : pN + pS + pE + pW - 4 * ph;
			const double dx = (mm & 3u) == 3u ? .5 * (pE - pW) : 0.;
			const double dy = (mm & 12u) == 12u ? .5 * (pN - pS) : 0.;
			const double p2q = p2[u];
			const double CH = P.gamma * (8 * ph * (1 - ph) * (1 - 2 * ph) / P.lambda - 2 * P.lambda * lap);
			const double A = -4 * P.mu_a0 * (1 - C.area / P.a0) * ph;
			const double Rep = P.rcoef * ph * (p2q - ph * ph);
			double Adh = 0.;
			if(P.omega != 0) {
				const double laplacianSquare = (mm & 16u) ? phi2[kE[u]] + phi2[kN[u]] + phi2[kW[u]] + phi2[kS[u]] - 4. * p2q
						: phi2[kE[u]] + phi2[kW[u]] - 2. * p2q;   // a wall site
				const double lsquare = 2 * ph * lap + 2 * (dx * dx + dy * dy);
				const double dd = laplacianSquare - lsquare;
				Adh = -4 * P.lambda * P.omega * (dd / sqrt(1 + dd * dd)) * ph;
			}
			double mu = CH + A + Rep + Adh;
			if(a.wmask) mu = f_ext(P, ph, w[u]) + mu;   // freeEnergy starts at F_ext
			const double CH_d = (P.gamma / P.lambda) * 4 * ph * ph * (1 - ph) * (1 - ph) + P.gamma * P.lambda * (dx * dx + dy * dy);
			const double A_d = -P.mu / P.a0 * ph * ph;
			const double Rep_d = (P.kappa / P.lambda) * ph * ph * (p2q - ph * ph);
			double fd = CH_d + A_d + Rep_d;
			if(a.wmask) fd = f_density_ext(P, ph, w[u]) + fd;
			s_s[q] = fd - mu * ph;
			if(walls) {
				s_dx[q] = dx;
				s_dy[q] = dy;
			}
		}
	}
	__syncthreads();
	const double gl = 2 * P.gamma * P.lambda;
	const double *sq00 = a.sq00 + B.box_off, *sq01 = a.sq01 + B.box_off;
	const bool dry = P.friction_cell == 0;
	for(int q0 = tid; q0 < S; q0 += NB * 256) {
		double e00[NB], w00[NB], n01[NB], s01[NB], e01[NB], w01[NB], n00[NB], s00[NB];
		unsigned m[NB];
#pragma unroll
		for(int u = 0; u < NB; ++u) {
			const int q = min(q0 + u * 256, S - 1);
			const int qy = (int) __umulhi((unsigned) q, mlx), qx = q - qy * LX;
			int kE, kW, kN, kS;
			const int k = box_site(qx, qy, kE, kW, kN, kS);
			e00[u] = sq00[kE]; w00[u] = sq00[kW]; n00[u] = sq00[kN]; s00[u] = sq00[kS];
			e01[u] = sq01[kE]; w01[u] = sq01[kW]; n01[u] = sq01[kN]; s01[u] = sq01[kS];
			m[u] = wmb ? wmb[k] : 31u;
		}
#pragma unroll
		for(int u = 0; u < NB; ++u) {
			const int q = q0 + u * 256;
			if(q >= S) break;
			const int qy = (int) __umulhi((unsigned) q, mlx), qx = q - qy * LX;
			const int xE = qx + 1 == LX ? 0 : qx + 1, xW = qx == 0 ? LX - 1 : qx - 1;
			const int yN = qy + 1 == LY ? 0 : qy + 1, yS = qy == 0 ? LY - 1 : qy - 1;
			const int qE = qy * LX + xE, qW = qy * LX + xW, qN = yN * LX + qx, qS = yS * LX + qx;
			// the central differences at q and its four neighbours: from LDS in a
			// wall box, else .5 (phi_E - phi_W), .5 (phi_N - phi_S) as in phase 1
			double dxq, dyq, dxE, dyE, dxW, dyW, dxN, dyN, dxS, dyS;
			if(walls) {
				dxq = s_dx[q]; dyq = s_dy[q];
				dxE = s_dx[qE]; dyE = s_dy[qE];
				dxW = s_dx[qW]; dyW = s_dy[qW];
				dxN = s_dx[qN]; dyN = s_dy[qN];
				dxS = s_dx[qS]; dyS = s_dy[qS];
			}
			else {
				const int xEE = xE + 1 == LX ? 0 : xE + 1, xWW = xW == 0 ? LX - 1 : xW - 1;
				const int yNN = yN + 1 == LY ? 0 : yN + 1, ySS = yS == 0 ? LY - 1 : yS - 1;
				auto d = [&](int x, int y) { return s_phi[y * LX + x]; };
				dxq = .5 * (d(xE, qy) - d(xW, qy));
				dyq = .5 * (d(qx, yN) - d(qx, yS));
				dxE = .5 * (d(xEE, qy) - d(qx, qy));
				dyE = .5 * (d(xE, yN) - d(xE, yS));
				dxW = .5 * (d(qx, qy) - d(xWW, qy));
				dyW = .5 * (d(xW, yN) - d(xW, yS));
				dxN = .5 * (d(xE, yN) - d(xW, yN));
				dyN = .5 * (d(qx, yNN) - d(qx, qy));
				dxS = .5 * (d(xE, yS) - d(xW, yS));
				dyS = .5 * (d(qx, qy) - d(qx, ySS));
			}
			const double f_passive_x = -0.5 * ((s_s[qE] + dxE * (gl * dxE)) - (s_s[qW] + dxW * (gl * dxW)))
					- 0.5 * ((dxN * (gl * dyN)) - (dxS * (gl * dyS)));
			const double f_passive_y = -0.5 * ((s_s[qN] + dyN * (gl * dyN)) - (s_s[qS] + dyS * (gl * dyS)))
					- 0.5 * ((dyE * (gl * dxE)) - (dyW * (gl * dxW)));
			const double dx = dxq, dy = dyq;
			const double fQ_self_x = -(C.Q00 * dx + C.Q01 * dy);
			const double fQ_self_y = -(C.Q01 * dx - C.Q00 * dy);
			const double fQ_inter_x = -(0.5 * (e00[u] - w00[u]) + 0.5 * (n01[u] - s01[u])) - fQ_self_x;
			const double fQ_inter_y = -(0.5 * (e01[u] - w01[u]) - 0.5 * (n00[u] - s00[u])) - fQ_self_y;
			double fx = f_passive_x * P.passive_alpha + fQ_self_x * P.zetaQ_self + fQ_inter_x * P.zetaQ_inter;
			double fy = f_passive_y * P.passive_alpha + fQ_self_y * P.zetaQ_self + fQ_inter_y * P.zetaQ_inter;
			if(a.wmask && !(m[u] & 16u)) fx = fy = 0.;   // wall site
			if(dry) a.vrows[C.start + q] = make_double2(fx / P.friction, fy / P.friction);
			else a.frows[C.start + q] = make_double2(fx, fy);
		}
	}
}

hipError_t launch_nem_force(const NemArgs &a, int ncells, size_t lds, hipStream_t s) {
	if(ncells <= 0) return hipSuccess;
	static size_t lds_set = 0;
	if(lds > lds_set) {
		hipError_t e = hipFuncSetAttribute((const void *) k_nem_force, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
		if(e != hipSuccess) return e;
		lds_set = lds;
	}
	RC_LAUNCH(k_nem_force, dim3(ncells), dim3(256), lds, s, a.cells, a.st, a.phi, ncells, a);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wet velocity solve: A v = f, (A v)_i = D(k_i) v_i - eta * N4(k_i), with
// Sv(k) = sum of v over the rows living at box site k and
// N4(k) = Sv(S) + Sv(W) + Sv(E) + Sv(N).  Per CG iteration:
//   k_cg_gather<P>  p = z + beta p (z = r / D), Sp(k), P2(k) = sum p^2 at k
//   k_cg_box<1>     N4(k);  p.Ap = sum_k D(k) P2(k) - eta Sp(k) N4(k)  -> alpha
//   k_cg_row<UPD>   x += alpha p, r -= alpha (D p - eta N4), r.r, r.z  -> beta
// Reductions: per-workgroup partials, summed in a fixed order.  alpha is
// re-derived by every k_cg_row workgroup from the p.Ap partials; beta and the
// convergence flags by every k_cg_gather<P_UPDATE> workgroup from the r.r,
// r.z partials, which only read the state: an extra workgroup of the next
// k_cg_box<1> writes that update (apply_update), after the gather has
// finished, and a one-workgroup k_cg_reduce<2> at the end of a batch of
// iterations.  All three sum in the same 256-thread order, so they agree bit
// for bit.
// ---------------------------------------------------------------------------
// D(k) of every box site of every box (the site's box found by its offset)
__global__ __launch_bounds__(256) void k_cg_diag(CGArgs a, int nsites) {
	const int kg = blockIdx.x * 256 + threadIdx.x;
	if(kg >= nsites) return;
	int b = 0;
	while(b + 1 < a.nboxes && a.boxes[b + 1].box_off <= kg) ++b;
	const DevBox B = a.boxes[b];
	const DevParams P = a.params[2 * b + a.set];
	const int k = kg - B.box_off;
	const int L = B.L, y = k / L, x = k - y * L;
	const int xe = x + 1 == L ? 0 : x + 1, xw = x == 0 ? L - 1 : x - 1;
	const int yn = y + 1 == L ? 0 : y + 1, ys = y == 0 ? L - 1 : y - 1;
	// triplet order of WetPolarModel.cpp:206-215: neighbours 1,3,5,7 = S,W,E,N
	const int nb[4] = { ys * L + x, y * L + xw, y * L + xe, yn * L + x };
	double d = P.friction;
	for(int j = 0; j < 4; ++j) {
		const int m = a.cover[B.box_off + nb[j]];
		for(int l = 0; l < m; ++l) d += P.friction_cell * 1.0;
	}
	a.diag[B.box_off + k] = d;
}

hipError_t launch_cg_diag(const CGArgs &a, int nsites, hipStream_t s) {
	RC_LAUNCH(k_cg_diag, dim3((nsites + 255) / 256), dim3(256), 0, s, a, nsites);
	return hipGetLastError();
}

// Block partials (valid in thread 0) -> part[j*nblk + blockIdx] (SoA); the
// next launch, k_cg_reduce, sums them in a fixed order (deterministic).  A
// last-workgroup ticket would save that launch, but its agent-scope release
// costs an L2 write-back per workgroup on this multi-XCD part (measured 2x
// slower per CG iteration).
template <int NV>
__device__ inline void block_partials(const double (&v)[NV], double *part) {
	if(threadIdx.x == 0) {
#pragma unroll
		for(int j = 0; j < NV; ++j) part[j * gridDim.x + blockIdx.x] = v[j];
	}
}

// a NaN in the wet solve (no counterpart: Eigen would iterate to its cap):
// halt the device; first_cell -1 names the solve rather than a cell
__device__ inline void cg_flag_nan(DevStatus *st) {
	if(!(atomicOr(&st->flags, ST_NAN) & ST_NAN)) st->first_cell = -1;
}

// scalar recurrences of Eigen's ConjugateGradient loop
__device__ void cg_after_resid(CGState *cg, const double *tot, double tol, int max_iter, DevStatus *st) {
	cg->max_iter = max_iter;
	for(int c = 0; c < 2; ++c) {
		const double bb = tot[c], rr = tot[2 + c], rz = tot[4 + c];
		if(!(bb == bb) || !(rr == rr) || !(rz == rz)) {
			cg_flag_nan(st);
			cg->done[c] = 1;
			continue;
		}
		cg->bb[c] = bb;
		cg->iters[c] = 0;
		cg->beta[c] = 0.;
		cg->alpha[c] = 0.;
		cg->zero_rhs[c] = 0;
		if(bb == 0) {
			cg->zero_rhs[c] = 1;
			cg->done[c] = 1;
			cg->rn2[c] = 0.;
			continue;
		}
		double thr = tol * tol * bb;
		if(thr < DBL_MIN) thr = DBL_MIN;
		cg->thr[c] = thr;
		cg->rn2[c] = rr;
		if(rr < thr) {
			cg->done[c] = 1;
			continue;
		}
		cg->done[c] = 0;
		cg->absNew[c] = rz;
	}
	cg->active = !(cg->done[0] && cg->done[1]);
}
// alpha0_c = (e.r)/(e.A e): the A-norm optimal step along the learned
// correction e (zero when e.A e is not positive or the solve is done)
__device__ void cg_after_step(CGState *cg, const double *er, const double *eae) {
	for(int c = 0; c < 2; ++c) cg->alpha0[c] = cg->active && eae[c] > 0 ? er[c] / eae[c] : 0.;
}
__device__ void cg_after_pap(CGState *cg, const double *tot) {
	for(int c = 0; c < 2; ++c)
		if(!cg->done[c]) cg->alpha[c] = cg->absNew[c] / tot[c];
}
// the state the update recurrence reads, taken before the sums are ready
struct CGUpdIn {
	int done[2], iters[2], max_iter;
	double thr[2], absNew[2];
};
__device__ inline CGUpdIn cg_update_inputs(const CGState *cg) {
	CGUpdIn u;
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		u.done[c] = cg->done[c];
		u.iters[c] = cg->iters[c];
		u.thr[c] = cg->thr[c];
		u.absNew[c] = cg->absNew[c];
	}
	u.max_iter = cg->max_iter;
	return u;
}
__device__ void cg_after_update(CGState *cg, const double *tot, const CGUpdIn &u, DevStatus *st) {
	int done[2] = { u.done[0], u.done[1] };
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		if(done[c]) continue;
		const int it = u.iters[c] + 1;
		cg->iters[c] = it;
		const double rn2 = tot[c];
		cg->rn2[c] = rn2;
		if(!(rn2 == rn2)) cg_flag_nan(st);   // later launches return at entry
		if(rn2 < u.thr[c] || it >= u.max_iter) {
			done[c] = 1;
			cg->done[c] = 1;
			continue;
		}
		const double absOld = u.absNew[c];
		cg->absNew[c] = tot[2 + c];
		cg->beta[c] = tot[2 + c] / absOld;
	}
	cg->active = !(done[0] && done[1]);
}
// beta and the convergence flags cg_after_update would set, without writing
// the state (the same expressions, so the same bits)
__device__ inline bool cg_update_peek(const CGUpdIn &u, const double *tot, bool (&done)[2], double (&beta)[2]) {
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		done[c] = u.done[c];
		beta[c] = 0.;
		if(done[c]) continue;
		const int it = u.iters[c] + 1;
		if(tot[c] < u.thr[c] || it >= u.max_iter) {
			done[c] = true;
			continue;
		}
		beta[c] = tot[2 + c] / u.absNew[c];
	}
	return !(done[0] && done[1]);
}

// the per-thread sums acc of a block of NT threads added up: each wave's lanes
// (wave_sum), then the waves in order; every thread gets tot
template <int NV, int NT>
__device__ inline void reduce_tail(double (&acc)[NV], double (&tot)[NV], double *s_red) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for(int j = 0; j < NV; ++j) acc[j] = wave_sum(acc[j]);
	if(lane == 0) {
#pragma unroll
		for(int j = 0; j < NV; ++j) s_red[wave * NV + j] = acc[j];
	}
	__syncthreads();
#pragma unroll
	for(int j = 0; j < NV; ++j) {
		double t = 0.;
#pragma unroll 4
		for(int w = 0; w < NT / 64; ++w) t += s_red[w * NV + j];
		tot[j] = t;
	}
	__syncthreads();
}

// sums SoA partials of the previous kernel(s) over a block of NT threads in a
// fixed order (deterministic, the same in every workgroup); every thread gets
// tot.  U partials per thread and value are requested at once: U changes
// only how the loads are grouped, not the order of the additions.
template <int NV, int NT = 1024, int U = 4>
__device__ inline void reduce_parts(const double *part, int nblk, double (&tot)[NV], double *s_red) {
	double acc[NV];
#pragma unroll
	for(int j = 0; j < NV; ++j) acc[j] = 0.;
	// Up to U partials per thread and value requested at once, then added in
	// order; a slot past nblk reads the last partial and adds +0 instead
	// (every sum unchanged).  32-bit byte offsets from the uniform base keep
	// one address register per load.
	const char *base = reinterpret_cast<const char *>(part);
	for(int b = threadIdx.x; b < nblk; b += U * NT) {
		double v[U][NV];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			const int bu = min(b + u * NT, nblk - 1);
#pragma unroll
			for(int j = 0; j < NV; ++j) {
				const unsigned off = (unsigned) (j * nblk + bu) * 8u;
				v[u][j] = *reinterpret_cast<const double *>(base + off);
			}
		}
#pragma unroll
		for(int u = 0; u < U; ++u)
#pragma unroll
			for(int j = 0; j < NV; ++j) acc[j] += b + u * NT < nblk ? v[u][j] : 0.;
	}
	reduce_tail<NV, NT>(acc, tot, s_red);
}

template <int WHAT>
__global__ __launch_bounds__(1024) void k_cg_reduce(CGArgs a, int nblk) {
	CGState *cg = a.cg;
	constexpr int NV = WHAT == 0 ? 6 : WHAT == 1 || WHAT == 3 ? 2 : 4;
	__shared__ double s_red[16 * 6];
	// The status word and the solve's flag only decide whether the sums are
	// used: they, and the state the recurrence reads, are requested with the
	// partials instead of ahead of them (partials a skipped launch left are
	// read and discarded).
	const unsigned flags = a.st->flags;
	const int active = cg->active;
	CGUpdIn upd;
	if(WHAT == 2 && threadIdx.x == 0) upd = cg_update_inputs(cg);
	double tot[NV];
	// WHAT 2 runs on 256 threads: the order of k_cg_gather<P_UPDATE> and k_cg_box<1>
	if(WHAT == 2) reduce_parts<NV, 256>(a.part, nblk, tot, s_red);
	else reduce_parts<NV>(WHAT == 1 ? a.part_box : a.part, nblk, tot, s_red);
	if(flags) return;
	// (a learned stage after convergence leaves the state as it is: WHAT 3, 4)
	if((WHAT == 1 || WHAT == 2 || WHAT == 3 || WHAT == 4) && !active) return;
	if(WHAT == 3) {
		double eae[2];
		reduce_parts<2>(a.part_box, a.nbox_blk, eae, s_red);
		if(threadIdx.x == 0) cg_after_step(cg, tot, eae);
		return;
	}
	if(threadIdx.x != 0) return;
	if(WHAT == 5) {
		// after the fallback pass: its partials are valid only if a component reverted
		const int r0 = cg->rn2[0] > cg->rn2_start[0], r1 = cg->rn2[1] > cg->rn2_start[1];
		cg->reverted[0] = r0;
		cg->reverted[1] = r1;
		if(!r0 && !r1) return;
		const double t6[6] = { cg->bb[0], cg->bb[1], tot[0], tot[1], tot[2], tot[3] };
		cg_after_resid(cg, t6, a.tol, a.max_iter, a.st);
		return;
	}
	if(WHAT == 0) cg_after_resid(cg, tot, a.tol, a.max_iter, a.st);
	else if(WHAT == 1) cg_after_pap(cg, tot);
	else if(WHAT == 4) {
		// residual of the new start; |b|^2 is that of the solve's first residual pass
		const double t6[6] = { cg->bb[0], cg->bb[1], tot[0], tot[1], tot[2], tot[3] };
		cg_after_resid(cg, t6, a.tol, a.max_iter, a.st);
	}
	else cg_after_update(cg, tot, upd, a.st);
}

hipError_t launch_cg_reduce(int what, const CGArgs &a, int nblk, hipStream_t s) {
	switch(what) {
	case 0: RC_LAUNCH(k_cg_reduce<0>, dim3(1), dim3(1024), 0, s, a, nblk); break;
	case 1: RC_LAUNCH(k_cg_reduce<1>, dim3(1), dim3(1024), 0, s, a, nblk); break;
	case 3: RC_LAUNCH(k_cg_reduce<3>, dim3(1), dim3(1024), 0, s, a, nblk); break;
	case 4: RC_LAUNCH(k_cg_reduce<4>, dim3(1), dim3(1024), 0, s, a, nblk); break;
	case 5: RC_LAUNCH(k_cg_reduce<5>, dim3(1), dim3(1024), 0, s, a, nblk); break;
	default: RC_LAUNCH(k_cg_reduce<2>, dim3(1), dim3(256), 0, s, a, nblk); break;
	}
	return hipGetLastError();
}

// residual scale of component c: sqrt(|r_c|^2 / rows) (zero for a zero residual)
__device__ inline double resid_scale(const CGState *cg, int c, double rows) {
	const double r2 = cg->rn2[c];
	return r2 > 0 ? sqrt(r2 / rows) : 0.;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_cg_gather(CGArgs a) {
	if(a.st->flags) return;
	const CGState *cg = a.cg;
	if(MODE != CG_GATHER_X && MODE != CG_GATHER_BX && !cg->active) return;
	const int tile = xcd_item(blockIdx.x, gridDim.x), tid = threadIdx.x;
	const DevBox B = a.boxes[a.tile_box[tile]];
	const int n = list_count(a.tile_count, tile);
	const int t = tile - B.tile_off, L = B.L;
	const int x0 = (t % B.ntx) * TW, y0 = (t / B.ntx) * TH;
	const int tw = min(TW, L - x0), th = min(TH, L - y0);
	const int sx = tid % TW, sy = tid / TW;
	if(MODE == CG_GATHER_E) {
		// e = (r + eta N4) / D per covering row (A e = r row by row); box sums of e, e^2; e.r -> partials
		__shared__ double s_red[4 * 2];
		double er[2] = { 0., 0. };
		if(sx < tw && sy < th) {
			const int k = B.box_off + (y0 + sy) * L + x0 + sx;
			const double D = a.diag[k], eta = a.params[a.set].friction_cell;
			const double2 n4 = a.n4[k];
			double2 acc = make_double2(0., 0.), acc2 = make_double2(0., 0.);
			// batches of GCHUNK covering rows (the lists are padded to 16 with
			// entries that cover nothing), every load issued before the first use
			for(int base = 0; base < n; base += GCHUNK) {
				int row[GCHUNK];
				double2 rv[GCHUNK];
#pragma unroll
				for(int q = 0; q < GCHUNK; ++q) row[q] = entry_row(list_entry(a.tile_geo, tile, base + q), sx, sy, L);
#pragma unroll
				for(int q = 0; q < GCHUNK; ++q) rv[q] = row[q] >= 0 ? a.r[row[q]] : make_double2(0., 0.);
				__builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), before the stores below
#pragma unroll
				for(int q = 0; q < GCHUNK; ++q) {
					if(row[q] < 0) continue;
					const double2 r = rv[q];
					const double2 e = make_double2((r.x + eta * n4.x) / D, (r.y + eta * n4.y) / D);
					a.p[row[q]] = e;
					acc.x += e.x; acc.y += e.y;
					acc2.x += e.x * e.x; acc2.y += e.y * e.y;
					er[0] += e.x * r.x; er[1] += e.y * r.y;
				}
			}
			a.sp[k] = acc;
			a.p2[k] = acc2;
		}
		block_sum<2>(er, s_red);
		block_partials<2>(er, a.part);
		return;
	}
	bool dn[2] = { false, false };
	double bt[2] = { 0., 0. };
	if(MODE == CG_GATHER_P_UPDATE) {
		// beta and the convergence flags of this iteration from the r.r, r.z
		// partials of k_cg_row<UPDATE> (the state itself is written later)
		__shared__ double s_upd[4 * 4];
		const CGUpdIn u = cg_update_inputs(cg);
		double tot[4];
		reduce_parts<4, 256, 8>(a.part, a.ncells, tot, s_upd);
		if(!cg_update_peek(u, tot, dn, bt)) return;
	}
	if(sx >= tw || sy >= th) return;
	const int x = x0 + sx, y = y0 + sy;
	const int k = B.box_off + y * L + x;
	double inv = 0.;
	const bool d0 = dn[0], d1 = dn[1];
	const double b0 = bt[0], b1 = bt[1];
	if(MODE != CG_GATHER_X && MODE != CG_GATHER_BX) inv = 1. / a.diag[k];   // Jacobi: z = invdiag .* r
	double2 acc = make_double2(0., 0.), acc2 = make_double2(0., 0.);
	for(int base = 0; base < n; base += GCHUNK) {
		int row[GCHUNK];
		double2 rv[GCHUNK], pv[GCHUNK];
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) row[q] = entry_row(list_entry(a.tile_geo, tile, base + q), sx, sy, L);
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) {
			rv[q] = make_double2(0., 0.);
			pv[q] = make_double2(0., 0.);
			if(row[q] >= 0) {
				if(MODE == CG_GATHER_X) rv[q] = a.x[row[q]];
				else if(MODE == CG_GATHER_FEAT) rv[q] = a.r[row[q]];
				else if(MODE == CG_GATHER_BX) {
					rv[q] = a.b[row[q]];
					pv[q] = a.x[row[q]];
				}
				else {
					rv[q] = a.r[row[q]];
					if(MODE == CG_GATHER_P_UPDATE) pv[q] = a.p[row[q]];
				}
			}
		}
		// one wait for the batch: the memory counter also counts the stores
		// below, which would otherwise put a wait behind each of them
		__builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
#pragma unroll
		for(int q = 0; q < GCHUNK; ++q) {
			if(row[q] < 0) continue;
			double2 v;
			if(MODE == CG_GATHER_BX) {
				v = rv[q];
				acc2.x += pv[q].x;
				acc2.y += pv[q].y;
			}
			else if(MODE == CG_GATHER_X || MODE == CG_GATHER_FEAT) v = rv[q];
			else {
				if(MODE == CG_GATHER_P_FIRST) v = make_double2(inv * rv[q].x, inv * rv[q].y);
				else {
					v.x = d0 ? pv[q].x : inv * rv[q].x + b0 * pv[q].x;
					v.y = d1 ? pv[q].y : inv * rv[q].y + b1 * pv[q].y;
				}
				a.p[row[q]] = v;
				acc2.x += v.x * v.x;
				acc2.y += v.y * v.y;
			}
			acc.x += v.x;
			acc.y += v.y;
		}
	}
	if(MODE == CG_GATHER_FEAT) {
		// network input: residual sums over the residual scale, sum phi^2 (bf16 channels 0..7)
		const double n0 = resid_scale(cg, 0, a.rows), n1 = resid_scale(cg, 1, a.rows);
		const float f0 = n0 > 0 ? (float) (acc.x / n0) * a.sf : 0.f;
		const float f1 = n1 > 0 ? (float) (acc.y / n1) * a.sf : 0.f;
		const float f2 = (float) a.phi2[k] * a.sp2;
		uint4 o;
		o.x = (unsigned) f2bf(f0) | ((unsigned) f2bf(f1) << 16);
		o.y = (unsigned) f2bf(f2);
		o.z = o.w = 0u;
		*reinterpret_cast<uint4 *>(a.act0 + (size_t) (k - B.box_off) * 32) = o;
		return;
	}
	a.sp[k] = acc;
	if(MODE != CG_GATHER_X) a.p2[k] = acc2;   // BX: the box sums of x
}

hipError_t launch_cg_gather(int mode, const CGArgs &a, int ntiles, hipStream_t s) {
	switch(mode) {
	case CG_GATHER_X: RC_LAUNCH(k_cg_gather<CG_GATHER_X>, dim3(ntiles), dim3(256), 0, s, a); break;
	case CG_GATHER_P_FIRST: RC_LAUNCH(k_cg_gather<CG_GATHER_P_FIRST>, dim3(ntiles), dim3(256), 0, s, a); break;
	case CG_GATHER_FEAT: RC_LAUNCH(k_cg_gather<CG_GATHER_FEAT>, dim3(ntiles), dim3(256), 0, s, a); break;
	case CG_GATHER_E: RC_LAUNCH(k_cg_gather<CG_GATHER_E>, dim3(ntiles), dim3(256), 0, s, a); break;
	case CG_GATHER_BX: RC_LAUNCH(k_cg_gather<CG_GATHER_BX>, dim3(ntiles), dim3(256), 0, s, a); break;
	default: RC_LAUNCH(k_cg_gather<CG_GATHER_P_UPDATE>, dim3(ntiles), dim3(256), 0, s, a); break;
	}
	return hipGetLastError();
}

// N4(k) and, for PAP, p.Ap reduced to alpha (Eigen: alpha = absNew / p.dot(A p))
// MODE 0: N4 of sp; 1: N4 and the p.Ap partials; 2: N4 of the network output
// sp = out * sv * (residual scale) of a learned stage
template <int MODE>
__global__ __launch_bounds__(256) void k_cg_box(CGArgs a, int nsites) {
	if(a.st->flags) return;
	CGState *cg = a.cg;
	if(MODE != 0 && !cg->active) return;
	__shared__ double s_red[4 * 2];
	const int nblk = (nsites + 255) / 256;
	if(MODE == 1 && blockIdx.x == nblk) {
		// one extra workgroup (launched with apply_update only): the previous
		// iteration's update of the CG state, whose beta and flags
		// k_cg_gather<P_UPDATE> derived from the same partials.  Beside the box
		// workgroups rather than ahead of one of them, so the pass's duration is
		// not the sum of the two chains.
		__shared__ double s_upd[4 * 4];
		const CGUpdIn u = cg_update_inputs(cg);
		double tot[4];
		reduce_parts<4, 256, 8>(a.part, a.ncells, tot, s_upd);
		if(threadIdx.x == 0) cg_after_update(cg, tot, u, a.st);
		return;
	}
	const int k = xcd_item(blockIdx.x, nblk) * 256 + threadIdx.x;
	const DevBox B = a.boxes[0];
	const double eta = a.params[a.set].friction_cell;
	double part[2] = { 0., 0. };
	if(k < nsites) {
		const int L = B.L, y = k / L, x = k - y * L;
		const int xe = x + 1 == L ? 0 : x + 1, xw = x == 0 ? L - 1 : x - 1;
		const int yn = y + 1 == L ? 0 : y + 1, ys = y == 0 ? L - 1 : y - 1;
		const int o = B.box_off;
		double2 n4;
		if(MODE == 2) {
			const float2 *out = reinterpret_cast<const float2 *>(a.net_out);
			const float2 sS = out[ys * L + x], sW = out[y * L + xw], sE = out[y * L + xe], sN = out[yn * L + x];
			const double f0 = a.sv * resid_scale(cg, 0, a.rows), f1 = a.sv * resid_scale(cg, 1, a.rows);
			n4 = make_double2(f0 * ((double) sS.x + (double) sW.x + (double) sE.x + (double) sN.x),
					f1 * ((double) sS.y + (double) sW.y + (double) sE.y + (double) sN.y));
		}
		else {
			const double2 sS = a.sp[o + ys * L + x], sW = a.sp[o + y * L + xw], sE = a.sp[o + y * L + xe],
					sN = a.sp[o + yn * L + x];
			n4 = make_double2(sS.x + sW.x + sE.x + sN.x, sS.y + sW.y + sE.y + sN.y);
		}
		a.n4[o + k] = n4;
		if(MODE == 1) {
			const double D = a.diag[o + k];
			const double2 sp = a.sp[o + k], p2 = a.p2[o + k];
			part[0] = D * p2.x - eta * sp.x * n4.x;
			part[1] = D * p2.y - eta * sp.y * n4.y;
		}
	}
	if(MODE != 1) return;
	block_sum<2>(part, s_red);
	// stride nblk: the update workgroup is not a box workgroup
	if(threadIdx.x == 0) {
		a.part_box[blockIdx.x] = part[0];
		a.part_box[nblk + blockIdx.x] = part[1];
	}
}

hipError_t launch_cg_box(int mode, const CGArgs &a, int nsites, hipStream_t s) {
	const int nblk = (nsites + 255) / 256;
	const dim3 g(nblk);
	if(mode == 1) RC_LAUNCH(k_cg_box<1>, dim3(nblk + (a.apply_update ? 1 : 0)), dim3(256), 0, s, a, nsites);
	else if(mode == 2) RC_LAUNCH(k_cg_box<2>, g, dim3(256), 0, s, a, nsites);
	else RC_LAUNCH(k_cg_box<0>, g, dim3(256), 0, s, a, nsites);
	return hipGetLastError();
}

template <int MODE>
__global__ __launch_bounds__(256) void k_cg_row(CGArgs a) {
	if(a.st->flags) return;
	CGState *cg = a.cg;
	// a learned stage after the solve has converged is a no-op (alpha0 = 0): skipped
	if((MODE == CG_ROW_UPDATE || MODE == CG_ROW_STEP) && !cg->active) return;
	// fallback: a component whose learned start has a larger residual than the previous solution
	const bool rev0 = MODE == CG_ROW_REVERT && cg->rn2[0] > cg->rn2_start[0];
	const bool rev1 = MODE == CG_ROW_REVERT && cg->rn2[1] > cg->rn2_start[1];
	if(MODE == CG_ROW_REVERT && !rev0 && !rev1) return;
	__shared__ double s_red[4 * 6];
	const int ci = xcd_item(blockIdx.x, gridDim.x), tid = threadIdx.x;
	const DevCell C = a.cells[ci];
	const DevBox B = a.boxes[C.box];
	const double eta = a.params[a.set].friction_cell;
	const int LX = C.LX, LY = C.LY, S = C.S, L = B.L;
	const int cx = C.corner % L, cy = C.corner / L;
	double al0 = 0., al1 = 0.;
	bool dn0 = false, dn1 = false;
	if(MODE == CG_ROW_UPDATE) {
		// alpha = absNew / p.Ap (cg_after_pap), every workgroup summing the box
		// pass's partials in the same fixed order: no separate reduction launch
		double pap[2];
		reduce_parts<2, 256>(a.part_box, a.nbox_blk, pap, s_red);
		dn0 = cg->done[0]; dn1 = cg->done[1];
		if(!dn0) al0 = cg->absNew[0] / pap[0];
		if(!dn1) al1 = cg->absNew[1] / pap[1];
		if(ci == 0 && tid == 0) {
			if(!dn0) cg->alpha[0] = al0;
			if(!dn1) cg->alpha[1] = al1;
		}
	}
	double red[6] = { 0., 0., 0., 0., 0., 0. };
	auto box_site = [&](int q) {
		const int qy = q / LX, qx = q - qy * LX;
		int bx = qx + C.offx;
		if(bx >= LX) bx -= LX;
		bx += cx;
		if(bx >= L) bx -= L;
		int by = qy + C.offy;
		if(by >= LY) by -= LY;
		by += cy;
		if(by >= L) by -= L;
		return B.box_off + by * L + bx;
	};
	if(MODE == CG_ROW_STEP) {
		// the learned stage's step along e (held in p): alpha0 per component
		al0 = cg->alpha0[0];
		al1 = cg->alpha0[1];
	}
	if constexpr(MODE == CG_ROW_UPDATE || MODE == CG_ROW_STEP) {
		// x += alpha p, r -= alpha Ap on two rows per pass, every load of both
		// issued before either row is updated (rows q, q + 256, ... in order)
		auto upd = [&](double D, double2 n4, double2 vin, double2 xx, double2 rr, int row) {
			double2 t;
			t.x = D * vin.x - eta * n4.x;
			t.y = D * vin.y - eta * n4.y;
			const double inv = 1. / D;
			if(!dn0) { xx.x += al0 * vin.x; rr.x -= al0 * t.x; }
			if(!dn1) { xx.y += al1 * vin.y; rr.y -= al1 * t.y; }
			a.x[row] = xx;
			a.r[row] = rr;
			red[0] += rr.x * rr.x; red[1] += rr.y * rr.y;
			red[2] += rr.x * (inv * rr.x); red[3] += rr.y * (inv * rr.y);
		};
		for(int q0 = tid; q0 < S; q0 += 512) {
			const bool two = q0 + 256 < S;
			const int q1 = two ? q0 + 256 : q0;   // past the end: row q0 again, read and not written
			const int k0 = box_site(q0), k1 = box_site(q1);
			const int r0 = C.start + q0, r1 = C.start + q1;
			const double D0 = a.diag[k0], D1 = a.diag[k1];
			const double2 n40 = a.n4[k0], n41 = a.n4[k1];
			const double2 p0 = a.p[r0], p1 = a.p[r1];
			const double2 x0 = a.x[r0], x1 = a.x[r1];
			const double2 rr0 = a.r[r0], rr1 = a.r[r1];
			upd(D0, n40, p0, x0, rr0, r0);
			if(two) upd(D1, n41, p1, x1, rr1, r1);
		}
	}
	else for(int q = tid; q < S; q += 256) {
		const int qy = q / LX, qx = q - qy * LX;
		int bx = qx + C.offx;
		if(bx >= LX) bx -= LX;
		bx += cx;
		if(bx >= L) bx -= L;
		int by = qy + C.offy;
		if(by >= LY) by -= LY;
		by += cy;
		if(by >= L) by -= L;
		const int k = B.box_off + by * L + bx;
		const double D = a.diag[k];
		const double2 n4 = a.n4[k];
		const int row = C.start + q;
		if(MODE == CG_ROW_REVERT) {
			// back to the previous solution (and its residual) in the components that got worse
			const double inv = 1. / D;
			double2 xx = a.x[row], rn = a.r[row];
			if(rev0) { xx.x = a.xprev[row].x; rn.x = a.rprev[row].x; }
			if(rev1) { xx.y = a.xprev[row].y; rn.y = a.rprev[row].y; }
			a.x[row] = xx;
			a.r[row] = rn;
			red[0] += rn.x * rn.x; red[1] += rn.y * rn.y;
			red[2] += rn.x * (inv * rn.x); red[3] += rn.y * (inv * rn.y);
			continue;
		}
		const double2 vin = MODE == CG_ROW_RESID ? a.x[row] : a.p[row];
		double2 t;
		t.x = D * vin.x - eta * n4.x;
		t.y = D * vin.y - eta * n4.y;
		const double inv = 1. / D;
		if(MODE == CG_ROW_RESID) {
			const double2 bb = a.b[row];
			const double2 rr = make_double2(bb.x - t.x, bb.y - t.y);
			a.r[row] = rr;
			red[0] += bb.x * bb.x; red[1] += bb.y * bb.y;
			red[2] += rr.x * rr.x; red[3] += rr.y * rr.y;
			red[4] += rr.x * (inv * rr.x); red[5] += rr.y * (inv * rr.y);
		}
		else {
			double2 xx = a.x[row], rr = a.r[row];
			if(!dn0) { xx.x += al0 * vin.x; rr.x -= al0 * t.x; }
			if(!dn1) { xx.y += al1 * vin.y; rr.y -= al1 * t.y; }
			a.x[row] = xx;
			a.r[row] = rr;
			red[0] += rr.x * rr.x; red[1] += rr.y * rr.y;
			red[2] += rr.x * (inv * rr.x); red[3] += rr.y * (inv * rr.y);
		}
	}
	if(MODE == CG_ROW_RESID) {
		block_sum<6>(red, s_red);
		block_partials<6>(red, a.part);
	}
	else {
		double r4[4] = { red[0], red[1], red[2], red[3] };
		block_sum<4>(r4, s_red);
		block_partials<4>(r4, a.part);
	}
}

hipError_t launch_cg_row(int mode, const CGArgs &a, hipStream_t s) {
	if(mode == CG_ROW_RESID) RC_LAUNCH(k_cg_row<CG_ROW_RESID>, dim3(a.ncells), dim3(256), 0, s, a);
	else if(mode == CG_ROW_STEP) RC_LAUNCH(k_cg_row<CG_ROW_STEP>, dim3(a.ncells), dim3(256), 0, s, a);
	else if(mode == CG_ROW_REVERT) RC_LAUNCH(k_cg_row<CG_ROW_REVERT>, dim3(a.ncells), dim3(256), 0, s, a);
	else RC_LAUNCH(k_cg_row<CG_ROW_UPDATE>, dim3(a.ncells), dim3(256), 0, s, a);
	return hipGetLastError();
}

// snapshot channels: chan 0 -> sp into [k][0..1], chan 2 -> phi2 into [k][2],
// chan 3 -> sp into [k][3..4], chan 5 -> [k][3..4] -= sp; with rn2 given, sp_c
// is divided by sqrt(rn2_c / rows) (zero when rn2_c is zero)
__global__ void k_box_snapshot(CGArgs a, const double *phi2, float *snap, int chan, const double *rn2, double rows,
		int nsites) {
	const int k = blockIdx.x * 256 + threadIdx.x;
	if(k >= nsites) return;
	if(chan == 2) {
		snap[5 * k + 2] = (float) phi2[k];
		return;
	}
	double2 v = a.sp[k];
	if(rn2) {
		v.x = rn2[0] > 0 ? v.x / sqrt(rn2[0] / rows) : 0.;
		v.y = rn2[1] > 0 ? v.y / sqrt(rn2[1] / rows) : 0.;
	}
	if(chan == 5) {
		snap[5 * k + 3] -= (float) v.x;
		snap[5 * k + 4] -= (float) v.y;
	}
	else {
		snap[5 * k + chan] = (float) v.x;
		snap[5 * k + chan + 1] = (float) v.y;
	}
}
hipError_t launch_box_snapshot(const CGArgs &a, const double *phi2, float *snap, int chan, const double *rn2, double rows,
		int nsites, hipStream_t s) {
	RC_LAUNCH(k_box_snapshot, dim3((nsites + 255) / 256), dim3(256), 0, s, a, phi2, snap, chan, rn2, rows, nsites);
	return hipGetLastError();
}

// Row repack after a re-layout: for listed cell j, copy its map[j].z rows
// from src + map[j].x to dst + map[j].y (and the same for src2 / dst2 when set).
__global__ __launch_bounds__(256) void k_repack_rows(const double *__restrict__ src, double *__restrict__ dst,
		const double *__restrict__ src2, double *__restrict__ dst2, const int4 *__restrict__ map, int n) {
	const int j = blockIdx.x;
	if(j >= n) return;
	const int4 m = map[j];
	for(int q = threadIdx.x; q < m.z; q += 256) {
		dst[m.y + q] = src[m.x + q];
		if(src2) dst2[m.y + q] = src2[m.x + q];
	}
}
hipError_t launch_repack_rows(const double *src, double *dst, const double *src2, double *dst2, const int4 *map, int n,
		hipStream_t s) {
	if(n <= 0) return hipSuccess;
	RC_LAUNCH(k_repack_rows, dim3(n), dim3(256), 0, s, src, dst, src2, dst2, map, n);
	return hipGetLastError();
}

// The re-cut of a growing patch (MultiPhaseField::set_positions, growth
// branch, MultiPhaseField.cpp:248-310) on the device, one workgroup per
// growing cell.  ent[4 j .. 4 j + 3] = {old start, new start, LX, LY}, {offx,
// offy, corner, box side}, {new LX, new LY, new corner, sx}, {sy, dimx, dimy,
// 0}: the host's geometry of the re-cut.  Every old site (i, j) of rows sy ..
// dimy - 1 and columns sx .. dimx - 1 goes to new site |dx| + |dy| new LX
// (dx, dy: the old site's offset from the new corner, nearest image); where
// the reference's loop writes one new site twice the later (i, j) in its
// order (rows, then columns) wins: the largest key i LX + j, taken with an
// LDS atomicMax.  Sites nothing maps to are 0.  area_out[j] = sum of phi^2
// over the copied sites (the reference adds every copy, overwritten or not),
// in a fixed order, or NaN if a site maps outside the new patch.  The
// Repulsion rows (src2, when set) keep their first min(old, new) entries
// (std::vector::resize).
__global__ __launch_bounds__(256) void k_grow_rows(const double *__restrict__ src, double *__restrict__ dst,
		const double *__restrict__ src2, double *__restrict__ dst2, const int4 *__restrict__ ent, double *area_out) {
	extern __shared__ int s_key[];
	__shared__ double s_part[4];
	__shared__ int s_bad;
	const int tid = threadIdx.x;
	const int4 e0 = ent[4 * blockIdx.x], e1 = ent[4 * blockIdx.x + 1], e2 = ent[4 * blockIdx.x + 2], e3 = ent[4 * blockIdx.x + 3];
	const int ostart = e0.x, nstart = e0.y, LX = e0.z, LY = e0.w;
	const int offx = e1.x, offy = e1.y, corner = e1.z, side = e1.w;
	const int nLX = e2.x, nLY = e2.y, ncorner = e2.z, sx = e2.w;
	const int sy = e3.x, dimx = e3.y, dimy = e3.z;
	const int nS = nLX * nLY, S = LX * LY;
	for(int q = tid; q < nS; q += 256) s_key[q] = -1;
	if(tid == 0) s_bad = 0;
	__syncthreads();
	const int ncx = ncorner - (ncorner / side) * side, ncy = ncorner / side;
	const int w = dimx - sx, hg = dimy - sy;
	const int cnt = w > 0 && hg > 0 ? w * hg : 0;
	auto source = [&](int i, int j) {
		int off_y = i - offy;
		if(off_y < 0) off_y += LY;
		int off_x = j - offx;
		if(off_x < 0) off_x += LX;
		return ostart + off_x + off_y * LX;
	};
	double acc = 0.;
	for(int t = tid; t < cnt; t += 256) {
		const int i = sy + t / w, j = sx + t % w;
		const int yy = box_y(side, corner, i), xx = box_x(side, corner, j);
		int dx = ncx - xx;
		if(dx > side / 2) dx -= side;
		else if(dx < -side / 2) dx += side;
		int dy = ncy - yy;
		if(dy > side / 2) dy -= side;
		else if(dy < -side / 2) dy += side;
		const int d = abs(dx) + abs(dy) * nLX;
		if(d < 0 || d >= nS) s_bad = 1;
		else atomicMax(&s_key[d], i * LX + j);
		const double v = src[source(i, j)];
		acc += v * v;
	}
	__syncthreads();
	for(int q = tid; q < nS; q += 256) {
		const int k = s_key[q];
		dst[nstart + q] = k >= 0 ? src[source(k / LX, k % LX)] : 0.;
		if(src2) dst2[nstart + q] = q < S ? src2[ostart + q] : 0.;
	}
	const double ws = wave_sum(acc);
	if((tid & 63) == 0) s_part[tid >> 6] = ws;
	__syncthreads();
	if(tid == 0) area_out[blockIdx.x] = s_bad ? NAN : (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]);
}
hipError_t launch_grow_rows(const double *src, double *dst, const double *src2, double *dst2, const int4 *ent, int n,
		int max_sites, double *area_out, hipStream_t s) {
	if(n <= 0) return hipSuccess;
	const size_t lds = sizeof(int) * (size_t) max_sites;
	if(lds > GROW_LDS_MAX) return hipErrorInvalidValue;
	static size_t lds_set = 0;
	if(lds > lds_set) {
		hipError_t e = hipFuncSetAttribute((const void *) k_grow_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
		if(e != hipSuccess) return e;
		lds_set = lds;
	}
	RC_LAUNCH(k_grow_rows, dim3(n), dim3(256), lds, s, src, dst, src2, dst2, ent, area_out);
	return hipGetLastError();
}

__global__ void k_zero_rows(double2 *x, long long rows, int mask) {
	const long long i = (long long) blockIdx.x * 256 + threadIdx.x;
	if(i >= rows) return;
	double2 v = x[i];
	if(mask & 1) v.x = 0.;
	if(mask & 2) v.y = 0.;
	x[i] = v;
}

hipError_t launch_zero_rows(double2 *x, long long rows, int mask, hipStream_t s) {
	if(rows <= 0) return hipSuccess;
	RC_LAUNCH(k_zero_rows, dim3((unsigned) ((rows + 255) / 256)), dim3(256), 0, s, x, rows, mask);
	return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Site-space CG (SCGState in rc_device.h): Eigen's Jacobi-PCG iterates of the
// wet solve carried as scalars times the ker-P parts r_perp, z_perp = r_perp /
// D of the first residual plus L x L fields, in Chronopoulos and Gear's
// arrangement (one reduction per iteration), for every wet box of a handle at
// once (each box with its own scalars).  The iterates are the reference's in
// exact arithmetic; an iteration is one pass over the box sites instead of
// passes over every row.
// ---------------------------------------------------------------------------

// the four box neighbours of site k, in the triplet order of
// WetPolarModel.cpp:206-215 (neighbours 1, 3, 5, 7 = S, W, E, N)
__device__ inline void site_nbrs(int k, int L, int *nb) {
	const int y = k / L, x = k - y * L;
	const int xe = x + 1 == L ? 0 : x + 1, xw = x == 0 ? L - 1 : x - 1;
	const int yn = y + 1 == L ? 0 : y + 1, ys = y == 0 ? L - 1 : y - 1;
	nb[0] = ys * L + x;
	nb[1] = y * L + xw;
	nb[2] = y * L + xe;
	nb[3] = yn * L + x;
}

// the partial sums of one site: y.m y, y.m (y / D), (y / D).m w (x, y components)
__device__ inline void scg_part_i(double (&part)[6], int m, double inv, double2 y, double2 w) {
	const double fm = m;
	const double zx = inv * y.x, zy = inv * y.y;
	part[0] = fm * (y.x * y.x);
	part[1] = fm * (y.y * y.y);
	part[2] = fm * (y.x * zx);
	part[3] = fm * (y.y * zy);
	part[4] = fm * (zx * w.x);
	part[5] = fm * (zy * w.y);
}
__device__ inline void scg_part(double (&part)[6], int m, double D, double2 y, double2 w) {
	scg_part_i(part, m, 1. / D, y, w);
}
// The sums of one box's partials (6 values per item, items begin .. begin +
// count - 1 of a [6][stride] array) over NT threads in one fixed order, split
// so that a launch requests this thread's first partials with its other loads
// and tests the status word before using any: parts_load, then parts_sum.
template <int NT>
__device__ inline void parts_load(const double *part, int stride, int begin, int count, double (&v)[6]) {
	const int b = begin + min((int) threadIdx.x, count - 1);
#pragma unroll
	for(int j = 0; j < 6; ++j) v[j] = part[j * stride + b];
}
template <int NT>
__device__ inline void parts_sum(const double *part, int stride, int begin, int count, const double (&v)[6],
		double (&tot)[6], double *s_red) {
	double acc[6];
	const bool own = (int) threadIdx.x < count;
#pragma unroll
	for(int j = 0; j < 6; ++j) {
		acc[j] = 0.;
		acc[j] += own ? v[j] : 0.;
	}
	for(int b = threadIdx.x + NT; b < count; b += NT)
#pragma unroll
		for(int j = 0; j < 6; ++j) acc[j] += part[j * stride + begin + b];
	reduce_tail<6, NT>(acc, tot, s_red);
}

// The site form's sums, in one fixed order wherever they are formed:
// a tile's six partials (scg_tile_partials) and a box's sums of its tiles'
// partials (scg_sums_load, scg_sums_add; the same in every workgroup of the box
// and in k_scg_check).  Wave w < 6 of the workgroup adds value w: each lane its
// items lane, lane + 64, ... in order, then the lanes (wave_sum); the other
// waves idle.  One value per wave keeps the cross-lane work to six wave sums
// per workgroup.
// items per lane requested with a launch's first loads: 8 covers the 512
// tiles of a box up to 1024 x 512, 16 the 1024 of a 1024^2 box (the host picks
// by the largest box; the order of the additions is the same either way)
constexpr int SCG_PU = 8, SCG_PU_BIG = 16;
static_assert(SCG_THREADS >= 6 * 64, "one wave per value");
// a tile's partials from each thread's six values (s_part: [6][SCG_THREADS]
// doubles of LDS) -> part[j * ntiles + tile]
__device__ inline void scg_tile_partials(const double (&v)[6], double *s_part, double *part, int ntiles, int tile) {
#pragma unroll
	for(int j = 0; j < 6; ++j) s_part[j * SCG_THREADS + threadIdx.x] = v[j];
	__syncthreads();
	const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
	if(w < 6) {
		double acc = 0.;
#pragma unroll
		for(int i = 0; i < SCG_THREADS / 64; ++i) acc += s_part[w * SCG_THREADS + lane + 64 * i];
		acc = wave_sum(acc);
		if(lane == 0) part[w * ntiles + tile] = acc;
	}
}
// a box's partials: items begin .. begin + count - 1 of a [6][stride] array.
// The first PU items of the lane are requested by scg_sums_load (waves 0-5
// only; items past the end read a valid item that is not added), the rest by
// scg_sums_add, which gives every thread the six sums (s_red: 6 doubles of
// LDS).
template <int PU>
__device__ inline void scg_sums_load(const double *part, int stride, int begin, int count, double (&v)[PU]) {
	const int w = (int) threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
	for(int i = 0; i < PU; ++i) v[i] = 0.;
	if(w < 6) {
#pragma unroll
		for(int i = 0; i < PU; ++i) v[i] = part[w * stride + begin + min(lane + 64 * i, count - 1)];
	}
}
template <int PU>
__device__ inline void scg_sums_add(const double *part, int stride, int begin, int count, const double (&v)[PU],
		double (&tot)[6], double *s_red) {
	const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
	if(w < 6) {
		double acc = 0.;
#pragma unroll
		for(int i = 0; i < PU; ++i)
			if(lane + 64 * i < count) acc += v[i];
		for(int it = lane + 64 * PU; it < count; it += 64) acc += part[w * stride + begin + it];
		acc = wave_sum(acc);
		if(lane == 0) s_red[w] = acc;
	}
	__syncthreads();
#pragma unroll
	for(int j = 0; j < 6; ++j) tot[j] = s_red[j];
}

// The tile of a site pass: workgroup b takes tile xcd_item(b) of the handle's
// tiles (box-major), a tile of SCG_TX x SCG_TY sites of its box.  Slots at tile
// coordinates (hx, hy) stand for the box site ((x0 + hx) mod L, (y0 + hy) mod
// L), so a stencil over the slots is the periodic stencil also in a tile the
// box edge cuts (slots past the edge hold wrapped sites and are not written).
// Sites are returned as box-array indices (box_off + site).
struct SCGTileBox {
	int tile, box, o, L, x0, y0;
	bool lead;    // the box's first tile (its workgroup writes the box's state)
};
// The tile with its box's tile range and viscosity: from the launch arguments
// with one box (ONE), else one load of the tile's SCGTileGeo
struct SCGIterTile {
	SCGTileBox t;
	int tile_off, ntiles;
	double eta;
};
template <bool ONE>
__device__ inline SCGIterTile scg_iter_tile(const SCGArgs &a, int one_L, int one_ntx, int ntiles) {
	SCGIterTile g;
	const int tile = xcd_item(blockIdx.x, ntiles);   // the grid is one workgroup per tile
	g.t.tile = tile;
	if(ONE) {
		g.t.box = 0;
		g.t.o = 0;
		g.t.L = one_L;
		g.t.x0 = (tile % one_ntx) * SCG_TX;
		g.t.y0 = (tile / one_ntx) * SCG_TY;
		g.t.lead = tile == 0;
		g.tile_off = 0;
		g.ntiles = ntiles;
		g.eta = a.one_eta;
	}
	else {
		const SCGTileGeo G = a.geo[tile];
		g.t.box = G.box;
		g.t.o = G.o;
		g.t.L = G.L;
		g.t.x0 = G.x0;
		g.t.y0 = G.y0;
		g.t.lead = G.lead != 0;
		g.tile_off = G.tile_off;
		g.ntiles = G.ntiles;
		g.eta = a.set ? G.eta[1] : G.eta[0];   // a select, not an indexed copy in scratch
	}
	return g;
}
// Slot coordinates lie in -2 .. L + SCG_TX (a tile starts inside its box),
// so for L > SCG_TX / 2 they wrap by selects (at most two subtractions of L);
// smaller boxes take the remainder (a branch on L, the same for the whole
// workgroup).
__device__ inline int scg_wrap(int g, int L) {
	return g < 0 ? g + L : (g < L ? g : (g < 2 * L ? g - L : g - 2 * L));
}
__device__ inline int scg_slot_site(const SCGTileBox &T, int hx, int hy) {
	const int gx = T.x0 + hx, gy = T.y0 + hy;
	if(T.L <= SCG_TX / 2) return T.o + (gy < 0 ? gy + T.L : gy % T.L) * T.L + (gx < 0 ? gx + T.L : gx % T.L);
	return T.o + scg_wrap(gy, T.L) * T.L + scg_wrap(gx, T.L);
}

// Set-up, per site: D(k) = xi + eta (the rows of the four neighbours), added
// one triplet at a time in the order of WetPolarModel.cpp:206-215 (S, W, E, N);
// site means b_bar = P b / m, x_bar = P x0 / m; c = b_bar - D x_bar;
// y = (P r0) / m = c + eta N4(P x0), the site part of the first residual
// r0 = b - A x0; s = 0.
template <bool ONE>
__global__ __launch_bounds__(SCG_THREADS) void k_scg_init(SCGArgs a) {
	const unsigned flags = a.st->flags;   // tested before the first store, not ahead of the loads
	const SCGTileBox T = scg_iter_tile<ONE>(a, a.one_L, a.one_ntx, a.ntiles).t;
	const int lx = threadIdx.x % SCG_TX, ly = threadIdx.x / SCG_TX;
	if(T.x0 + lx >= T.L || T.y0 + ly >= T.L) return;
	const int kl = (T.y0 + ly) * T.L + T.x0 + lx, k = T.o + kl;
	const DevParams &P = a.params[2 * T.box + a.set];
	const double eta = P.friction_cell;
	int nb[4];
	site_nbrs(kl, T.L, nb);
	const int m = a.cover[k];
	const int mn[4] = { a.cover[T.o + nb[0]], a.cover[T.o + nb[1]], a.cover[T.o + nb[2]], a.cover[T.o + nb[3]] };
	double D = P.friction;
	for(int d = 0; d < 4; ++d)
		for(int l = 0; l < mn[d]; ++l) D += P.friction_cell * 1.0;
	const int n = (mn[0] + mn[1]) + (mn[2] + mn[3]);
	double2 c = make_double2(0., 0.), y = make_double2(0., 0.);
	if(m > 0) {
		const double fm = (double) m;
		const double2 sb = a.sb[k], sx = a.sx[k];
		const double2 gS = a.sx[T.o + nb[0]], gW = a.sx[T.o + nb[1]], gE = a.sx[T.o + nb[2]], gN = a.sx[T.o + nb[3]];
		c.x = sb.x / fm - D * (sx.x / fm);
		c.y = sb.y / fm - D * (sx.y / fm);
		y.x = c.x + eta * (((gS.x + gW.x) + gE.x) + gN.x);
		y.y = c.y + eta * (((gS.y + gW.y) + gE.y) + gN.y);
	}
	const double2 zero = make_double2(0., 0.);
	if(flags) return;
	a.diag[k] = D;
	a.mn[k] = m | n << 8;
	a.c[k] = c;
	a.y[0][k] = y;
	a.s[k] = zero;
}

// The one-site ring of a tile (init_w): slots at (hx, hy) in -1 .. SCG_TX x
// -1 .. SCG_TY, LDS index (hy + 1) SCG_HW + hx + 1.  Thread t owns slot
// (t mod SCG_TX, t / SCG_TX); threads 0 .. SCG_RING - 1 also take one ring slot.
constexpr int SCG_HW = SCG_TX + 2, SCG_RING = 2 * SCG_HW + 2 * SCG_TY;
struct SCGTile {
	int k;        // own slot's site
	bool in;      // the own slot is a box site of this tile
	int slot;     // own slot's LDS index
	int hk, hslot;   // ring slot's site and LDS index (hslot < 0: none)
};
__device__ inline SCGTile scg_tile(const SCGTileBox &T) {
	SCGTile t;
	const int lx = threadIdx.x % SCG_TX, ly = threadIdx.x / SCG_TX;
	t.in = T.x0 + lx < T.L && T.y0 + ly < T.L;
	t.k = scg_slot_site(T, lx, ly);
	t.slot = (ly + 1) * SCG_HW + lx + 1;
	t.hslot = -1;
	t.hk = 0;
	const int i = threadIdx.x;
	if(i < SCG_RING) {
		int hx, hy;
		if(i < SCG_HW) { hx = i - 1; hy = -1; }
		else if(i < 2 * SCG_HW) { hx = i - SCG_HW - 1; hy = SCG_TY; }
		else if(i < 2 * SCG_HW + SCG_TY) { hx = -1; hy = i - 2 * SCG_HW; }
		else { hx = SCG_TX; hy = i - 2 * SCG_HW - SCG_TY; }
		t.hk = scg_slot_site(T, hx, hy);
		t.hslot = (hy + 1) * SCG_HW + hx + 1;
	}
	return t;
}
// w = D z - eta N4(m z) at the own slot, from the slots' m z in LDS (S, W, E, N
// in the triplet order of WetPolarModel.cpp:206-215)
__device__ inline double2 scg_w_lds(const double2 *s_mz, int slot, double D, double2 z, int m, double eta) {
	if(m == 0) return make_double2(0., 0.);
	const double2 mS = s_mz[slot - SCG_HW], mW = s_mz[slot - 1], mE = s_mz[slot + 1], mN = s_mz[slot + SCG_HW];
	return make_double2(D * z.x - eta * (((mS.x + mW.x) + mE.x) + mN.x), D * z.y - eta * (((mS.y + mW.y) + mE.y) + mN.y));
}
// z = y / D where a row lives (0 elsewhere) and its product with m
__device__ inline double2 scg_z(double2 y, double D, int m) {
	const double inv = 1. / D;
	return m > 0 ? make_double2(inv * y.x, inv * y.y) : make_double2(0., 0.);
}
__device__ inline double2 scg_mz(double2 z, int m) {
	const double fm = m;
	return make_double2(fm * z.x, fm * z.y);
}

// Set-up, per tile: the partials of parity 0, with w = A z of the first
// residual.  The workgroup of each box's first tile also sums the box's row
// partials (k_scg_row_init) and writes Eigen's start (ConjugateGradient.h:
// |b|^2 and the threshold max(tol^2 |b|^2, DBL_MIN)) into SCGState [2 box],
// with rho = 1, pi = xi = 0; the first stop test is iteration 0's.
template <bool ONE>
__global__ __launch_bounds__(SCG_THREADS) void k_scg_init_w(SCGArgs a) {
	const unsigned flags = a.st->flags;   // tested before the first store, not ahead of the loads
	__shared__ double s_red[(SCG_THREADS / 64) * 6];
	__shared__ double2 s_mz[SCG_HW * (SCG_TY + 2)];
	__shared__ double s_part[6 * SCG_THREADS];
	const SCGTileBox T = scg_iter_tile<ONE>(a, a.one_L, a.one_ntx, a.ntiles).t;
	if(T.lead) {
		const SCGBox Bx = a.cgb[T.box];
		double t6[6], v[6];
		parts_load<SCG_THREADS>(a.part, a.ncells, Bx.cell_begin, Bx.cell_end - Bx.cell_begin, v);
		parts_sum<SCG_THREADS>(a.part, a.ncells, Bx.cell_begin, Bx.cell_end - Bx.cell_begin, v, t6, s_red);
		if(!flags && threadIdx.x == 0) {
			SCGState S{};
			S.max_iter = Bx.max_iter;
			for(int c = 0; c < 2; ++c) {
				const double bb = t6[c];
				S.rho[c] = 1.;
				S.bb[c] = bb;
				S.rperp2[c] = t6[2 + c];
				S.rzperp[c] = t6[4 + c];
				if(!(bb == bb) || !(t6[2 + c] == t6[2 + c]) || !(t6[4 + c] == t6[4 + c])) {
					cg_flag_nan(a.st);
					S.done[c] = 1;
					continue;
				}
				if(bb == 0) {
					S.zero_rhs[c] = 1;
					S.done[c] = 1;
					continue;
				}
				double thr = Bx.tol * Bx.tol * bb;
				if(thr < DBL_MIN) thr = DBL_MIN;
				S.thr[c] = thr;
			}
			S.active = !(S.done[0] && S.done[1]);
			for(int c = 0; c < 2; ++c) S.done_now[c] = S.done[c];
			S.active_now = S.active;
			// both copies: an iteration rewrites only the scalars that change
			a.s2[2 * T.box] = S;
			a.s2[2 * T.box + 1] = S;
		}
	}
	const double eta = a.params[2 * T.box + a.set].friction_cell;
	const SCGTile t = scg_tile(T);
	const int m = a.cover[t.k];
	const double D = a.diag[t.k];
	const double2 y = a.y[0][t.k];
	const double2 z = scg_z(y, D, m);
	if(flags) return;
	s_mz[t.slot] = scg_mz(z, m);
	if(t.hslot >= 0) {
		const int hm = a.cover[t.hk];
		s_mz[t.hslot] = scg_mz(scg_z(a.y[0][t.hk], a.diag[t.hk], hm), hm);
	}
	__syncthreads();
	const double2 w = scg_w_lds(s_mz, t.slot, D, z, m, eta);
	double part[6] = { 0., 0., 0., 0., 0., 0. };
	if(t.in) scg_part(part, m, D, y, w);
	scg_tile_partials(part, s_part, a.psum[0], a.ntiles, T.tile);
}

// box site of storage row q of cell C (MultiPhaseField ring buffer, SquareBox wrap)
__device__ inline int cell_site(const DevCell &C, int L, int q) {
	const int qy = q / C.LX, qx = q - qy * C.LX;
	int bx = qx + C.offx;
	if(bx >= C.LX) bx -= C.LX;
	bx += C.corner % L;
	if(bx >= L) bx -= L;
	int by = qy + C.offy;
	if(by >= C.LY) by -= C.LY;
	by += C.corner / L;
	if(by >= L) by -= L;
	return by * L + bx;
}

// Set-up, per cell: the ker-P part of the first residual, r_perp_i = (b_i -
// D x0_i) - c(k_i), summed as |b|^2, |r_perp|^2, r_perp.(r_perp / D) -> part
// (6 per cell, [6][ncells]: b.b x, y; r_perp.r_perp x, y; r_perp.z_perp x, y)
__global__ __launch_bounds__(256) void k_scg_row_init(SCGArgs a) {
	const unsigned flags = a.st->flags;   // tested before the partials are written, not ahead of the loads
	__shared__ double s_red[4 * 6];
	const int ci = xcd_item(blockIdx.x, gridDim.x);
	const DevCell C = a.cells[ci];
	const DevBox B = a.boxes[C.box];
	double red[6] = { 0., 0., 0., 0., 0., 0. };
	for(int q = threadIdx.x; q < C.S; q += 256) {
		const int k = B.box_off + cell_site(C, B.L, q), row = C.start + q;
		const double D = a.diag[k], inv = 1. / D;
		const double2 c = a.c[k], bb = a.b[row], x0 = a.x[row];
		const double rx = (bb.x - D * x0.x) - c.x, ry = (bb.y - D * x0.y) - c.y;
		red[0] += bb.x * bb.x; red[1] += bb.y * bb.y;
		red[2] += rx * rx; red[3] += ry * ry;
		red[4] += rx * (inv * rx); red[5] += ry * (inv * ry);
	}
	if(flags) return;
	block_sum<6>(red, s_red);
	// indexed by cell, so that each box's cells are one range
	if(threadIdx.x == 0) {
#pragma unroll
		for(int j = 0; j < 6; ++j) a.part[j * a.ncells + ci] = red[j];
	}
}

// The scalars of SCGState an iteration reads, requested with its first loads,
// and the ones it rewrites (scg_put; the others are the set-up's, the same in
// both copies).
struct SCGVar {
	double rn2[2], gamma[2], alpha[2], beta[2], pi[2], rho[2], xi[2];
	double thr[2], rperp2[2], rzperp[2];
	int done[2], iters[2], pend[2], max_iter, active;
};
__device__ inline SCGVar scg_var(const SCGState *S) {
	SCGVar v;
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		v.rn2[c] = S->rn2[c];
		v.gamma[c] = S->gamma[c];
		v.alpha[c] = S->alpha[c];
		v.beta[c] = S->beta[c];
		v.pi[c] = S->pi[c];
		v.rho[c] = S->rho[c];
		v.xi[c] = S->xi[c];
		v.thr[c] = S->thr[c];
		v.rperp2[c] = S->rperp2[c];
		v.rzperp[c] = S->rzperp[c];
		v.done[c] = S->done[c];
		v.iters[c] = S->iters[c];
		v.pend[c] = S->pend[c];
	}
	v.max_iter = S->max_iter;
	v.active = S->active;
	return v;
}
__device__ inline void scg_put(SCGState *S, const SCGVar &v) {
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		S->rn2[c] = v.rn2[c];
		S->gamma[c] = v.gamma[c];
		S->alpha[c] = v.alpha[c];
		S->beta[c] = v.beta[c];
		S->pi[c] = v.pi[c];
		S->rho[c] = v.rho[c];
		S->xi[c] = v.xi[c];
		S->done[c] = v.done[c];
		S->iters[c] = v.iters[c];
		S->pend[c] = v.pend[c];
	}
	S->active = v.active;
}

struct SCGCoef {
	bool done[2];
	double rn2[2], gamma[2], alpha[2], beta[2], pi[2];
};
// Iteration j's scalars from the sums tot of the residual entering it:
//   |r|^2 = rho^2 |r_perp|^2 + y.m y  (Eigen's stop test: of the first residual
//   at j = 0, of the last update after that)
//   gamma = r.z  = rho^2 r_perp.z_perp + y.m (y / D)
//   delta = z.Az = rho^2 r_perp.z_perp + (y / D).m w      (A z_perp = r_perp)
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//   (j = 0: beta = 0, alpha = gamma / delta), pi = rho + beta pi_prev
// (a NaN residual is not "done": the writer raises ST_NAN, which halts the device)
template <class State>
__device__ inline SCGCoef scg_coef(const State &S, const double (&tot)[6], bool first) {
	SCGCoef k;
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		k.done[c] = S.done[c];
		k.rn2[c] = S.rn2[c];
		k.gamma[c] = S.gamma[c];
		k.alpha[c] = 0.;
		k.beta[c] = 0.;
		k.pi[c] = S.pi[c];
		if(k.done[c]) continue;
		const double r2 = S.rho[c] * S.rho[c];
		const double rn2 = r2 * S.rperp2[c] + tot[c];
		k.rn2[c] = rn2;
		if(rn2 < S.thr[c] || S.iters[c] >= S.max_iter) {
			k.done[c] = true;
			continue;
		}
		const double gamma = r2 * S.rzperp[c] + tot[2 + c];
		const double delta = r2 * S.rzperp[c] + tot[4 + c];
		k.gamma[c] = gamma;
		if(first) k.alpha[c] = gamma / delta;
		else {
			k.beta[c] = gamma / S.gamma[c];
			k.alpha[c] = gamma / (delta - k.beta[c] * gamma / S.alpha[c]);
		}
		k.pi[c] = S.rho[c] + k.beta[c] * S.pi[c];
	}
	return k;
}

// One component's scalars of scg_coef (component c, chosen per lane): the
// same expressions, so that one wave forms both components' at once.
struct SCGCoef1 {
	double rn2, gamma, alpha, beta, pi;
	int done;
};
__device__ inline SCGCoef1 scg_coef1(const SCGVar &S, const double (&tot)[6], bool first, int c) {
	const bool x = c == 0;
	const double rho = x ? S.rho[0] : S.rho[1], rperp2 = x ? S.rperp2[0] : S.rperp2[1];
	const double rzperp = x ? S.rzperp[0] : S.rzperp[1], thr = x ? S.thr[0] : S.thr[1];
	const double g_prev = x ? S.gamma[0] : S.gamma[1], a_prev = x ? S.alpha[0] : S.alpha[1];
	const double pi_prev = x ? S.pi[0] : S.pi[1];
	const int iters = x ? S.iters[0] : S.iters[1];
	SCGCoef1 k;
	k.done = x ? S.done[0] : S.done[1];
	k.rn2 = x ? S.rn2[0] : S.rn2[1];
	k.gamma = g_prev;
	k.alpha = 0.;
	k.beta = 0.;
	k.pi = pi_prev;
	if(k.done) return k;
	const double r2 = rho * rho;
	const double rn2 = r2 * rperp2 + (x ? tot[0] : tot[1]);
	k.rn2 = rn2;
	if(rn2 < thr || iters >= S.max_iter) {
		k.done = 1;
		return k;
	}
	const double gamma = r2 * rzperp + (x ? tot[2] : tot[3]);
	const double delta = r2 * rzperp + (x ? tot[4] : tot[5]);
	k.gamma = gamma;
	if(first) k.alpha = gamma / delta;
	else {
		k.beta = gamma / g_prev;
		k.alpha = gamma / (delta - k.beta * gamma / a_prev);
	}
	k.pi = rho + k.beta * pi_prev;
	return k;
}

// The iteration's tile with a two-site ring: slots at tile coordinates (hx, hy)
// in -2 .. SCG_TX + 1 x -2 .. SCG_TY + 1, grid index (hy + 2) SCG_G2W + hx + 2.
// Thread t owns slot (t mod SCG_TX, t / SCG_TX) and threads 0 .. SCG_RING2 - 1
// one ring slot each: the two rows below, the two above, then the two columns
// on either side, row by row.
constexpr int SCG_G2W = SCG_TX + 4, SCG_G2H = SCG_TY + 4, SCG_RING2 = SCG_G2W * SCG_G2H - SCG_THREADS;
static_assert(SCG_RING2 <= SCG_THREADS, "one ring slot per thread");
struct SCGTile2 {
	int k;         // own slot's site
	bool in;       // the own slot is a box site of this tile
	int g;         // own slot's grid index
	int hk, hg;    // ring slot's site and grid index (hg < 0: none)
	bool hin;      // the ring slot lies in the one-site ring
};
__device__ inline SCGTile2 scg_tile2(const SCGTileBox &T) {
	SCGTile2 t;
	const int lx = threadIdx.x % SCG_TX, ly = threadIdx.x / SCG_TX;
	t.in = T.x0 + lx < T.L && T.y0 + ly < T.L;
	t.k = scg_slot_site(T, lx, ly);
	t.g = (ly + 2) * SCG_G2W + lx + 2;
	t.hg = -1;
	t.hk = 0;
	t.hin = false;
	const int i = threadIdx.x;
	if(i < SCG_RING2) {
		int hx, hy;
		if(i < 2 * SCG_G2W) {
			hy = i / SCG_G2W - 2;
			hx = i % SCG_G2W - 2;
		}
		else if(i < 4 * SCG_G2W) {
			const int r = i - 2 * SCG_G2W;
			hy = SCG_TY + r / SCG_G2W;
			hx = r % SCG_G2W - 2;
		}
		else {
			const int r = i - 4 * SCG_G2W, c = r & 3;
			hy = r >> 2;
			hx = c < 2 ? c - 2 : SCG_TX + c - 2;
		}
		t.hk = scg_slot_site(T, hx, hy);
		t.hg = (hy + 2) * SCG_G2W + hx + 2;
		t.hin = hx >= -1 && hx <= SCG_TX && hy >= -1 && hy <= SCG_TY;
	}
	return t;
}

// one slot's data entering an iteration: r's and p's site parts, m and n
// (D and 1 / D follow from n by the box's table, scg_diag)
struct SCGSite {
	double2 y, q;
	double D, inv;
	int m, n;
};
__device__ inline SCGSite scg_load(const int *mnp, const double2 *yin, const double2 *qin, int k, bool first) {
	SCGSite v;
	const int mn = mnp[k];
	v.m = mn & 255;
	v.n = min(mn >> 8, SCG_NTAB - 1);
	v.y = yin[k];
	v.q = first ? make_double2(0., 0.) : qin[k];
	return v;
}
__device__ inline void scg_diag(SCGSite &v, const double *s_tab) {
	v.D = s_tab[v.n];
	v.inv = s_tab[SCG_NTAB + v.n];
}
// z = (1 / D) y where a row lives, as scg_z, from the tabulated 1 / D
__device__ inline double2 scg_zi(double2 y, double inv, int m) {
	return m > 0 ? make_double2(inv * y.x, inv * y.y) : make_double2(0., 0.);
}
// the m-weighted four-neighbour sum of a grid of m v values (S, W, E, N in the
// triplet order of WetPolarModel.cpp:206-215) and D v - eta times it: the site
// part of A applied to v (0 where no row lives)
__device__ inline double2 scg_apply(const double2 *s_mv, int g, double D, double2 v, int m, double eta) {
	if(m == 0) return make_double2(0., 0.);
	const double2 mS = s_mv[g - SCG_G2W], mW = s_mv[g - 1], mE = s_mv[g + 1], mN = s_mv[g + SCG_G2W];
	return make_double2(D * v.x - eta * (((mS.x + mW.x) + mE.x) + mN.x), D * v.y - eta * (((mS.y + mW.y) + mE.y) + mN.y));
}

// A store marked non-temporal (streaming): the fields an iteration writes are
// read by the next launch only, which may run on another XCD, so keeping them
// in this XCD's L2 only adds to the write-back at the end of the launch; the
// streaming stores measured 9 % less CG time per step at 512^2.
typedef double rc_dbl2 __attribute__((ext_vector_type(2)));
__device__ inline void nt_store2(double2 *p, double2 v) {
	const rc_dbl2 w = { v.x, v.y };
	__builtin_nontemporal_store(w, reinterpret_cast<rc_dbl2 *>(p));
}

// Iteration j (one launch, one workgroup per tile of every box): the sums of
// the box's tiles' partials (every workgroup of the box, the same fixed
// order), the scalars of the iteration from them (two lanes, shared through
// LDS), then
//   q = z + beta q (p),  A p from q over the one-site ring,  s += alpha q,
//   y -= alpha A p,  z = y / D,  w = A z (for the sums only),
// q over the tile and its two-site ring, the new y over the one-site ring,
// each exchanged through LDS as m q and m z.  Only y, q (both parities) and s
// are stored: A p and w are formed from them.  s moves at odd iterations, by
// the even iteration's step and its own, so an iteration reads and writes
// 68 B per box site (even) or 100 B (odd).  The workgroup of the box's first
// tile writes SCGState [2 box + ((j + 1) & 1)]: rho -= alpha pi, xi += alpha
// pi and the recurrence scalars (nobody reads that copy during iteration j).
// The pointers and sizes the first loads need lead the arguments, so that
// they arrive preloaded in SGPRs (kernel-argument preloading) and the loads
// go out at the first instructions, not after a read of the argument block.
template <bool FIRST, bool ONE, int PU>
__global__ __launch_bounds__(SCG_THREADS) void k_scg_iter(const double2 *yin, const double2 *qin, const int *mnp,
		double2 *sp, const double *psin, const SCGState *Sp0, int lntx, int ntp, SCGArgs a, int j) {
	// lntx = box 0's side | its tiles per row << 16; ntp = tiles | (j & 1) << 31
	const int one_L = lntx & 0xFFFF, one_ntx = lntx >> 16;
	const int ntiles = ntp & 0x7FFFFFFF, pin = (int) ((unsigned) ntp >> 31), pout = pin ^ 1;
	__shared__ double s_red[6];
	__shared__ double2 s_mq[SCG_G2W * SCG_G2H];
	__shared__ double2 s_mz[SCG_G2W * SCG_G2H];
	__shared__ double s_part[6 * SCG_THREADS];
	__shared__ double s_tab[2 * SCG_NTAB];
	__shared__ SCGCoef1 s_cf[2];
	const SCGIterTile G = scg_iter_tile<ONE>(a, one_L, one_ntx, ntiles);
	const SCGTileBox &T = G.t;
	const SCGState *Sp = Sp0 + 2 * T.box;
	SCGState *Sn = a.s2 + 2 * T.box + pout;
	const bool writer = T.lead && threadIdx.x == 0;
	const double eta = G.eta;
	const SCGTile2 t = scg_tile2(T);
	// Every load is requested before any is used: the site data, this lane's
	// partials, the status word and the state's scalars (all addresses are
	// valid whatever they hold), so the launch waits for one memory round
	// trip, and a launch after convergence returns after that one.  The
	// state's scalars and the site data come first (their pointers arrive
	// preloaded), the loads behind a read of the argument block last.
	SCGVar V = scg_var(Sp);
	SCGSite c0 = scg_load(mnp, yin, qin, t.k, FIRST);
	SCGSite ch{};
	if(t.hg >= 0) ch = scg_load(mnp, yin, qin, t.hk, FIRST);
	// s moves at odd iterations only, by the previous iteration's pending
	// step and its own: (s + alpha_prev q_prev) + alpha q, the order of a
	// step per iteration, with q_prev the q this iteration reads anyway
	const bool s_pass = pin != 0;
	double2 s = make_double2(0., 0.);
	if(s_pass) s = sp[t.k];
	double pre[PU];
	scg_sums_load<PU>(psin, ntiles, G.tile_off, G.ntiles, pre);
	const int tab_i = min((int) threadIdx.x, 2 * SCG_NTAB - 1);
	const double tab = a.dtab[(2 * T.box + a.set) * 2 * SCG_NTAB + tab_i];
	const unsigned flags = a.st->flags;
	// an ordering point: the compiler keeps the loads above it, where they
	// were requested, instead of moving them under the branches below
	asm volatile("" ::: "memory");
	if(flags) return;
	if(!V.active) {
		if(writer) scg_put(Sn, V);
		return;
	}
	if(threadIdx.x < 2 * SCG_NTAB) s_tab[threadIdx.x] = tab;
	double tot[6];
	scg_sums_add<PU>(psin, ntiles, G.tile_off, G.ntiles, pre, tot, s_red);
	// (the table is in LDS: scg_sums_add ends with a barrier)
	// the scalars once, lanes 0 and 1 of wave 0 (one component each), for
	// every thread through LDS
	if(threadIdx.x < 2) s_cf[threadIdx.x] = scg_coef1(V, tot, FIRST, threadIdx.x);
	scg_diag(c0, s_tab);
	if(t.hg >= 0) scg_diag(ch, s_tab);
	__syncthreads();
	SCGCoef cf;
#pragma unroll
	for(int c = 0; c < 2; ++c) {
		const SCGCoef1 k = s_cf[c];
		cf.done[c] = k.done != 0;
		cf.rn2[c] = k.rn2;
		cf.gamma[c] = k.gamma;
		cf.alpha[c] = k.alpha;
		cf.beta[c] = k.beta;
		cf.pi[c] = k.pi;
	}
	if(cf.done[0] && cf.done[1]) {
		if(writer) {
#pragma unroll
			for(int c = 0; c < 2; ++c) {
				V.rn2[c] = cf.rn2[c];
				V.done[c] = 1;
			}
			V.active = 0;
			scg_put(Sn, V);
		}
		return;
	}
	const bool d0 = cf.done[0], d1 = cf.done[1];
	const double al0 = cf.alpha[0], al1 = cf.alpha[1], b0 = cf.beta[0], b1 = cf.beta[1];
	// p = z + beta p_prev at a slot (z = y / D where a row lives); a converged
	// component keeps its p
	auto pdir = [&](const SCGSite &v) {
		const double2 z = scg_zi(v.y, v.inv, v.m);
		double2 qn = v.q;
		if(!d0) qn.x = z.x + b0 * v.q.x;
		if(!d1) qn.y = z.y + b1 * v.q.y;
		return qn;
	};
	const double2 q0 = pdir(c0);
	s_mq[t.g] = scg_mz(q0, c0.m);
	// q and s are final here: their stores go out now, overlapping the rest
	// of the pass (every workgroup stores at once; the write-out, not the
	// issue, is what the end of the launch waits for)
	if(s_pass) {
		if(V.pend[0]) s.x += V.alpha[0] * c0.q.x;
		if(V.pend[1]) s.y += V.alpha[1] * c0.q.y;
		if(!d0) s.x += al0 * q0.x;
		if(!d1) s.y += al1 * q0.y;
	}
	if(t.in) {
		nt_store2(a.q[pout] + t.k, q0);
		if(s_pass) nt_store2(sp + t.k, s);
	}
	double2 qh = make_double2(0., 0.);
	if(t.hg >= 0) {
		qh = pdir(ch);
		s_mq[t.hg] = scg_mz(qh, ch.m);
	}
	__syncthreads();
	// A p and the new r over the one-site ring; m z of the new r into LDS
	auto rstep = [&](const SCGSite &v, double2 qn, int g) {
		const double2 ap = scg_apply(s_mq, g, v.D, qn, v.m, eta);
		double2 yn = v.y;
		if(!d0) yn.x = v.y.x - al0 * ap.x;
		if(!d1) yn.y = v.y.y - al1 * ap.y;
		return yn;
	};
	const double2 yn0 = rstep(c0, q0, t.g);
	const double2 z0 = scg_zi(yn0, c0.inv, c0.m);
	s_mz[t.g] = scg_mz(z0, c0.m);
	if(t.hin) s_mz[t.hg] = scg_mz(scg_zi(rstep(ch, qh, t.hg), ch.inv, ch.m), ch.m);
	__syncthreads();
	const double2 wn = scg_apply(s_mz, t.g, c0.D, z0, c0.m, eta);
	double part[6] = { 0., 0., 0., 0., 0., 0. };
	if(t.in) {
		nt_store2(a.y[pout] + t.k, yn0);
		scg_part_i(part, c0.m, c0.inv, yn0, wn);
	}
	if(writer) {
#pragma unroll
		for(int c = 0; c < 2; ++c) {
			V.rn2[c] = cf.rn2[c];
			if(!(cf.rn2[c] == cf.rn2[c])) cg_flag_nan(a.st);
			V.pend[c] = !s_pass && !cf.done[c];
			if(cf.done[c]) {
				V.done[c] = 1;
				continue;
			}
			V.gamma[c] = cf.gamma[c];
			V.alpha[c] = cf.alpha[c];
			V.beta[c] = cf.beta[c];
			V.rho[c] = V.rho[c] - cf.alpha[c] * cf.pi[c];
			V.xi[c] = V.xi[c] + cf.alpha[c] * cf.pi[c];
			V.pi[c] = cf.pi[c];
			V.iters[c] = V.iters[c] + 1;
		}
		V.active = 1;
		scg_put(Sn, V);
	}
	scg_tile_partials(part, s_part, a.psum[pout], a.ntiles, T.tile);
}

// The stop test of each box's state entering iteration j (after a batch),
// one workgroup per box, written into SCGState [2 box + (j & 1)] as rn2_now,
// done_now, active_now for the host; the sums in the order of the iteration
// that would use them
__global__ __launch_bounds__(SCG_THREADS) void k_scg_check(SCGArgs a, int j) {
	__shared__ double s_red[(SCG_THREADS / 64) * 6];
	const int box = blockIdx.x;
	const SCGBox Bx = a.cgb[box];
	SCGState *S = a.s2 + 2 * box + (j & 1);
	const int active = S->active;
	double tot[6] = { 0., 0., 0., 0., 0., 0. };
	if(active && j > 0) {
		double pre[SCG_PU];
		scg_sums_load<SCG_PU>(a.psum[j & 1], a.ntiles, Bx.tile_off, Bx.ntiles, pre);
		scg_sums_add<SCG_PU>(a.psum[j & 1], a.ntiles, Bx.tile_off, Bx.ntiles, pre, tot, s_red);
	}
	if(threadIdx.x != 0) return;
	if(!active || j == 0) {
		for(int c = 0; c < 2; ++c) {
			S->rn2_now[c] = S->rn2[c];
			S->done_now[c] = S->done[c];
		}
		S->active_now = active;
		return;
	}
	const SCGCoef cf = scg_coef(*S, tot, false);
	for(int c = 0; c < 2; ++c) {
		S->rn2_now[c] = cf.rn2[c];
		S->done_now[c] = cf.done[c];
		if(!(cf.rn2[c] == cf.rn2[c])) cg_flag_nan(a.st);
	}
	S->active_now = !(cf.done[0] && cf.done[1]);
}

// x_i = x0_i + xi (r_perp_i / D) + s(k_i), r_perp as in k_scg_row_init, with
// the scalars of the cell's box; a component with a zero right-hand side is 0
// (Eigen's solve of b = 0)
__global__ __launch_bounds__(256) void k_scg_final(SCGArgs a, int j) {
	const unsigned flags = a.st->flags;   // tested before the first store, not ahead of the loads
	const int ci = xcd_item(blockIdx.x, gridDim.x);
	const DevCell C = a.cells[ci];
	const DevBox B = a.boxes[C.box];
	const SCGState *S = a.s2 + 2 * C.box + (j & 1);
	const double xi0 = S->xi[0], xi1 = S->xi[1];
	const bool z0 = S->zero_rhs[0], z1 = S->zero_rhs[1];
	// the step an even last iteration left pending (its q in the parity 1 copy)
	const bool pe0 = S->pend[0], pe1 = S->pend[1];
	const double pa0 = S->alpha[0], pa1 = S->alpha[1];
	for(int q = threadIdx.x; q < C.S; q += 256) {
		const int k = B.box_off + cell_site(C, B.L, q), row = C.start + q;
		const double D = a.diag[k], inv = 1. / D;
		const double2 c = a.c[k], bb = a.b[row], x0 = a.x[row];
		double2 sv = a.s[k];
		if(pe0 || pe1) {
			const double2 qp = a.q[1][k];
			if(pe0) sv.x += pa0 * qp.x;
			if(pe1) sv.y += pa1 * qp.y;
		}
		const double rx = (bb.x - D * x0.x) - c.x, ry = (bb.y - D * x0.y) - c.y;
		double2 x;
		x.x = z0 ? 0. : (x0.x + xi0 * (inv * rx)) + sv.x;
		x.y = z1 ? 0. : (x0.y + xi1 * (inv * ry)) + sv.y;
		if(flags) return;
		a.x[row] = x;
	}
}

hipError_t launch_scg(int which, const SCGArgs &a, int j, hipStream_t s) {
	switch(which) {
	case SCG_INIT:
		if(a.nboxes == 1) RC_LAUNCH(k_scg_init<true>, dim3(a.ntiles), dim3(SCG_THREADS), 0, s, a);
		else RC_LAUNCH(k_scg_init<false>, dim3(a.ntiles), dim3(SCG_THREADS), 0, s, a);
		break;
	case SCG_INIT_W:
		if(a.nboxes == 1) RC_LAUNCH(k_scg_init_w<true>, dim3(a.ntiles), dim3(SCG_THREADS), 0, s, a);
		else RC_LAUNCH(k_scg_init_w<false>, dim3(a.ntiles), dim3(SCG_THREADS), 0, s, a);
		break;
	case SCG_ROW_INIT: RC_LAUNCH(k_scg_row_init, dim3(a.ncells), dim3(256), 0, s, a); break;
	case SCG_ITER: {
		const double2 *yin = a.y[j & 1], *qin = a.q[j & 1];
		const double *psin = a.psum[j & 1];
		const int lntx = a.one_L | a.one_ntx << 16, ntp = (int) ((unsigned) a.ntiles | (unsigned) (j & 1) << 31);
		const SCGState *Sp0 = a.s2 + (j & 1);
#define RC_SCG_ITER(F, O, U) RC_LAUNCH((k_scg_iter<F, O, U>), dim3(a.ntiles), dim3(SCG_THREADS), 0, s, yin, qin, a.mn, a.s, psin, \
		Sp0, lntx, ntp, a, j)
		const bool big = a.max_box_tiles > 64 * SCG_PU, one = a.nboxes == 1, first = j == 0;
		if(big) {
			if(one) { if(first) RC_SCG_ITER(true, true, SCG_PU_BIG); else RC_SCG_ITER(false, true, SCG_PU_BIG); }
			else { if(first) RC_SCG_ITER(true, false, SCG_PU_BIG); else RC_SCG_ITER(false, false, SCG_PU_BIG); }
		}
		else {
			if(one) { if(first) RC_SCG_ITER(true, true, SCG_PU); else RC_SCG_ITER(false, true, SCG_PU); }
			else { if(first) RC_SCG_ITER(true, false, SCG_PU); else RC_SCG_ITER(false, false, SCG_PU); }
		}
#undef RC_SCG_ITER
		break;
	}
	case SCG_CHECK: RC_LAUNCH(k_scg_check, dim3(a.nboxes), dim3(SCG_THREADS), 0, s, a, j); break;
	case SCG_FINAL: RC_LAUNCH(k_scg_final, dim3(a.ncells), dim3(256), 0, s, a, j); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

}  // namespace rc
