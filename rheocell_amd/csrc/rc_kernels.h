// rc_kernels.h -- launch interface of the HIP kernels (rc_kernels.hip).
#pragma once

#include "rc_device.h"

namespace rc {

enum StageMode { MODE_FUSED = 0, MODE_CF = 1, MODE_FS = 2 };

// Arguments of the per-cell stage kernel (one workgroup per cell).
struct StageArgs {
	const DevBox *boxes;
	DevCell *cells;
	const DevParams *params;     // params[2*box + set]
	int cf_set, fs_set;          // parameter sets of the force evaluation / field update
	int store;                   // 1: predictor (first_step(true)), 0: corrector
	int step;                    // step index, recorded with a growth request
	const int *step_base;        // added to step when set (graph replays: the window's first step)
	int pure;                    // no side effects on cells / Repulsion (MODE_CF: materialise;
	                             // predictor modes: shape tuner, outputs to the predictor's scratch rows)
	int jq;                      // accumulate Repulsion and CIL sums (J_Q != 0)
	int write_v;                 // MODE_FUSED: also write v rows
	const double *phi_c;         // phi entering the force evaluation
	double *phi_out;             // field update output
	const double *phi_old;       // corrector: phi_n
	double *dphi_old;            // predictor: written, corrector: read
	const double *phi2;          // box sum of phi^2
	const double *trig;          // patch trig tables
	double2 *vrows;              // velocity rows (read in MODE_FS, written by dry MODE_CF / write_v)
	double2 *frows;              // wet MODE_CF: force rows (CG right-hand side)
	double *rep;                 // Repulsion rows (jq)
	double *out_mu, *out_dx, *out_dy, *out_lap;   // optional site outputs (MODE_CF)
	const double *gauss;         // predictor noise, one per cell
	int4 *geo;                   // compact geometry, rewritten at recentering
	int threads, npt;            // stage shape: threads per cell, rows per thread strip (>= every patch's)
	const int *cell_list;        // cell of each workgroup (null: workgroup i takes cell i)
	int rcp;                     // lambda and every friction are powers of two (DevParams *_rcp)
	// walls and external forces (null: none): per box site, bit 1 / 2 / 4 / 8 =
	// the E / W / N / S neighbour is open (getWalls < wall_slip), 16 = the site
	// itself is; and the wall field getWalls
	const unsigned char *wmask;
	const double *walls;
	int win;                     // LDS holds the box window of sum phi^2 (some omega != 0)
	int p_tile, p_rows, p_sites; // k_stage_p: {phi, mu} tile capacity (double2), row-table length, largest patch
	int p_nbox;                  // k_stage_p: boxes (their parameters are staged in LDS)
	DevStatus *st;
};

// compact per-cell geometry: {corner, LX | LY << 16, offx | offy << 16, row start}
struct GatherArgs {
	const DevBox *boxes;
	const int *tile_box;
	const int4 *geo;
	const double *phi;
	double *phi2;
	int *cover;
	int *tile_cells;             // cells intersecting each tile, cell order
	int4 *tile_geo;              // their entries relative to the tile (tile_entry), written with the list
	int *tile_count;
	int build_lists;             // 1: rebuild the lists from geo, 0: reuse them
	int nboxes;                  // 1: box 0's geometry from the fields below, no table load
	int one_L, one_ntx, one_cells;
	int nem;                     // also the nematic box sums of phi Q00, phi Q01 (Q of cells[])
	const DevCell *cells;
	double *sq00, *sq01;
	DevStatus *st;
};

// CG_GATHER_BX: box sums of b (into sp) and of x (into p2) in one pass
enum CGGatherMode { CG_GATHER_X = 0, CG_GATHER_P_FIRST = 1, CG_GATHER_P_UPDATE = 2, CG_GATHER_FEAT = 3, CG_GATHER_E = 4,
	CG_GATHER_BX = 5 };
enum CGRowMode { CG_ROW_RESID = 0, CG_ROW_UPDATE = 2, CG_ROW_STEP = 5, CG_ROW_REVERT = 6 };

struct CGArgs {
	const DevBox *boxes;
	const int *tile_box;
	const DevCell *cells;
	const int4 *geo;
	const DevParams *params;     // wet parameters (params[2*0 + set])
	int set;
	const int *tile_cells;
	const int4 *tile_geo;
	const int *tile_count;
	const int *cover;
	double *diag;                // D(k) = xi + eta * sum of the 4 neighbour covers
	double2 *sp;                 // box sums of the gathered vector
	double2 *p2;                 // box sums of p^2
	double2 *n4;                 // sum of sp over the 4 box neighbours
	double2 *x, *r, *p, *b;
	// learned refinement stage: network input (bf16 [L*L][32]) and output (fp32 [L*L][2]),
	// feature / output multipliers, row count (residual scale sqrt(|r_c|^2 / rows))
	unsigned short *act0;
	const double *phi2;
	const float *net_out;
	float sf, sp2, sv;
	double rows;
	int nbox_blk;                // workgroups of a box-site pass (partials of k_cg_box<1>)
	int apply_update;            // k_cg_box<1>: an extra workgroup applies the previous k_cg_row<UPDATE>'s state update
	const double2 *xprev, *rprev;  // previous-solution start and its residual (learned fallback)
	double *part;                // per-cell partial sums, 6 per cell
	double *part_box;            // per-workgroup partials of the box pass, 2 each
	CGState *cg;
	double tol;
	int max_iter;
	int ncells;
	int nboxes;                  // boxes of the handle (k_cg_diag; the other row-form kernels take box 0)
	DevStatus *st;
};

hipError_t launch_gather(const GatherArgs &a, int ntiles, hipStream_t s);

// Box fields of the VelocityField / ShapeField observables: per box site, the
// sum over the covering patches (cell order) of w * phi, w = the velocity row
// (which 0) or the cell's (S00, S01) (which 1), over the tiles of one box
// (tile0 .. tile0 + ntiles - 1); the tile lists must describe the current patches.
struct FieldArgs {
	const DevBox *boxes;
	const int *tile_box;
	const int4 *tile_geo;
	const int *tile_count;
	const int *tile_cells;
	const DevCell *cells;
	const double *phi;
	const double2 *v;
	int which;                   // 0 velocity rows, 1 the cell's (S00, S01), 2 the cell's (Q00, Q01)
	int tile0;
	int global;                  // 1: outputs indexed by the box-array site (box_off + x + L y)
	double *fx, *fy;             // L*L sums, box-local site index x + L y (global: every box)
};
hipError_t launch_box_field(const FieldArgs &a, int ntiles, hipStream_t s);
// size_grid x size_grid periodic box average of fx, fy (L*L each) into cx, cy
hipError_t launch_box_coarse(const double *fx, const double *fy, double *cx, double *cy, int L, int size_grid,
		hipStream_t s);
// Dynamic LDS of the stage kernel for patches up to LX x LY and strips of npt
// rows: {phi, mu} tile of LY + npt rows of LX + 2 columns (periodic columns on
// both sides, periodic images of the first rows below), per-row trig and box
// offsets, the Repulsion tile (jq) and the box window of sum phi^2 (win).
constexpr size_t STAGE_LDS_MAX = 160 * 1024 - 4096;   // LDS per CU less the kernel's static arrays
inline size_t stage_lds_bytes(int LX, int LY, int npt, bool jq, bool win) {
	const size_t tr = (size_t) LY + npt, pw = (size_t) LX + 2;
	size_t b = 16 * tr * pw + 16 * tr + 16 * tr;
	if(jq) b += 8 * ((tr * pw + 1) & ~(size_t) 1);
	if(win) b += 8 * pw * ((size_t) LY + 2);
	return b;
}
hipError_t launch_stage(int mode, const StageArgs &a, int ncells, size_t lds_bytes, hipStream_t s);

// The dry fused stage as one persistent launch over every cell (k_stage_p):
// grid workgroups of 1024 threads, one per CU, each walking a contiguous range
// of cells with the next cell's loads in flight.  For parameter sets without
// J_Q, omega and beta, periodic boxes, and patches whose strips of 1024
// threads have at most P_STRIP_ROWS rows (p_rows_of); StageArgs::npt is the
// launch's strip rows (2 or 3), at least every cell's.
constexpr int P_THREADS = 1024, P_STRIP_ROWS = 3;
// the tuner considers it only with at least this many cells per CU (its
// overlap needs several cells per workgroup); below, results do not depend
// on timing: the groups run
constexpr int P_MIN_CELLS_PER_CU = 2;
inline int p_rows_of(int LX, int LY) { return LX >= 1 && LX <= P_THREADS ? (LY + P_THREADS / LX - 1) / (P_THREADS / LX) : 1 << 20; }
// tile capacity (double2) of one patch at npt strip rows; LDS of a launch (StageArgs p_*, store)
inline int p_tile_of(int LX, int LY, int npt) { return (LY + npt) * (LX + 2); }
inline size_t stage_p_lds_bytes(const StageArgs &a, int cells_per_wg) {
	return 16 * (size_t) a.p_tile + 16 * (size_t) a.p_rows + 8 * (size_t) a.p_sites * (a.store ? 2 : 1) +
			8 * 12 * 2 * (size_t) a.p_nbox + 8 * (size_t) (a.p_nbox + 1) + 64 * (size_t) cells_per_wg + 8 * (size_t) cells_per_wg;
}
hipError_t launch_stage_p(const StageArgs &a, int ncells, int grid, hipStream_t s);

// WetModel::f_interaction + calc_internal_forces (WetModel.cpp:488-641) of
// every cell, one workgroup per cell: the stress-divergence passive force and
// the self / inter-cell nematic forces from the box sums of phi Q; writes the
// force rows (wet) or v = f / friction (dry).  LDS: nem_lds_bytes(max sites).
struct NemArgs {
	const DevBox *boxes;
	const DevCell *cells;
	const DevParams *params;     // params[2*box + set]
	int set;
	const double *phi;
	const double *phi2, *sq00, *sq01;
	double2 *frows;              // wet: the CG right-hand side
	double2 *vrows;              // dry: the velocity
	const unsigned char *wmask;  // walls and external forces (StageArgs), or null
	const double *walls;
	DevStatus *st;
};
constexpr size_t NEM_LDS_MAX = 160 * 1024 - 1024;
// phi and f - mu phi per site; in a wall box also the two masked differences
inline size_t nem_lds_bytes(int max_sites, bool walls) { return (walls ? 32 : 16) * (size_t) max_sites; }
hipError_t launch_nem_force(const NemArgs &a, int ncells, size_t lds_bytes, hipStream_t s);
hipError_t launch_cg_diag(const CGArgs &a, int nbox_sites, hipStream_t s);
hipError_t launch_cg_gather(int mode, const CGArgs &a, int ntiles, hipStream_t s);
// mode 0: N4 of sp; 1: N4 and the p.Ap partials; 2: N4 of the network output (learned stage)
hipError_t launch_cg_box(int mode, const CGArgs &a, int nsites, hipStream_t s);
hipError_t launch_cg_row(int mode, const CGArgs &a, hipStream_t s);
// what: 0 after the residual pass, 1 after p.Ap, 2 after the x,r update,
// 3 step of a learned stage, 4 after a learned stage's x,r update,
// 5 after the learned fallback pass
hipError_t launch_cg_reduce(int what, const CGArgs &a, int nblk, hipStream_t s);
// snapshot channels: chan 0 -> sp into [k][0..1], chan 2 -> phi2 into [k][2],
// chan 3 -> sp into [k][3..4], chan 5 -> [k][3..4] -= sp; sp is divided by
// sqrt(rn2_c / rows) when rn2 is given
hipError_t launch_box_snapshot(const CGArgs &a, const double *phi2, float *snap, int chan, const double *rn2, double rows,
		int nsites, hipStream_t s);
hipError_t launch_zero_rows(double2 *x, long long rows, int comp_mask, hipStream_t s);

// Site-space CG (SCGState): every wet box of the handle (cgb[box]), box sums
// of b and x0 in sb and sx (k_cg_gather<CG_GATHER_BX>), D in diag, rows per
// site in cover; the site fields are indexed like the box arrays (box_off +
// site).  Site passes run one workgroup per SCG_TX x SCG_TY tile of a box
// (ntiles over all boxes, box-major; geo[tile] = its box and corner), one site
// per thread.  Each box solves on its own: its own scalars (s2[2 box + parity]) and
// partial sums (its tiles' range of psum).  Iteration j reads the fields of
// parity j & 1 and writes those of the other parity (its neighbours' old
// values stay readable), and the partial sums likewise.
constexpr int SCG_TX = 64, SCG_TY = 16, SCG_THREADS = SCG_TX * SCG_TY;   // iteration tile: SCG_TX x SCG_TY sites
// One tile of the iteration: its box, the box's first site, side, first tile
// and tile count, the tile's corner, whether it is the box's first tile, and
// the box's viscosity (friction_cell) in both parameter sets.  The iteration
// reads it in one load (with one box, it builds it from the launch arguments
// instead and needs no load before the site data).
struct SCGTileGeo {
	int box, o, L, tile_off, ntiles, x0, y0, lead;
	double eta[2];
};
// D(k) = friction + friction_cell added once per row of the four neighbours,
// so it depends only on the neighbours' row count n <= 4 STORE_MAX: the
// iteration takes D and 1 / D from a table by n (the same additions and the
// same division as per site) and reads m and n packed in one int per site.
constexpr int SCG_NTAB = 4 * STORE_MAX + 1;
struct SCGArgs {
	const DevBox *boxes;
	const DevCell *cells;
	const DevParams *params;
	int set;
	const int *cover;
	double *diag;                // D(k), written by SCG_INIT
	int *mn;                     // m | n << 8 per site (rows here, rows of the four neighbours), written by SCG_INIT
	const double *dtab;          // [nboxes][2 parameter sets][D, 1 / D][SCG_NTAB] by n
	const double2 *sb, *sx;      // box sums of b and of x0
	double2 *c;                  // b_bar - D x0_bar per site (site means of b, x0)
	double2 *s;                  // site part of x - x0
	double2 *y[2];               // site part of r
	double2 *q[2];               // site part of p
	const double2 *b;            // right-hand side rows
	double2 *x;                  // start rows in, solution rows out
	double *part;                // row-pass partials, 6 per cell
	double *psum[2];             // site partials y.m y, y.m (y / D), (y / D).m w: [6][ntiles] each
	SCGState *s2;                // [2 * nboxes]
	const SCGBox *cgb;           // [nboxes]
	const SCGTileGeo *geo;       // [ntiles]
	int one_L, one_ntx;          // box 0's side and tiles per row (read when nboxes == 1)
	double one_eta;              // box 0's friction_cell in the current parameter set
	int max_box_tiles;           // tiles of the largest box (the iteration's partial-sum preload)
	int ncells, ntiles, nboxes;
	DevStatus *st;
};
// launches of the site-space CG, in solve order:
//   SCG_INIT          per site: D, c, y = c + eta N4(sx), s = 0
//   SCG_ROW_INIT      per cell: |b|^2 and the ker-P residual sums over the rows
//   SCG_INIT_W        per site: the partials of parity 0 (with w = the site part of A z);
//                     per box: Eigen's start into SCGState [2 box] and [2 box + 1]
//   SCG_ITER          iteration j: gamma, delta, the stop test, alpha, beta; p and A p
//                     from it, x and r updated; the partials of the new r; SCGState [(j + 1) & 1]
//   SCG_CHECK         one workgroup per box: the stop test of the state entering iteration j
//   SCG_FINAL         x = x0 + xi z_perp + P^T s from the state entering iteration j
enum SCGLaunch { SCG_INIT = 0, SCG_ROW_INIT, SCG_INIT_W, SCG_ITER, SCG_CHECK, SCG_FINAL };
hipError_t launch_scg(int which, const SCGArgs &a, int j, hipStream_t s);
// map[j] = {old row start, new row start, rows}: dst[new + q] = src[old + q] (and src2 -> dst2 if set)
// growth re-cut of listed patches (k_grow_rows; 4 int4 per entry), at most
// GROW_LDS_MAX / 4 sites per new patch
constexpr size_t GROW_LDS_MAX = 128 * 1024;
hipError_t launch_grow_rows(const double *src, double *dst, const double *src2, double *dst2, const int4 *ent, int n,
		int max_sites, double *area_out, hipStream_t s);
hipError_t launch_repack_rows(const double *src, double *dst, const double *src2, double *dst2, const int4 *map, int n,
		hipStream_t s);
hipError_t stage_set_lds_limit(size_t bytes);
// sites-per-thread instantiation for a largest patch of max_sites (-1: too large)
// stage kernel shapes {threads per cell, rows per thread strip}; returns their count
int stage_shapes(int (*shapes)[2]);
// rows per thread strip of an LX x LY patch with the given threads per cell
int stage_strip_rows(int LX, int LY, int threads);
// events recorded by the next launches at kernel start / end (null: off)
void set_launch_events(hipEvent_t start, hipEvent_t stop);

}  // namespace rc
