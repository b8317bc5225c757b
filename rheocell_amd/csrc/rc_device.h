// rc_device.h -- device-side data model of the MI355X RheoCell hot path.
//
// HBM layout (one handle):
//   rows   : every per-site vector is one array over all patch sites of all
//            cells, cell-major in cell-index order, storage order inside a
//            patch (the field_start_index layout of WetPolarModel.cpp:160-162).
//   cells  : array of DevCell (per-cell scalars, BaseField.h:122-225).
//   boxes  : each replica's L x L box-grid arrays at box_off (row-major y*L+x).
//   tiles  : 32 x 8 box tiles, tile_off per box; list of intersecting cells.
#pragma once

#include <hip/hip_runtime.h>

namespace rc {

constexpr float PI_F = 3.141592653589793238462643f;  // src/defs.h:4 keeps PI a float
constexpr int TW = 32;           // box tile width  (one 64-lane wave = 2 rows)
constexpr int TH = 8;            // box tile height (256 sites per tile)
constexpr int MAXT = 128;        // cells intersecting one tile (bounded by 20 x 6)
constexpr int STORE_MAX = 20;    // WetPolarModel.cpp:51
constexpr int BORDER = 4;        // BaseField.cpp:5
constexpr int CELL_THREADS = 256;

// status flags written by kernels; every kernel returns at once when set
constexpr unsigned ST_OVERLAP = 1u;
constexpr unsigned ST_GROWTH = 2u;
constexpr unsigned ST_TILE = 4u;
constexpr unsigned ST_NAN = 8u;
// a growth request also ORs (step + 1) << ST_STEP_SHIFT into the flags: every
// cell of one corrector stage writes the same bits, so one plain load tells a
// later launch of that stage that the request is its own
constexpr unsigned ST_KIND_MASK = 0xFFu;
constexpr int ST_STEP_SHIFT = 8;

struct DevParams {
	double gamma, lambda, omega, mu, kappa, a0, p0, beta, passive_alpha;
	double zetaQ_self, zetaQ_inter, J_Q, friction, friction_cell, J0, dt, sqrt_dt;
	double lambda_rcp, friction_rcp;   // exact 1/lambda, 1/friction when a power of two, else 0
	double mu_a0, rcoef, pcoef;        // mu / a0, 4 (kappa / lambda), 4 (beta / p0), formed on the host
	int anchoring;
	int nematic;                       // interaction_type = wetmodel (WetModel.cpp)
	// external forces (BaseField::set_F_ext / set_f_density_ext / set_V_ext):
	// free energy sum_f ((fF[f] phi) w) w, density sum_f (((fD[f] phi) phi) w) w,
	// and the last force's velocity profile (vext_kind: 0 none, 1 shear, 2
	// Poiseuille, 3 stretching) with its lambda and current shear rate
	int nforce, vext_kind;
	double fF[4], fD[4];
	double vext_lambda, vext_rate;
};

struct DevCell {
	int LX, LY, S, start;
	int offx, offy, corner, corner_old;
	int x_sub_left, y_sub_bottom, box, trig_x;
	int trig_y, pad0;
	double CoMx, CoMy, CoM_oldx, CoM_oldy;
	double unrapx, unrapy;
	double Q00, Q01, thetaQ, thetaQ_old;
	double area, area_old, sumF, sumF_old, perimeter, S00, S01;
	double trx, try_;
	double nem0, nem1, nem_old0, nem_old1;   // nematic director nemQ, nemQ_old (WetModel)
};

struct DevBox {
	int L, box_off, tile_off, ntx, nty, cell_begin, cell_end, pad;
};

struct DevStatus {
	unsigned flags;
	int first_cell;
	int site;
	int step;        // step index (within the current rc_step call) that raised the flag
};

// CG state of the two simultaneous solves (x and y components share A)
struct CGState {
	double bb[2], thr[2], absNew[2], alpha[2], beta[2], rn2[2];
	double alpha0[2];   // step along the correction of a learned refinement stage
	double rn2_start[2];  // |r_c|^2 of the previous-solution start (learned fallback test)
	int done[2], iters[2], zero_rhs[2], max_iter, active;
	int reverted[2];    // learned start dropped for the previous solution
};

// Site-space form of the same CG.  Every row sits at one box site k_i, and the
// diagonal D(k_i) depends on the site only, so A = D - eta P^T Nb P (P: sum of
// the rows at each site, Nb: the four box neighbours) maps the rows that sum
// to zero at every site (ker P) onto themselves (as D) and the site-constant
// rows P^T y onto themselves (as P^T (D y - eta Nb(m y)), m = rows per site);
// the Jacobi preconditioner D keeps both too.  Eigen's iterates are therefore
//   r = rho r_perp + P^T y,   p = pi z_perp + P^T q,   x = x0 + xi z_perp + P^T s
// with r_perp the ker-P part of the first residual and z_perp = r_perp / D:
// scalars rho, pi, xi per component and L x L fields y, q, s.  The inner
// products split the same way (r_perp . P^T y = 0, P^T a . P^T b = a . m b).
// The recurrences are Chronopoulos and Gear's arrangement of the same PCG
// (gamma = r.z and delta = z.Az reduced together, alpha = gamma / (delta -
// beta gamma / alpha_prev); A p formed from p in the pass): one reduction, one
// pass per iteration.  Two copies, read from [j & 1] and written to
// [(j + 1) & 1] by iteration j; the set-up writes both, and an iteration
// rewrites only the scalars that change.
// One box of the site-space CG: side, first box site, tiles (SCG_TX x SCG_TY,
// ntx per tile row, the box's first tile in the handle's tile order), cells,
// and the solver tolerance and iteration cap of the box's parameters.
struct SCGBox {
	int L, box_off, ntx, tile_off, ntiles, cell_begin, cell_end, max_iter;
	double tol;
};

struct SCGState {
	double bb[2], thr[2], rn2[2];
	double rperp2[2], rzperp[2];   // |r_perp|^2, r_perp . (r_perp / D)
	double rho[2], pi[2], xi[2];
	double gamma[2], alpha[2], beta[2];   // r.z, alpha and beta of the last iteration
	int done[2], iters[2], zero_rhs[2], max_iter, active;
	// s lacks the step alpha q of the last iteration (an even one: even
	// iterations leave their step of s to the next, which reads q anyway;
	// that q is in the parity 1 copy)
	int pend[2];
	// stop test of the last update, written by k_scg_check after a batch
	double rn2_now[2];
	int done_now[2], active_now, pad;
};

}  // namespace rc
