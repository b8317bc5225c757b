/*
 * rc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the RheoCell
 * per-step hot path.  A plain-C restatement of the reference algorithm:
 *
 *   FD_CPUBackend::init / first_step / compute_forces / sim_step
 *       (src/Backends/FD_CPUBackend.cpp:19-160)
 *   WetPolarModel::begin_energy_computation, computeGlobalSums,
 *       f_interaction, calc_internal_forces, updateDirectedActiveForces
 *       (src/Interactions/WetPolarModel.cpp:84-430)
 *   BaseInteraction::derivX/derivY/Laplacian/updateFieldProperties
 *       (src/Interactions/BaseInteraction.h:196-263, BaseInteraction.cpp:62-69)
 *   MultiPhaseField ring buffer / recentering / growth
 *       (src/Fields/MultiPhaseField.cpp:27-359)
 *   SquareBox periodic map (src/Boxes/SquareBox.cpp:21-82, BaseBox.cpp:29-50)
 *   WetModel, the nematic variant (interaction_type = wetmodel):
 *       computeGlobalSums, f_interaction, calc_internal_forces,
 *       updateDirectedActiveForces and the solver restart rule
 *       (src/Interactions/WetModel.cpp:194-233, 423-747), and the director
 *       set up by MultiPhaseField::init (src/Fields/MultiPhaseField.cpp:68-98)
 *   walls and external forces: BaseInteraction's wall-masked stencils
 *       (src/Interactions/BaseInteraction.h:196-263), the wall branches of both
 *       models, BaseField::set_F_ext / set_f_density_ext / set_V_ext
 *       (src/Fields/BaseField.h:54-120) with the forces of src/Forces
 *   Utils::gaussian (src/Utilities/Utils.h:210-234) on POSIX drand48
 *   Eigen::ConjugateGradient<SparseMatrix, Upper> with the default Jacobi
 *       preconditioner (Eigen 3, un-vendored; restated from its published
 *       algorithm: r=b-Ax, stop when |r|^2 < max(tol^2 |b|^2, DBL_MIN),
 *       z = invdiag.*r, p = z + (rz_new/rz_old) p).
 *
 * Expressions keep the reference's operand order so that, compiled with
 * -ffp-contract=off, every per-site value is the IEEE evaluation of the
 * reference source.  The float PI of src/defs.h:4 is kept as float.
 *
 * Parity pins (tests/): trig tables, ring-buffer map, recentering and growth
 * against the reference's own MultiPhaseField/SquareBox compiled from
 * source (oracle/build_ref.sh); drand48 against glibc; formulas against
 * closed forms; the wet operator against explicit triplet assembly; the CG
 * against dense numpy solves.  Eigen's exact CG iterate sequence is parity
 * unpinned (Eigen is absent from this image).
 */
#define _GNU_SOURCE 1
#include "rc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.141592653589793238462643f   /* src/defs.h:4, a float literal */
#define STORE_MAX 20                          /* WetPolarModel.cpp:51 */
#define BORDER 4                              /* BaseField.cpp:5 */

typedef struct ocell {
	int index;
	int LX, LY, S;
	int offset[2];
	int corner, corner_old;
	double unrap[2];
	int x_sub_left, y_sub_bottom;
	double CoM[2], CoM_old[2];
	double Q00, Q01, thetaQ, thetaQ_old;
	double area, area_old, sumF, sumF_old, perimeter, S00, S01;
	double tf[2], tr[2];
	int run_clock;
	double *phi, *phi_old, *dphi_old, *mu, *dx, *dy, *lap, *rep, *pressure;
	double *fd;                /* freeEnergyDensity (WetModel) */
	double nem[2], nem_old[2]; /* nemQ, nemQ_old (WetModel) */
	int *nbr;                  /* 9*S, periodic in the patch */
	int *map;                  /* map_sub_to_box */
	double *cosx, *sinx, *cosy, *siny;
} ocell;

struct orc_sim {
	int L, N;
	orc_params act;            /* values as read from the input */
	/* currently applied values (WetPolarModel members) */
	double zetaQ_self, zetaQ_inter, J_Q, friction, friction_cell, beta;
	double a0, p0, tolerance;
	ocell *c;
	int *bnbr;                 /* 9*L*L box neighbours */
	double *phi2, *sum_phi;
	double *sq00, *sq01;       /* sumQ00, sumQ01: box sums of phi Q (WetModel) */
	double *walls;             /* getWalls of the box (NULL: 0 everywhere) */
	int nforce;
	struct { int kind; double kappa_wall, lambda, rate_active, rate; } force[ORC_MAX_FORCES];
	int nematic;               /* interaction_type = wetmodel */
	int *store_size, *store_idx;
	int *start;
	int size_rows, size_rows_old, restart_solver;
	int rows_cap;
	double *vx, *vy, *fx, *fy;
	orc_rng rng;
	int warm_mode;
	orc_cg_stats cg;
	double U;
	char err[512];
};

/* ------------------------------------------------------------------ */
static int set_err(orc_sim *s, const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(s->err, sizeof(s->err), fmt, ap);
	va_end(ap);
	return -1;
}

const char *orc_last_error(orc_sim *s) { return s->err; }

/* ---------------- RNG: POSIX drand48 + Utils::gaussian -------------- */
void orc_rng_seed(orc_rng *r, long seed) {
	/* srand48: high 32 bits = seed, low 16 bits = 0x330E */
	r->x = (((unsigned long long) (unsigned int) seed) << 16) | 0x330Eull;
	r->has_next = 0;
	r->next = -1e6;
}

double orc_rng_drand48(orc_rng *r) {
	r->x = (0x5DEECE66Dull * r->x + 0xBull) & 0xFFFFFFFFFFFFull;
	return ldexp((double) r->x, -48);
}

/* Marsaglia polar method with the cached second variate (Utils.h:210-234) */
double orc_rng_gaussian(orc_rng *r) {
	double u, v, w, toRet;
	if(r->has_next) {
		r->has_next = 0;
		return r->next;
	}
	w = 2.;
	while(w >= 1.0) {
		u = 2. * orc_rng_drand48(r) - 1.0;
		v = 2. * orc_rng_drand48(r) - 1.0;
		w = u * u + v * v;
	}
	w = sqrt((-2. * log(w)) / w);
	toRet = u * w;
	r->next = v * w;
	r->has_next = 1;
	return toRet;
}

void orc_trig_table(int L, int which, double *out) {
	/* MultiPhaseField.cpp:58-65: cos(2*PI*i/LsubX) and sin(...) of the same
	 * float argument; the reference build merges the pair into one glibc
	 * sincos call, which is what is evaluated here. */
	for(int i = 0; i < L; i++) {
		float a = 2 * PI_F * i / L;
		double sv, cv;
		sincos((double) a, &sv, &cv);
		out[i] = which == 0 ? cv : sv;
	}
}

/* ---------------- SquareBox (SquareBox.cpp:31-45, 71-76) ------------ */
static int box_x(int side, int site, int distX) {
	int x = (site - ((site / side) * side)) + distX;
	while(x < 0) x += side;
	while(x >= side) x -= side;
	return x;
}
static int box_y(int side, int site, int distY) {
	int y = (site / side) + distY;
	while(y < 0) y += side;
	while(y >= side) y -= side;
	return y;
}
static int box_elem(int side, int site, int dX, int dY) {
	return box_x(side, site, dX) + box_y(side, site, dY) * side;
}
static void min_image(int side, const double *v1, const double *v2, double *out) {
	out[0] = v2[0] - v1[0] - rint((v2[0] - v1[0]) / side) * side;
	out[1] = v2[1] - v1[1] - rint((v2[1] - v1[1]) / side) * side;
}

/* BaseBox::setNeighborsPeriodic (BaseBox.cpp:29-50) and
 * MultiPhaseField::setNeighborsSubSquarePeriodic (MultiPhaseField.cpp:188-208) */
static void periodic_nbrs(int Lx, int Ly, int *nb) {
	for(int i = 0; i < Lx * Ly; i++) {
		int y = i / Lx, x = i - (i / Lx) * Lx, ss = 0;
		for(int j = -1; j <= 1; j++)
			for(int k = -1; k <= 1; k++) {
				int xx = x + k, yy = y + j;
				if(x == 0 && k == -1) xx = Lx - 1;
				else if(x == Lx - 1 && k == 1) xx = 0;
				if(y == 0 && j == -1) yy = Ly - 1;
				else if(y == Ly - 1 && j == 1) yy = 0;
				nb[ss + i * 9] = xx + yy * Lx;
				ss++;
			}
	}
}

/* ---------------- MultiPhaseField ------------------------------------ */
static void *xrealloc_keep(void *p, size_t old_n, size_t new_n, size_t sz) {
	/* std::vector::resize: keep the first min(old,new) elements, zero the rest */
	void *q = realloc(p, new_n * sz ? new_n * sz : sz);
	if(new_n > old_n) memset((char *) q + old_n * sz, 0, (new_n - old_n) * sz);
	return q;
}

static void cell_resizing(ocell *p, int oldS, int newS) {
	p->phi = xrealloc_keep(p->phi, oldS, newS, sizeof(double));
	p->phi_old = xrealloc_keep(p->phi_old, oldS, newS, sizeof(double));
	p->dphi_old = xrealloc_keep(p->dphi_old, oldS, newS, sizeof(double));
	p->mu = xrealloc_keep(p->mu, oldS, newS, sizeof(double));
	p->dx = xrealloc_keep(p->dx, oldS, newS, sizeof(double));
	p->dy = xrealloc_keep(p->dy, oldS, newS, sizeof(double));
	p->lap = xrealloc_keep(p->lap, oldS, newS, sizeof(double));
	p->rep = xrealloc_keep(p->rep, oldS, newS, sizeof(double));
	p->pressure = xrealloc_keep(p->pressure, oldS, newS, sizeof(double));
	p->fd = xrealloc_keep(p->fd, oldS, newS, sizeof(double));
	p->map = xrealloc_keep(p->map, oldS, newS, sizeof(int));
	free(p->nbr);
	p->nbr = malloc(sizeof(int) * 9 * (size_t) newS);
	free(p->cosx); free(p->sinx); free(p->cosy); free(p->siny);
	p->cosx = malloc(sizeof(double) * p->LX);
	p->sinx = malloc(sizeof(double) * p->LX);
	p->cosy = malloc(sizeof(double) * p->LY);
	p->siny = malloc(sizeof(double) * p->LY);
	orc_trig_table(p->LX, 0, p->cosx);
	orc_trig_table(p->LX, 1, p->sinx);
	orc_trig_table(p->LY, 0, p->cosy);
	orc_trig_table(p->LY, 1, p->siny);
}

static void cell_free(ocell *p) {
	free(p->phi); free(p->phi_old); free(p->dphi_old); free(p->mu);
	free(p->dx); free(p->dy); free(p->lap); free(p->rep); free(p->pressure); free(p->fd);
	free(p->map); free(p->nbr);
	free(p->cosx); free(p->sinx); free(p->cosy); free(p->siny);
}

/* GetSubIndex (MultiPhaseField.cpp:348-350) */
static int sub_index(const orc_sim *s, const ocell *p, int q) {
	return box_elem(s->L, p->corner, ((q - (q / p->LX) * p->LX) + p->offset[0]) % p->LX,
			((q / p->LX) + p->offset[1]) % p->LY);
}

static void set_properties_to_zero(ocell *p) {
	p->S00 = 0; p->S01 = 0;
	p->x_sub_left = p->LX;
	p->y_sub_bottom = p->LY;
	p->area = 0; p->perimeter = 0; p->sumF = 0;
	p->CoM[0] = 0.; p->CoM[1] = 0.;
}

static void check_borders(ocell *p, int q) {
	int y = q / p->LX;
	int x = q - y * p->LX;
	x = (x + p->offset[0]) % p->LX;
	y = (y + p->offset[1]) % p->LY;
	if(x < p->x_sub_left) p->x_sub_left = x;
	if(y < p->y_sub_bottom) p->y_sub_bottom = y;
}

/* MultiPhaseField::set_positions(BaseBox*) (MultiPhaseField.cpp:223-310) */
static int set_positions(orc_sim *s, ocell *p) {
	const int side = s->L;
	if(p->x_sub_left >= BORDER && p->y_sub_bottom >= BORDER) {
		int site = box_elem(side, p->corner, (int) p->CoM[0], (int) p->CoM[1]);
		p->CoM[1] = (int) (site / side) + (p->CoM[1] - (int) p->CoM[1]);
		p->CoM[0] = (int) (site - (int) (site / side) * side) + (p->CoM[0] - (int) p->CoM[0]);
		double corner_old[2] = { (double) box_x(side, p->corner, 0), (double) box_y(side, p->corner, 0) };
		p->corner_old = p->corner;
		p->corner = box_elem(side, site, (int) -p->LX / 2, (int) -p->LY / 2);
		double corner_new[2] = { (double) box_x(side, p->corner, 0), (double) box_y(side, p->corner, 0) };
		double d[2];
		min_image(side, corner_old, corner_new, d);
		p->unrap[0] += d[0];
		p->unrap[1] += d[1];
		p->offset[0] = (int) (p->offset[0] + p->LX - d[0]) % p->LX;
		p->offset[1] = (int) (p->offset[1] + p->LY - d[1]) % p->LY;
		return 0;
	}
	/* patch growth / reshaping */
	int LX = p->LX, LY = p->LY;
	int nLX = LX - (2 * p->x_sub_left) + (2 * BORDER);
	int nLY = LY - (2 * p->y_sub_bottom) + (2 * BORDER);
	int nS = nLX * nLY;
	int ncorner = box_elem(side, p->corner, -(-(2 * p->x_sub_left) + (2 * BORDER)) / 2,
			-(-(2 * p->y_sub_bottom) + (2 * BORDER)) / 2);
	double *nf = calloc((size_t) nS, sizeof(double));
	p->area = 0;
	int sy, sx, dimy, dimx;
	if(nLY < LY) { sy = (LY - nLY) / 2; dimy = nLY; } else { sy = 0; dimy = LY; }
	if(nLX < LX) { sx = (LX - nLX) / 2; dimx = nLX; } else { sx = 0; dimx = LX; }
	for(int i = sy; i < dimy; i++) {
		int yy = box_y(side, p->corner, i);
		int off_y = i - p->offset[1];
		if(off_y < 0) off_y += LY;
		for(int j = sx; j < dimx; j++) {
			int xx = box_x(side, p->corner, j);
			int dx = (ncorner - (ncorner / side) * side) - xx;
			if(dx > side / 2) dx -= side;
			else if(dx < -side / 2) dx += side;
			int dy = (ncorner / side) - yy;
			if(dy > side / 2) dy -= side;
			else if(dy < -side / 2) dy += side;
			int off_x = j - p->offset[0];
			if(off_x < 0) off_x += LX;
			int dst = abs(dx) + abs(dy) * nLX;
			if(dst < 0 || dst >= nS) {
				free(nf);
				return set_err(s, "patch growth of cell %d writes outside the new patch", p->index);
			}
			double v = p->phi[off_x + off_y * LX];
			nf[dst] = v;
			p->area += v * v;
		}
	}
	double c0[2] = { (double) box_x(side, p->corner, 0), (double) box_y(side, p->corner, 0) };
	double c1[2] = { (double) box_x(side, ncorner, 0), (double) box_y(side, ncorner, 0) };
	double d[2];
	min_image(side, c0, c1, d);
	p->unrap[0] += d[0];
	p->unrap[1] += d[1];
	int oldS = p->S;
	p->LX = nLX; p->LY = nLY;
	p->corner_old = p->corner;
	p->corner = ncorner;
	p->offset[0] = 0; p->offset[1] = 0;
	p->S = nLX * nLY;
	cell_resizing(p, oldS, p->S);
	memcpy(p->phi, nf, sizeof(double) * (size_t) nS);
	free(nf);
	periodic_nbrs(p->LX, p->LY, p->nbr);
	int site = box_elem(side, p->corner, p->LX / 2, p->LY / 2);
	p->CoM[1] = (int) (site / side);
	p->CoM[0] = (int) (site - (int) (site / side) * side);
	return 0;
}

/* ---------------- WetPolarModel -------------------------------------- */
orc_sim *orc_create(int L, int N, const orc_params *prm, long seed) {
	orc_sim *s = calloc(1, sizeof(orc_sim));
	if(!(prm->wall_slip > 0)) {
		/* getWalls()==0 in a periodic box; wall_slip<=0 would mask every stencil */
		snprintf(s->err, sizeof(s->err), "wall_slip must be > 0 for a periodic square box");
	}
	s->L = L;
	s->N = N;
	s->act = *prm;
	/* constructor values before apply_changes_after_equilibration
	 * (WetPolarModel.cpp:4-21) */
	s->zetaQ_self = 0; s->zetaQ_inter = 0; s->J_Q = 0;
	s->friction = 1.; s->friction_cell = 0.; s->beta = 0.;
	s->p0 = 12.;
	s->a0 = PI_F * prm->R * prm->R;           /* float arithmetic, WetPolarModel.cpp:50 */
	s->tolerance = prm->CGtolerance > 0 ? prm->CGtolerance : DBL_EPSILON;
	s->c = calloc((size_t) N, sizeof(ocell));
	s->start = calloc((size_t) N, sizeof(int));
	s->nematic = prm->model == ORC_WETNEMATIC;
	orc_rng_seed(&s->rng, seed);
	for(int i = 0; i < N; i++) {             /* read_topology, WetPolarModel.cpp:69-75 */
		s->c[i].index = i;
		/* WetModel::read_topology (WetModel.cpp:67-81) draws nothing */
		if(!s->nematic) s->c[i].run_clock = (int) (2 * orc_rng_drand48(&s->rng) * prm->J_Q);
	}
	s->bnbr = malloc(sizeof(int) * 9 * (size_t) L * L);
	periodic_nbrs(L, L, s->bnbr);
	s->phi2 = calloc((size_t) L * L, sizeof(double));
	s->sum_phi = calloc((size_t) L * L, sizeof(double));
	s->sq00 = calloc((size_t) L * L, sizeof(double));
	s->sq01 = calloc((size_t) L * L, sizeof(double));
	s->store_size = calloc((size_t) L * L, sizeof(int));
	s->store_idx = calloc((size_t) L * L * STORE_MAX, sizeof(int));
	return s;
}

void orc_destroy(orc_sim *s) {
	if(!s) return;
	for(int i = 0; i < s->N; i++) cell_free(&s->c[i]);
	free(s->c); free(s->start); free(s->bnbr); free(s->phi2); free(s->sum_phi); free(s->sq00); free(s->sq01);
	free(s->walls);
	free(s->store_size); free(s->store_idx);
	free(s->vx); free(s->vy); free(s->fx); free(s->fy);
	free(s);
}

int orc_set_cell(orc_sim *s, int i, int LX, int LY, double comx, double comy, int offx, int offy,
		int corner, double unrapx, double unrapy, double Q00, double Q01, const double *phi) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	/* MultiPhaseField::init(Lx,Ly) (MultiPhaseField.cpp:68-107) */
	int oldS = p->S;
	p->LX = LX; p->LY = LY; p->S = LX * LY;
	p->x_sub_left = LX; p->y_sub_bottom = LY;
	cell_resizing(p, oldS, p->S);
	p->S00 = 0.; p->S01 = 0.;
	p->thetaQ = p->thetaQ_old = PI_F * (1 - 2 * orc_rng_drand48(&s->rng));
	{   /* the director of this draw (MultiPhaseField.cpp:80-95); the reader
	     * below replaces Q00 / Q01 but not nemQ (SimBackend.cpp:283-286) */
		double q01, q00;
		sincos(2 * p->thetaQ, &q01, &q00);
		double nemQ_mod = sqrt(q00 * q00 + q01 * q01);
		double nx = sqrt((1 + q00 / nemQ_mod) / 2);
		double sgn = q01 > 0 ? 1 : q01 < 0 ? -1 : 0;
		double ny = sgn * sqrt((1 - q00 / nemQ_mod) / 2);
		p->nem[0] = p->nem_old[0] = nx;
		p->nem[1] = p->nem_old[1] = ny;
	}
	p->area = 0; p->perimeter = 0; p->sumF = 0;
	periodic_nbrs(LX, LY, p->nbr);
	/* SimBackend.cpp:274-296 */
	p->CoM[0] = comx; p->CoM[1] = comy;
	p->CoM_old[0] = comx; p->CoM_old[1] = comy;
	p->unrap[0] = unrapx; p->unrap[1] = unrapy;
	p->corner = corner; p->corner_old = corner;
	p->offset[0] = offx; p->offset[1] = offy;
	p->Q00 = Q00; p->Q01 = Q01;
	p->thetaQ = 0.5 * atan2(2 * p->Q01, 2 * p->Q00);
	p->thetaQ_old = p->thetaQ;
	for(int k = 0; k < p->S; k++) {
		p->map[k] = sub_index(s, p, k);
		p->phi[k] = phi[k];
		p->area += phi[k] * phi[k];
		p->sumF += phi[k];
	}
	return 0;
}

static void resize_rows(orc_sim *s, int rows) {
	if(rows > s->rows_cap) {
		s->vx = realloc(s->vx, sizeof(double) * rows);
		s->vy = realloc(s->vy, sizeof(double) * rows);
		s->fx = realloc(s->fx, sizeof(double) * rows);
		s->fy = realloc(s->fy, sizeof(double) * rows);
		s->rows_cap = rows;
	}
}

/* ---------------- walls and external forces --------------------------- */
int orc_set_walls(orc_sim *s, const double *walls) {
	free(s->walls);
	s->walls = NULL;
	if(walls) {
		s->walls = malloc(sizeof(double) * (size_t) s->L * s->L);
		memcpy(s->walls, walls, sizeof(double) * (size_t) s->L * s->L);
	}
	return 0;
}

int orc_add_force(orc_sim *s, int kind, double kappa_wall, double lambda, double shear_rate) {
	if(s->nforce >= ORC_MAX_FORCES) return set_err(s, "at most %d external forces", ORC_MAX_FORCES);
	if(kind < ORC_FORCE_CHANNEL_WALLS || kind > ORC_FORCE_STRETCHING) return set_err(s, "unknown force kind %d", kind);
	s->force[s->nforce].kind = kind;
	s->force[s->nforce].kappa_wall = kappa_wall;
	s->force[s->nforce].lambda = lambda;
	s->force[s->nforce].rate_active = shear_rate;
	s->force[s->nforce].rate = 0.;      /* shear_rate is 0 until equilibration ends */
	s->nforce++;
	return 0;
}

static inline double get_walls(const orc_sim *s, int k) { return s->walls ? s->walls[k] : 0.; }
static inline int open_site(const orc_sim *s, int k) { return get_walls(s, k) < s->act.wall_slip; }

/* free_energy / free_energy_density of one force (ChannelWalls.cpp:26-31,
 * ShearFlowChannel.cpp:33-39, PoiseuilleFlowChannel.cpp, Stretching.cpp) */
static double force_free_energy(const orc_sim *s, int f, double phi, double walls) {
	double kappa_wall = s->force[f].kappa_wall, lambda = s->force[f].lambda;
	switch(s->force[f].kind) {
	case ORC_FORCE_CHANNEL_WALLS: return 2.0 * kappa_wall / lambda * phi * walls * walls;
	case ORC_FORCE_SHEAR_FLOW:
	case ORC_FORCE_POIS_FLOW: return 4.0 * kappa_wall / lambda * phi * walls * walls;
	default: return 0.;
	}
}
static double force_free_energy_density(const orc_sim *s, int f, double phi, double walls) {
	double kappa_wall = s->force[f].kappa_wall, lambda = s->force[f].lambda;
	if(s->force[f].kind == ORC_FORCE_SHEAR_FLOW) return 2.0 * kappa_wall / lambda * phi * phi * walls * walls;
	return 0.;
}

/* BaseField::set_F_ext (BaseField.h:54-66): the forces' free energy, also
 * added to Repulsion; 0 without forces */
static double set_F_ext(const orc_sim *s, ocell *p, int q, int k) {
	double F_ext = 0.;
	for(int f = 0; f < s->nforce; f++) {
		double force_calc = force_free_energy(s, f, p->phi[q], get_walls(s, k));
		F_ext += force_calc;
		p->rep[q] += force_calc;
	}
	return F_ext;
}
static double set_f_density_ext(const orc_sim *s, ocell *p, int q, int k) {
	double f_density_ext = 0.;
	for(int f = 0; f < s->nforce; f++) f_density_ext += force_free_energy_density(s, f, p->phi[q], get_walls(s, k));
	return f_density_ext;
}

/* BaseField::set_V_ext (BaseField.h:107-115): the last force's velocity
 * profile at box site k (ShearFlowChannel.cpp:41-45,
 * PoiseuilleFlowChannel.cpp, Stretching.cpp: of the cell index on a 10 x 10
 * grid); ChannelWalls has none */
static void set_V_ext(const orc_sim *s, const ocell *p, int k, double *V) {
	for(int f = 0; f < s->nforce; f++) {
		double lambda = s->force[f].lambda, shear_rate = s->force[f].rate;
		int lx = s->L, ly = s->L;
		V[0] = 0.;
		V[1] = 0.;
		if(s->force[f].kind == ORC_FORCE_SHEAR_FLOW) {
			int y = (k / lx);
			if(y > lambda && y < ly - lambda)
				V[0] = ((y - (lambda + 1) + 0.5) - 0.5 * (ly - 2 * (lambda + 1))) * shear_rate / (ly - 2 * (lambda + 1));
		}
		else if(s->force[f].kind == ORC_FORCE_POIS_FLOW) {
			int y = (k / lx);
			if(y > lambda && y < ly - lambda)
				V[0] = 4 * (y - (lambda + 1)) * shear_rate * (1 - (y - (lambda + 1)) / (ly - 2 * (lambda + 1))) / (ly - 2 * (lambda + 1));
		}
		else if(s->force[f].kind == ORC_FORCE_STRETCHING) {
			int kk = p->index, y = (kk / 10), x = kk - y * 10;
			if(x == 0) V[0] = -shear_rate;
			else if(x == 10 - 1) V[0] = shear_rate;
		}
	}
}

/* BaseInteraction::derivX/derivY/Laplacian (BaseInteraction.h:196-263): a
 * difference across a wall site (getWalls >= wall_slip) is dropped, a
 * one-sided Laplacian doubles the open side */
static inline double derivX(const orc_sim *s, const ocell *p, int q, int k) {
	const int *bn = s->bnbr + 9 * k;
	if(open_site(s, bn[5]) && open_site(s, bn[3]))
		return .5 * (p->phi[p->nbr[5 + q * 9]] - p->phi[p->nbr[3 + q * 9]]);
	return 0.;
}
static inline double derivY(const orc_sim *s, const ocell *p, int q, int k) {
	const int *bn = s->bnbr + 9 * k;
	if(open_site(s, bn[7]) && open_site(s, bn[1]))
		return .5 * (p->phi[p->nbr[7 + q * 9]] - p->phi[p->nbr[1 + q * 9]]);
	return 0.;
}
static inline double laplacian(const orc_sim *s, const ocell *p, int q, int k) {
	const int *bn = s->bnbr + 9 * k;
	const double *f = p->phi;
	const int *nb = p->nbr + 9 * q;
	const int N = open_site(s, bn[7]), S = open_site(s, bn[1]), E = open_site(s, bn[5]), W = open_site(s, bn[3]);
	if(N && S) {
		if(E && W) return f[nb[7]] + f[nb[1]] + f[nb[5]] + f[nb[3]] - 4 * f[q];
		else if(E) return f[nb[7]] + f[nb[1]] + 2 * f[nb[5]] - 4 * f[q];
		else if(W) return f[nb[7]] + f[nb[1]] + 2 * f[nb[3]] - 4 * f[q];
		else return f[nb[7]] + f[nb[1]] - 2 * f[q];
	}
	else if(N) {
		if(E && W) return 2 * f[nb[7]] + f[nb[5]] + f[nb[3]] - 4 * f[q];
		else if(E) return 2 * f[nb[7]] + 2 * f[nb[5]] - 4 * f[q];
		else if(W) return 2 * f[nb[7]] + 2 * f[nb[3]] - 4 * f[q];
		else return 2 * f[nb[7]] - 2 * f[q];
	}
	else if(S) {
		if(E && W) return 2 * f[nb[1]] + f[nb[5]] + f[nb[3]] - 4 * f[q];
		else if(E) return 2 * f[nb[1]] + 2 * f[nb[5]] - 4 * f[q];
		else if(W) return 2 * f[nb[1]] + 2 * f[nb[3]] - 4 * f[q];
		else return 2 * f[nb[1]] - 2 * f[q];
	}
	return 0.;
}

/* begin_energy_computation() without arguments (WetPolarModel.cpp:113-147) */
static void init_field_properties(orc_sim *s) {
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) {
			p->tf[0] = p->tf[1] = p->tr[0] = p->tr[1] = 0.;
			double dx = derivX(s, p, q, p->map[q]), dy = derivY(s, p, q, p->map[q]);
			/* BaseInteraction::updateFieldProperties (BaseInteraction.cpp:62-69) */
			p->mu[q] = set_F_ext(s, p, q, p->map[q]);
			p->fd[q] = set_f_density_ext(s, p, q, p->map[q]);
			p->dx[q] = dx;
			p->dy[q] = dy;
			p->S00 += -0.5 * (dx * dx - dy * dy);
			p->S01 += -dx * dy;
		}
	}
}

/* The terms WetPolarModel::f_interaction (WetPolarModel.cpp:262-301) and
 * WetModel::f_interaction (WetModel.cpp:488-600) share: Cahn-Hilliard, area
 * constraint, repulsion and adhesion */
typedef struct fterms { double ph, p2, CH, A, Rep, Adh; } fterms;

static fterms free_energy_terms(orc_sim *s, ocell *p, int q) {
	int k = p->map[q];
	const int *bn = s->bnbr + 9 * k;
	double *phi2 = s->phi2;
	double laplacianSquare;
	if(open_site(s, k)) laplacianSquare = phi2[bn[5]] + phi2[bn[7]] + phi2[bn[3]] + phi2[bn[1]] - 4. * phi2[k];
	else laplacianSquare = phi2[bn[5]] + phi2[bn[3]] - 2. * phi2[k];
	double ph = p->phi[q];
	double gamma = s->act.gamma, lambda = s->act.lambda, mu = s->act.mu, kappa = s->act.kappa,
			omega = s->act.omega, a0 = s->a0;
	fterms t;
	t.ph = ph;
	t.p2 = phi2[k];
	t.CH = gamma * (8 * ph * (1 - ph) * (1 - 2 * ph) / lambda - 2 * lambda * p->lap[q]);
	t.A = -4 * (mu / a0) * (1 - p->area / a0) * ph;
	t.Rep = 4 * (kappa / lambda) * ph * (phi2[k] - ph * ph);
	double lsquare = 2 * ph * p->lap[q] + 2 * (p->dx[q] * p->dx[q] + p->dy[q] * p->dy[q]);
	double suppress = (laplacianSquare - lsquare) / sqrt(1 + (laplacianSquare - lsquare) * (laplacianSquare - lsquare));
	t.Adh = -4 * lambda * omega * suppress * ph;
	return t;
}

static double f_interaction(orc_sim *s, ocell *p, int q) {
	fterms t = free_energy_terms(s, p, q);
	double P = 4 * (s->beta / s->p0) * (1 - p->perimeter / s->p0) * p->lap[q];
	double V = t.CH + t.A + t.Rep + t.Adh + P;
	p->mu[q] += V;
	p->pressure[q] = t.Rep - t.CH - t.A;
	p->rep[q] += t.Rep;
	return V;
}

static void calc_internal_forces(orc_sim *s, ocell *p, int q) {
	const int *nb = p->nbr + 9 * q;
	double f_passive_x = (-1) * 0.5 * (p->mu[nb[5]] - p->mu[nb[3]]);
	double f_passive_y = (-1) * 0.5 * (p->mu[nb[7]] - p->mu[nb[1]]);
	p->tr[0] += (-1) * 0.5 * (p->rep[nb[5]] - p->rep[nb[3]]) * p->phi[q];
	p->tr[1] += (-1) * 0.5 * (p->rep[nb[7]] - p->rep[nb[1]]) * p->phi[q];
	double lambda = s->act.lambda;
	double Theta = (-p->dx[q] * p->Q00) < 0 ? 0.0 : 1.0;
	double fQx = lambda * (-p->dx[q] * p->Q00) * Theta * p->Q00;
	Theta = (-p->dy[q] * p->Q01) < 0 ? 0.0 : 1.0;
	double fQy = lambda * (-p->dy[q] * p->Q01) * Theta * p->Q01;
	int row = q + s->start[p->index];
	if(open_site(s, p->map[q])) {
		s->fx[row] = f_passive_x * s->act.passive_alpha + fQx * s->zetaQ_self;
		s->fy[row] = f_passive_y * s->act.passive_alpha + fQy * s->zetaQ_self;
		p->tf[0] += s->fx[row] * p->phi[q];
		p->tf[1] += s->fy[row] * p->phi[q];
	}
	else {
		s->fx[row] = 0.;
		s->fy[row] = 0.;
	}
}

/* WetModel::f_interaction (WetModel.cpp:488-600): no perimeter term; the
 * free-energy density of each term is accumulated for the passive stress */
static double f_interaction_nem(orc_sim *s, ocell *p, int q) {
	fterms t = free_energy_terms(s, p, q);
	double gamma = s->act.gamma, lambda = s->act.lambda, mu = s->act.mu, kappa = s->act.kappa, a0 = s->a0;
	double ph = t.ph;
	double CH_density = (gamma / lambda) * 4 * ph * ph * (1 - ph) * (1 - ph)
			+ gamma * lambda * (p->dx[q] * p->dx[q] + p->dy[q] * p->dy[q]);
	double A_density = -mu / a0 * ph * ph;
	double Rep_density = (kappa / lambda) * ph * ph * (t.p2 - ph * ph);
	double V = t.CH + t.A + t.Rep + t.Adh;
	p->mu[q] += V;
	p->pressure[q] = t.Rep - t.CH - t.A;
	p->fd[q] += CH_density + A_density + Rep_density;
	return V;
}

/* WetModel::calc_internal_forces (WetModel.cpp:603-641): the passive force is
 * minus the divergence of the stress (f - mu phi) I + grad phi (x) df/d(grad phi)
 * with df/d(grad phi) = 2 gamma lambda grad phi (freeEnergyDensityGradient,
 * :597-598); the active forces are the cell's own nematic force and the one of
 * its neighbours from the box sums of phi Q */
static void calc_internal_forces_nem(orc_sim *s, ocell *p, int q) {
	const int *nb = p->nbr + 9 * q;
	const int *bn = s->bnbr + 9 * p->map[q];
	const double gl = 2 * s->act.gamma * s->act.lambda;
	const double *fd = p->fd, *mu = p->mu, *ph = p->phi, *dx = p->dx, *dy = p->dy;
#define GX(n) (gl * dx[n])
#define GY(n) (gl * dy[n])
	double f_passive_x = -0.5 * ((fd[nb[5]] - mu[nb[5]] * ph[nb[5]] + dx[nb[5]] * GX(nb[5]))
			- (fd[nb[3]] - mu[nb[3]] * ph[nb[3]] + dx[nb[3]] * GX(nb[3])))
			- 0.5 * ((dx[nb[7]] * GY(nb[7])) - (dx[nb[1]] * GY(nb[1])));
	double f_passive_y = -0.5 * ((fd[nb[7]] - mu[nb[7]] * ph[nb[7]] + dy[nb[7]] * GY(nb[7]))
			- (fd[nb[1]] - mu[nb[1]] * ph[nb[1]] + dy[nb[1]] * GY(nb[1])))
			- 0.5 * ((dy[nb[5]] * GX(nb[5])) - (dy[nb[3]] * GX(nb[3])));
#undef GX
#undef GY
	double fQ_self_x = -(p->Q00 * dx[q] + p->Q01 * dy[q]);
	double fQ_self_y = -(p->Q01 * dx[q] - p->Q00 * dy[q]);
	double fQ_inter_x = -(0.5 * (s->sq00[bn[5]] - s->sq00[bn[3]]) + 0.5 * (s->sq01[bn[7]] - s->sq01[bn[1]])) - fQ_self_x;
	double fQ_inter_y = -(0.5 * (s->sq01[bn[5]] - s->sq01[bn[3]]) - 0.5 * (s->sq00[bn[7]] - s->sq00[bn[1]])) - fQ_self_y;
	int row = q + s->start[p->index];
	if(open_site(s, p->map[q])) {
		s->fx[row] = f_passive_x * s->act.passive_alpha + fQ_self_x * s->zetaQ_self + fQ_inter_x * s->zetaQ_inter;
		s->fy[row] = f_passive_y * s->act.passive_alpha + fQ_self_y * s->zetaQ_self + fQ_inter_y * s->zetaQ_inter;
	}
	else {
		s->fx[row] = 0.;
		s->fy[row] = 0.;
	}
}

/* matrix-free form of the assembled operator (see orc_assemble_dense):
 * (A v)_i = D_i v_i - eta * sum_{j in 1,3,5,7} sum_{l at nbr_j(k_i)} v_l */
static void compute_diag(orc_sim *s, double *diag) {
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) {
			int row = q + s->start[i];
			int k = p->map[q];
			double d = s->friction;
			static const int nv[4] = { 1, 3, 5, 7 };
			for(int a = 0; a < 4; a++) {
				int o = s->bnbr[nv[a] + k * 9];
				for(int l = 0; l < s->store_size[o]; l++) d += s->friction_cell * 1.0;
			}
			diag[row] = d;
		}
	}
}

static void apply_A_mf(orc_sim *s, const double *diag, const double *x, double *y) {
	static const int nv[4] = { 1, 3, 5, 7 };
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) {
			int row = q + s->start[i];
			int k = p->map[q];
			double acc = diag[row] * x[row];
			for(int a = 0; a < 4; a++) {
				int o = s->bnbr[nv[a] + k * 9];
				for(int l = 0; l < s->store_size[o]; l++)
					acc += (-s->friction_cell * 1.0) * x[s->store_idx[l + o * STORE_MAX]];
			}
			y[row] = acc;
		}
	}
}

static double dotv(const double *a, const double *b, int n) {
	double acc = 0.;
	for(int i = 0; i < n; i++) acc += a[i] * b[i];
	return acc;
}

/* Eigen::ConjugateGradient (Jacobi preconditioner) on x = guess */
static int cg_solve(orc_sim *s, const double *diag, const double *b, double *x, int *iters_out, double *relres) {
	int n = s->size_rows;
	double *r = malloc(sizeof(double) * n), *pp = malloc(sizeof(double) * n),
			*z = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n), *inv = malloc(sizeof(double) * n);
	for(int i = 0; i < n; i++) inv[i] = 1. / diag[i];
	apply_A_mf(s, diag, x, t);
	for(int i = 0; i < n; i++) r[i] = b[i] - t[i];
	double rhsNorm2 = dotv(b, b, n);
	int iters = 0;
	if(rhsNorm2 == 0) {
		for(int i = 0; i < n; i++) x[i] = 0.;
		*relres = 0.;
		goto done;
	}
	double tol = s->tolerance;
	double threshold = tol * tol * rhsNorm2;
	if(threshold < DBL_MIN) threshold = DBL_MIN;
	double rn2 = dotv(r, r, n);
	if(rn2 < threshold) {
		*relres = sqrt(rn2 / rhsNorm2);
		goto done;
	}
	for(int i = 0; i < n; i++) pp[i] = inv[i] * r[i];
	double absNew = dotv(r, pp, n);
	int maxIters = s->act.cg_max_iter > 0 ? s->act.cg_max_iter : 2 * n;
	int i = 0;
	while(i < maxIters) {
		apply_A_mf(s, diag, pp, t);
		double alpha = absNew / dotv(pp, t, n);
		for(int j = 0; j < n; j++) x[j] += alpha * pp[j];
		for(int j = 0; j < n; j++) r[j] -= alpha * t[j];
		rn2 = dotv(r, r, n);
		iters++;
		if(rn2 < threshold) break;
		for(int j = 0; j < n; j++) z[j] = inv[j] * r[j];
		double absOld = absNew;
		absNew = dotv(r, z, n);
		double beta = absNew / absOld;
		for(int j = 0; j < n; j++) pp[j] = z[j] + beta * pp[j];
		i++;
	}
	*relres = sqrt(rn2 / rhsNorm2);
done:
	*iters_out = iters;
	free(r); free(pp); free(z); free(t); free(inv);
	return 0;
}

/* WetPolarModel::begin_energy_computation(fields) (WetPolarModel.cpp:149-234) */
static int compute_forces(orc_sim *s) {
	int LL = s->L * s->L;
	memset(s->store_size, 0, sizeof(int) * LL);
	/* cold solve after 10 warm ones (WetPolarModel.cpp:152-157), 100 for
	 * WetModel (WetModel.cpp:196-201) */
	if(s->restart_solver < (s->nematic ? 100 : 10) || s->warm_mode == 1) s->size_rows_old = s->size_rows;
	else {
		s->size_rows_old = 0;
		s->restart_solver = 0;
	}
	s->size_rows = 0;
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		s->start[i] = s->size_rows;
		s->size_rows += p->S;
		for(int q = 0; q < p->S; q++) {     /* computeGlobalSums */
			int k = sub_index(s, p, q);
			s->phi2[k] += p->phi[q] * p->phi[q];
			s->sum_phi[k] += p->phi[q];
			if(s->nematic) {           /* WetModel::computeGlobalSums (WetModel.cpp:423-431) */
				s->sq00[k] += p->phi[q] * p->Q00;
				s->sq01[k] += p->phi[q] * p->Q01;
			}
			p->map[q] = k;
			p->dx[q] = derivX(s, p, q, k);
			p->dy[q] = derivY(s, p, q, k);
			p->lap[q] = laplacian(s, p, q, k);
			if(!s->nematic) p->perimeter += p->dx[q] * p->dx[q] + p->dy[q] * p->dy[q];
			p->tf[0] = p->tf[1] = p->tr[0] = p->tr[1] = 0.;
			p->mu[q] = set_F_ext(s, p, q, k);
			p->fd[q] = set_f_density_ext(s, p, q, k);
			if(s->store_size[k] >= STORE_MAX)
				return set_err(s, "Too many field patches overlap: %d, %d, %d, %d", p->index, k, q, p->corner);
			s->store_idx[s->store_size[k] + k * STORE_MAX] = q + s->start[i];
			s->store_size[k]++;
		}
	}
	int cold = s->size_rows != s->size_rows_old;
	if(cold) {
		/* Eigen VectorXd::resize loses the contents */
		resize_rows(s, s->size_rows);
		for(int r = 0; r < s->size_rows; r++) s->vx[r] = s->vy[r] = 0.;
	}
	s->U = 0;
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) s->U += s->nematic ? f_interaction_nem(s, p, q) : f_interaction(s, p, q);
	}
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) {
			if(s->nematic) calc_internal_forces_nem(s, p, q);
			else calc_internal_forces(s, p, q);
			if(s->friction_cell == 0) {
				int row = q + s->start[i];
				s->vx[row] = s->fx[row] / s->friction;
				s->vy[row] = s->fy[row] / s->friction;
			}
		}
	}
	if(s->friction_cell != 0) {
		if(s->walls) return set_err(s, "the wet solve with walls is not restated");
		double *diag = malloc(sizeof(double) * s->size_rows);
		compute_diag(s, diag);
		int itx, ity;
		double rx, ry;
		if(cold) {
			for(int r = 0; r < s->size_rows; r++) s->vx[r] = s->vy[r] = 0.;
			s->restart_solver = 0;
		}
		else
			s->restart_solver += 1;
		cg_solve(s, diag, s->fx, s->vx, &itx, &rx);
		cg_solve(s, diag, s->fy, s->vy, &ity, &ry);
		s->cg.solves += 2;
		s->cg.iters_total += itx + ity;
		s->cg.last_iters[0] = itx;
		s->cg.last_iters[1] = ity;
		s->cg.last_relres[0] = rx;
		s->cg.last_relres[1] = ry;
		s->cg.last_cold = cold;
		free(diag);
	}
	return 0;
}

/* WetPolarModel::updateDirectedActiveForces (WetPolarModel.cpp:342-414) */
static void update_directed_active_forces(orc_sim *s, ocell *p, int store) {
	double dt = s->act.dt;
	if(store) p->thetaQ_old = p->thetaQ + sqrt(dt) * s->zetaQ_inter * orc_rng_gaussian(&s->rng);
	double f0 = p->tr[0], f1 = p->tr[1];
	p->thetaQ = p->thetaQ_old - dt * s->J_Q * (atan2(f0 * p->Q01 - f1 * p->Q00, f0 * p->Q00 + f1 * p->Q01));
	p->Q00 = cos(p->thetaQ);
	p->Q01 = sin(p->thetaQ);
	if(s->act.anchoring) {                 /* update_anchoring, WetPolarModel.cpp:417-430 */
		double walls_length = 8;
		double dist1 = ((double) p->CoM_old[1] - walls_length);
		double dist2 = ((double) s->L - walls_length) - (double) p->CoM_old[1];
		double theta;
		if(dist1 < dist2) theta = (PI_F / 2) - (PI_F / 2) * dist1 / (((double) s->L - 2 * walls_length) / 2);
		else theta = (PI_F / 2) + ((PI_F / 2) * dist2 / (((double) s->L - 2 * walls_length) / 2));
		double n0 = cos(theta), n1 = sin(theta);
		p->Q00 = 0.5 * (n0 * n0 - n1 * n1);
		p->Q01 = n0 * n1;
	}
}

/* WetModel::updateDirectedActiveForces (WetModel.cpp:644-700): the director
 * relaxes towards the long axis of the shape tensor; update_anchoring
 * (:703-718) */
static void update_directed_nem(orc_sim *s, ocell *p, int store) {
	double dt = s->act.dt;
	double S_mod = sqrt(p->S00 * p->S00 + p->S01 * p->S01);
	if(store) {
		p->nem_old[0] = p->nem[0];
		p->nem_old[1] = p->nem[1];
	}
	if(S_mod > 0.00000001) {
		double t = 0.5 * atan2(p->S01, p->S00);
		double d1, d0;
		sincos(t, &d1, &d0);
		double sgn = (d0 * p->nem[0] + d1 * p->nem[1] > 0.0) ? 1.0 : -1.0;
		p->nem[0] = p->nem_old[0] + dt * s->J_Q * (sgn * d0 - p->nem[0]);
		p->nem[1] = p->nem_old[1] + dt * s->J_Q * (sgn * d1 - p->nem[1]);
		p->Q00 = 0.5 * (p->nem[0] * p->nem[0] - p->nem[1] * p->nem[1]);
		p->Q01 = p->nem[0] * p->nem[1];
	}
	if(s->act.anchoring) {
		double walls_length = 8;
		double dist1 = ((double) p->CoM_old[1] - walls_length);
		double dist2 = ((double) s->L - walls_length) - (double) p->CoM_old[1];
		double theta;
		if(dist1 < dist2) theta = (PI_F / 2) - (PI_F / 2) * dist1 / (((double) s->L - 2 * walls_length) / 2);
		else theta = (PI_F / 2) + ((PI_F / 2) * dist2 / (((double) s->L - 2 * walls_length) / 2));
		double n1, n0;
		sincos(theta, &n1, &n0);
		p->nem[0] = n0;
		p->nem[1] = n1;
		p->Q00 = 0.5 * (p->nem[0] * p->nem[0] - p->nem[1] * p->nem[1]);
		p->Q01 = p->nem[0] * p->nem[1];
	}
}

/* FD_CPUBackend::first_step (FD_CPUBackend.cpp:44-120) */
static int first_step(orc_sim *s, int store) {
	double J0 = s->act.J0, dt = s->act.dt;
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		double temp[4] = { 0, 0, 0, 0 };
		if(store) {
			p->CoM_old[0] = p->CoM[0];
			p->CoM_old[1] = p->CoM[1];
			p->area_old = p->area;
			p->sumF_old = p->sumF;
		}
		set_properties_to_zero(p);
		for(int q = 0; q < p->S; q++) {
			int k = p->map[q];
			int row = q + s->start[i];
			double V[2] = { 0., 0. };   /* V_ext, kept by the field between sites */
			if(s->nforce > 0) set_V_ext(s, p, k, V);
			double dphi = -J0 * p->mu[q] - s->vx[row] * p->dx[q] - s->vy[row] * p->dy[q]
					- V[0] * p->dx[q] - V[1] * p->dy[q];
			if(store) {
				p->phi_old[q] = p->phi[q];
				p->dphi_old[q] = dphi;
			}
			double phi = p->phi_old[q] + dt * .5 * (dphi + p->dphi_old[q]);
			p->phi[q] = phi;
			p->area += phi * phi;
			p->sumF += phi;
			int x = (((q - (q / p->LX) * p->LX) + p->offset[0]) % p->LX);
			int y = (((q / p->LX) + p->offset[1]) % p->LY);
			temp[0] += phi * p->cosx[x];
			temp[1] += phi * p->sinx[x];
			temp[2] += phi * p->cosy[y];
			temp[3] += phi * p->siny[y];
			if(!store && phi < 0.09 && phi > 0.01) check_borders(p, q);
			s->phi2[k] = 0;                  /* resetSums */
			s->sum_phi[k] = 0;
			s->sq00[k] = 0;
			s->sq01[k] = 0;
			p->S00 += -0.5 * (p->dx[q] * p->dx[q] - p->dy[q] * p->dy[q]);
			p->S01 += -p->dx[q] * p->dy[q];
		}
		p->CoM[0] = p->LX * (atan2(-temp[1] / p->sumF, -temp[0] / p->sumF) + PI_F) / (2 * PI_F);
		p->CoM[1] = p->LY * (atan2(-temp[3] / p->sumF, -temp[2] / p->sumF) + PI_F) / (2 * PI_F);
		if(!store)
			if(set_positions(s, p)) return -1;
		if(s->nematic) update_directed_nem(s, p, store);
		else update_directed_active_forces(s, p, store);
	}
	return 0;
}

int orc_init(orc_sim *s) {
	for(int k = 0; k < s->L * s->L; k++) {
		s->phi2[k] = 0;
		s->sum_phi[k] = 0;
		s->sq00[k] = 0;
		s->sq01[k] = 0;
		s->store_size[k] = 0;
	}
	s->restart_solver = 0;
	s->size_rows = 0;
	init_field_properties(s);
	return compute_forces(s);
}

void orc_apply_changes_after_equilibration(orc_sim *s) {
	for(int f = 0; f < s->nforce; f++) s->force[f].rate = s->force[f].rate_active;   /* set_ext_properties */
	s->zetaQ_self = s->act.zetaQ_self;
	s->zetaQ_inter = s->act.zetaQ_inter;
	s->J_Q = s->act.J_Q;
	s->friction = s->act.friction;
	s->friction_cell = s->act.friction_cell;
	s->beta = s->act.beta;
}

int orc_step(orc_sim *s, int nsteps) {
	for(int n = 0; n < nsteps; n++) {
		if(first_step(s, 1)) return -1;
		if(compute_forces(s)) return -1;
		if(first_step(s, 0)) return -1;
		if(compute_forces(s)) return -1;
	}
	return 0;
}

void orc_set_warm_mode(orc_sim *s, int mode) { s->warm_mode = mode; }

/* ---------------- accessors ------------------------------------------ */
int orc_num_cells(orc_sim *s) { return s->N; }
int orc_size_rows(orc_sim *s) { return s->size_rows; }

int orc_get_cell(orc_sim *s, int i, orc_cell_info *o) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	o->LsubX = p->LX; o->LsubY = p->LY; o->subSize = p->S;
	o->offset[0] = p->offset[0]; o->offset[1] = p->offset[1];
	o->corner = p->corner; o->corner_old = p->corner_old;
	o->x_sub_left = p->x_sub_left; o->y_sub_bottom = p->y_sub_bottom;
	o->start = s->start[i];
	o->CoM[0] = p->CoM[0]; o->CoM[1] = p->CoM[1];
	o->CoM_old[0] = p->CoM_old[0]; o->CoM_old[1] = p->CoM_old[1];
	o->unrap[0] = p->unrap[0]; o->unrap[1] = p->unrap[1];
	o->Q00 = p->Q00; o->Q01 = p->Q01; o->thetaQ = p->thetaQ; o->thetaQ_old = p->thetaQ_old;
	o->area = p->area; o->area_old = p->area_old; o->sumF = p->sumF; o->sumF_old = p->sumF_old;
	o->perimeter = p->perimeter; o->S00 = p->S00; o->S01 = p->S01;
	o->total_force[0] = p->tf[0]; o->total_force[1] = p->tf[1];
	o->total_force_repulsion[0] = p->tr[0]; o->total_force_repulsion[1] = p->tr[1];
	o->nemQ[0] = p->nem[0]; o->nemQ[1] = p->nem[1];
	o->nemQ_old[0] = p->nem_old[0]; o->nemQ_old[1] = p->nem_old[1];
	return 0;
}

int orc_get_site_array(orc_sim *s, int i, int which, double *out) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	for(int q = 0; q < p->S; q++) {
		switch(which) {
		case 0: out[q] = p->phi[q]; break;
		case 1: out[q] = p->mu[q]; break;
		case 2: out[q] = p->dx[q]; break;
		case 3: out[q] = p->dy[q]; break;
		case 4: out[q] = p->lap[q]; break;
		case 5: out[q] = p->phi_old[q]; break;
		case 6: out[q] = p->dphi_old[q]; break;
		case 7: out[q] = p->rep[q]; break;
		case 8: out[q] = (double) p->map[q]; break;
		case 9: out[q] = p->fd[q]; break;
		default: return set_err(s, "bad site array %d", which);
		}
	}
	return 0;
}

int orc_get_rows(orc_sim *s, int which, double *out) {
	double *src = which == 0 ? s->vx : which == 1 ? s->vy : which == 2 ? s->fx : which == 3 ? s->fy : NULL;
	if(!src) return set_err(s, "bad row array %d", which);
	memcpy(out, src, sizeof(double) * s->size_rows);
	return 0;
}

int orc_get_box(orc_sim *s, int which, double *out) {
	int LL = s->L * s->L;
	for(int k = 0; k < LL; k++) {
		if(which == 0) out[k] = s->phi2[k];
		else if(which == 1) out[k] = s->sum_phi[k];
		else if(which == 2) out[k] = s->store_size[k];
		else if(which == 3) out[k] = s->sq00[k];
		else if(which == 4) out[k] = s->sq01[k];
		else return set_err(s, "bad box array %d", which);
	}
	return 0;
}

void orc_get_cg_stats(orc_sim *s, orc_cg_stats *o) { *o = s->cg; }

int orc_apply_A(orc_sim *s, const double *x, double *y) {
	double *diag = malloc(sizeof(double) * s->size_rows);
	compute_diag(s, diag);
	apply_A_mf(s, diag, x, y);
	free(diag);
	return 0;
}

/* Triplet rules of WetPolarModel.cpp:200-217 summed like setFromTriplets */
int orc_assemble_dense(orc_sim *s, double *A) {
	int n = s->size_rows;
	static const int nv[4] = { 1, 3, 5, 7 };
	static const double w[9] = { 0, 1, 0, 1, 0, 1, 0, 1, 0 };
	memset(A, 0, sizeof(double) * (size_t) n * n);
	for(int i = 0; i < s->N; i++) {
		ocell *p = &s->c[i];
		for(int q = 0; q < p->S; q++) {
			int row = q + s->start[i];
			A[(size_t) row * n + row] += s->friction;
			for(int a = 0; a < 4; a++) {
				int j = nv[a];
				int o = s->bnbr[j + p->map[q] * 9];
				for(int l = 0; l < s->store_size[o]; l++) {
					A[(size_t) row * n + row] += s->friction_cell * w[j];
					A[(size_t) row * n + s->store_idx[l + o * STORE_MAX]] += -s->friction_cell * w[j];
				}
			}
		}
	}
	return 0;
}

/* ---------------- test hooks ----------------------------------------- */
void orc_test_nbrs(int LX, int LY, int *out) { periodic_nbrs(LX, LY, out); }

int orc_test_set_positions(int Lbox, int *ints, double *dbl, double *phi, int phi_cap) {
	orc_sim tmp;
	memset(&tmp, 0, sizeof(tmp));
	tmp.L = Lbox;
	ocell p;
	memset(&p, 0, sizeof(p));
	p.LX = ints[0]; p.LY = ints[1]; p.S = p.LX * p.LY;
	p.offset[0] = ints[2]; p.offset[1] = ints[3];
	p.corner = ints[4]; p.corner_old = ints[5];
	p.x_sub_left = ints[6]; p.y_sub_bottom = ints[7];
	p.unrap[0] = dbl[0]; p.unrap[1] = dbl[1];
	p.CoM[0] = dbl[2]; p.CoM[1] = dbl[3];
	p.area = dbl[4];
	cell_resizing(&p, 0, p.S);
	memcpy(p.phi, phi, sizeof(double) * p.S);
	periodic_nbrs(p.LX, p.LY, p.nbr);
	int rc = set_positions(&tmp, &p);
	if(rc == 0) {
		if(p.S > phi_cap) rc = -1;
		else {
			ints[0] = p.LX; ints[1] = p.LY; ints[2] = p.offset[0]; ints[3] = p.offset[1];
			ints[4] = p.corner; ints[5] = p.corner_old; ints[6] = p.x_sub_left; ints[7] = p.y_sub_bottom;
			dbl[0] = p.unrap[0]; dbl[1] = p.unrap[1]; dbl[2] = p.CoM[0]; dbl[3] = p.CoM[1]; dbl[4] = p.area;
			memcpy(phi, p.phi, sizeof(double) * p.S);
		}
	}
	cell_free(&p);
	return rc;
}

int orc_test_set_site(orc_sim *s, int i, int which, const double *in) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	double *dst = which == 0 ? p->phi : which == 1 ? p->mu : which == 2 ? p->dx : which == 3 ? p->dy
			: which == 4 ? p->lap : which == 5 ? p->phi_old : which == 6 ? p->dphi_old : NULL;
	if(!dst) return set_err(s, "bad site array %d", which);
	memcpy(dst, in, sizeof(double) * (size_t) p->S);
	return 0;
}

int orc_test_set_rows(orc_sim *s, int which, const double *in) {
	double *dst = which == 0 ? s->vx : which == 1 ? s->vy : NULL;
	if(!dst) return set_err(s, "bad row array %d", which);
	memcpy(dst, in, sizeof(double) * (size_t) s->size_rows);
	return 0;
}

int orc_test_set_polarity(orc_sim *s, int i, double thetaQ, double thetaQ_old, double tr0, double tr1) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	p->thetaQ = thetaQ;
	p->thetaQ_old = thetaQ_old;
	p->Q00 = cos(thetaQ);
	p->Q01 = sin(thetaQ);
	p->tr[0] = tr0;
	p->tr[1] = tr1;
	return 0;
}

int orc_test_set_director(orc_sim *s, int i, double n0, double n1, double n0_old, double n1_old) {
	if(i < 0 || i >= s->N) return set_err(s, "cell index %d out of range", i);
	ocell *p = &s->c[i];
	p->nem[0] = n0;
	p->nem[1] = n1;
	p->nem_old[0] = n0_old;
	p->nem_old[1] = n1_old;
	return 0;
}

int orc_test_first_step(orc_sim *s, int store) { return first_step(s, store); }

int orc_test_forces(orc_sim *s) {
	for(int k = 0; k < s->L * s->L; k++) {
		s->phi2[k] = 0;
		s->sum_phi[k] = 0;
		s->sq00[k] = 0;
		s->sq01[k] = 0;
	}
	for(int i = 0; i < s->N; i++) s->c[i].perimeter = 0;
	return compute_forces(s);
}
