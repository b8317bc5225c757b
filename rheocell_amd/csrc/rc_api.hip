// rc_api.hip -- C ABI (include/rheocell.h) over the gfx950 kernels.
//
// Step schedule (FD_CPUBackend::sim_step, src/Backends/FD_CPUBackend.cpp:134-160,
// is first_step(true), compute_forces, first_step(false), compute_forces):
//
//  dry model (friction_cell == 0): the force evaluation and the following
//    field update of a stage are fused into one per-cell kernel, so a stage
//    is  k_box_gather(phi) -> k_stage<FUSED>.  The force evaluation that the
//    reference runs at the end of a step therefore runs at the start of the
//    next stage; mu, v, grad phi are never written to HBM and are
//    re-materialised (k_stage<CF>, no side effects) only when a caller asks
//    for them.
//  wet model: a stage is k_stage<FS> (field update, recomputing mu and
//    grad phi from phi and the box sums), then k_box_gather -> k_stage<CF>
//    (right-hand side rows) -> Jacobi-PCG (x and y together), exactly the
//    reference order, so the velocity rows always belong to the current phi.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rheocell.h"
#include "rc_kernels.h"
#include "rc_net.h"

using namespace rc;

struct TimedPair {
	int kind;
	hipEvent_t a, b;
};

struct rc_handle {
	// configuration
	std::vector<rc_box_desc> boxes;
	std::vector<DevBox> dboxes;
	std::vector<DevParams> dparams;   // 2 per box: [0] before equilibration, [1] active
	int nboxes = 0;
	int ncells = 0;
	long long rows = 0;
	int ntiles = 0;
	long long box_sites = 0;
	int max_S = 0, max_W = 0;
	size_t lds_stage = 0, lds_cg = 0;
	bool wet = false, jq = false, win = false;   // win: some omega != 0 (adhesion reads the box window)
	bool nematic = false;         // interaction_type = wetmodel (WetModel.cpp)
	bool ext = false;             // some box has walls or external forces
	// split schedule: the force evaluation is its own launch sequence and keeps
	// the velocity rows (the wet model, and every nematic box)
	bool split = false;

	// device memory
	hipStream_t stream = nullptr;
	DevBox *d_boxes = nullptr;
	int *d_tile_box = nullptr;
	DevCell *d_cells = nullptr;
	int4 *d_geo = nullptr;        // compact per-cell geometry for the box gathers
	DevParams *d_params = nullptr;
	double *d_phiA = nullptr, *d_phiB = nullptr, *d_dphi = nullptr, *d_rep = nullptr;
	double2 *d_v = nullptr, *d_f = nullptr, *d_r = nullptr, *d_p = nullptr;
	double *d_trig = nullptr;
	double *d_phi2 = nullptr, *d_diag = nullptr;
	double *d_sq00 = nullptr, *d_sq01 = nullptr;   // nematic: box sums of phi Q00, phi Q01
	double *d_walls = nullptr;                     // ext: getWalls of every box site
	unsigned char *d_wmask = nullptr;              // ext: open E/W/N/S/self bits (StageArgs::wmask)
	int *d_cover = nullptr, *d_tile_cells = nullptr, *d_tile_count = nullptr;
	int4 *d_tile_geo = nullptr;
	double2 *d_sp = nullptr, *d_p2 = nullptr, *d_n4 = nullptr;
	double *d_part_box = nullptr;
	double *d_part = nullptr;
	CGState *d_cg = nullptr;
	DevStatus *d_st = nullptr;
	double *d_gauss = nullptr;
	long long gauss_cap = 0;
	double *d_out[4] = { nullptr, nullptr, nullptr, nullptr };
	CGState *h_cg = nullptr;      // pinned
	// site-space form of the CG (SCGState)
	int cg_form = RC_CG_SITES;
	double2 *d_sc = nullptr, *d_ss = nullptr;
	double2 *d_sy[2] = {}, *d_sq[2] = {};
	double *d_psum[2] = {};
	SCGState *d_scg = nullptr;    // [2 * nboxes]
	SCGState *h_scg = nullptr;    // pinned, [2 * nboxes]
	SCGBox *d_cgb = nullptr;      // [nboxes]
	SCGTileGeo *d_cg_geo = nullptr;   // [cg_ntiles]
	int *d_mn = nullptr;          // [sites] rows here and at the four neighbours (site form)
	double *d_dtab = nullptr;     // [nboxes][2][2][SCG_NTAB] D and 1 / D by neighbour rows
	std::vector<SCGBox> cgb;      // host copy (its iteration caps follow the layout)
	int cg_ntiles = 0;
	int scg_parity = 0;           // h_scg[2 b + scg_parity]: each box's state after the last solve
	hipEvent_t cg_e0 = nullptr, cg_e1 = nullptr;
	bool cg_ms_pending = false;   // the last solve's time is still to be read from cg_e0 / cg_e1
	DevStatus *h_st = nullptr;    // pinned

	// state machine
	bool uploaded = false, initialized = false;
	bool lists_valid = false;     // tile lists match the current patch geometry
	// Field-update kernel groups: cells grouped by the smallest stage shape
	// that fits their patch, one launch per group with its own shape and LDS.
	struct StageGroup {
		int key = 0;                             // index of the first fitting shape (the group's size class)
		std::vector<int> cells;
		int *d_cells = nullptr;
		int mx = 0, my = 0;                      // largest patch sides of the group
		std::vector<std::pair<int, int>> valid;   // shapes that fit every patch of the group
		int threads = 256, rows = 7;             // shape in use
		double tune_us[8] = {};                  // tuner: mean pure force-evaluation time per shape
	};
	std::vector<StageGroup> groups;

	// HIP graph of GRAPH_STEPS dry steps, replayed per window of steps
	bool use_graphs = true, graph_valid = false, graph_gauss = false, capturing = false;
	int graph_set = -1;
	hipGraphExec_t graph_exec = nullptr;
	double *d_ggauss = nullptr;     // the window's predictor noise [GRAPH_STEPS][ncells]
	double *d_bf = nullptr;         // rc_box_field: four L x L arrays of the largest box, on first use
	int device = 0;                 // device current at rc_create
	double *scratch = nullptr;      // this device's static staging array (rc_box_field_noalloc)
	int *d_step_base = nullptr;     // first step of the window being replayed
	int *h_step_base = nullptr;     // pinned, one slot per window of a chunk
	long long graph_launches = 0, eager_steps = 0;
	long long growths = 0;          // growth requests served (host re-cut + re-layout)
	double growth_ms = 0.;          // host time spent serving them
	int stage_forced[2] = { 0, 0 };     // rc_set_stage_shape; {0, 0}: autotune
	bool stage_tuned = false;
	// the dry fused stage as one persistent launch over every cell (k_stage_p)
	// instead of one launch per group: eligible layouts, the choice, and the
	// tuner's times of both (predictor stage, us)
	bool persist_ok = false, persist = false;
	int p_npt = 0, p_tile = 0, p_rows = 0, p_sites = 0;
	int ncu = 0;                  // compute units of the device (k_stage_p grid)
	double persist_us = 0., groups_us = 0.;
	int cur_set = 0;              // parameter set currently applied
	int cf_set = 0;               // set of the force evaluation preceding the next field update
	// the reference's solver restart bookkeeping (WetPolarModel.cpp:151-157), one
	// simulation per box: rows of each box now and at its last solve, warm
	// solves since its last cold one; rows of box b are box_rows[b] .. box_rows[b + 1] - 1
	std::vector<int> restart_solver;
	std::vector<long long> size_rows, size_rows_old, box_rows;
	std::vector<char> cold_box;   // boxes whose current solve starts from zero
	int warm_mode = RC_WARM_REFERENCE;
	int last_iters[2] = { 8, 8 };   // iterations of the last warm / cold solve (first batch size)
	rc_cg_stats_t cgs{};
	std::string err;

	// learned warm start
	rc_net *net = nullptr;
	float scale[3] = { 1.f, 1.f, 1.f };
	int stages = 1;                 // network refinement stages per solve
	bool pending_learned_stat = false;
	double2 *d_xprev = nullptr, *d_rprev = nullptr;   // previous-solution start and residual (fallback)
	float *d_snap = nullptr;
	float *d_rec = nullptr;         // training records: [cap][L*L][5]
	int rec_cap = 0, rec_n = 0;
	int rec_open = 0;               // records of the running solve (one per stage)
	double2 *d_xrec = nullptr;      // x at each recorded stage [RC_MAX_STAGES][rows]
	double *d_rnrec = nullptr;      // |r_c|^2 at each recorded stage [RC_MAX_STAGES][2]

	// timing
	bool timing = false;
	std::vector<TimedPair> pairs;
	std::vector<hipEvent_t> pool;
	size_t pool_used = 0;
	double t_ms[RC_NKINDS] = {};
	long long t_n[RC_NKINDS] = {};
};

// ---------------------------------------------------------------------------
static int fail(rc_handle *h, int code, const char *fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	if(h) h->err = buf;
	return code;
}

#define HIPCHK(h, call)                                                                  \
	do {                                                                                 \
		hipError_t e_ = (call);                                                          \
		if(e_ != hipSuccess) return fail(h, RC_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
	} while(0)

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
	*p = nullptr;
	if(n == 0) n = 1;
	return hipMalloc((void **) p, n * sizeof(T));
}

static void dfree(void *p) {
	if(p) (void) hipFree(p);
}

static hipEvent_t ev_get(rc_handle *h) {
	if(h->pool_used == h->pool.size()) {
		hipEvent_t e;
		(void) hipEventCreate(&e);
		h->pool.push_back(e);
	}
	return h->pool[h->pool_used++];
}

struct TimeScope {
	rc_handle *h;
	int kind;
	bool group;
	hipEvent_t b;
	TimeScope(rc_handle *h_, int k, bool several = false)
		: h(h_), kind(k), group(several || k == RC_T_CG_SOLVE || k == RC_T_STEP || k == RC_T_CG_LEARNED), b(nullptr) {
		if(h->timing) {
			hipEvent_t a = ev_get(h);
			b = ev_get(h);
			h->pairs.push_back({ kind, a, b });
			if(group) (void) hipEventRecord(a, h->stream);
			else set_launch_events(a, b);
		}
	}
	~TimeScope() {
		if(h->timing) {
			if(group) (void) hipEventRecord(b, h->stream);
			else set_launch_events(nullptr, nullptr);
		}
	}
};

static void collect_timing(rc_handle *h) {
	if(!h->timing) return;
	for(auto &p : h->pairs) {
		float ms = 0.f;
		if(hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
			h->t_ms[p.kind] += ms;
			h->t_n[p.kind] += 1;
		}
	}
	h->pairs.clear();
	h->pool_used = 0;
}

// 1/x when x is a power of two with a normal reciprocal (then y / x == y * (1/x)
// exactly), else 0
static double exact_rcp(double x) {
	int e = 0;
	const double m = std::frexp(x, &e);
	if(!(x > 0) || m != 0.5 || e < -1020 || e > 1020) return 0.;
	return 1. / x;
}

static DevParams make_params(const rc_box_desc &d, bool active) {
	const rc_model_params &m = d.params;
	DevParams p{};
	p.gamma = m.gamma;
	p.lambda = m.lambda;
	p.omega = m.omega;
	p.mu = m.mu;
	p.kappa = m.kappa;
	p.a0 = PI_F * m.R * m.R;   // float arithmetic, WetPolarModel.cpp:50
	p.p0 = 12.;                // WetPolarModel.h:42
	p.passive_alpha = m.passive_alpha;
	p.J0 = m.J0;
	p.dt = m.dt;
	p.sqrt_dt = std::sqrt(m.dt);
	p.anchoring = m.anchoring;
	p.lambda_rcp = exact_rcp(m.lambda);
	p.mu_a0 = p.mu / p.a0;              // the grouped quotients of WetPolarModel.cpp:278-291
	p.rcoef = 4 * (p.kappa / p.lambda);
	if(active) {
		p.zetaQ_self = m.zetaQ_self;
		p.zetaQ_inter = m.zetaQ_inter;
		p.J_Q = m.J_Q;
		p.friction = m.friction;
		p.friction_cell = m.friction_cell;
		p.beta = m.beta;
	}
	else {  // WetPolarModel constructor values (WetPolarModel.cpp:4-21)
		p.zetaQ_self = 0;
		p.zetaQ_inter = 0;
		p.J_Q = 0;
		p.friction = 1.;
		p.friction_cell = 0.;
		p.beta = 0.;
	}
	p.nematic = m.interaction == RC_INTERACTION_WETNEMATIC ? 1 : 0;
	if(p.nematic) p.beta = 0.;   // WetModel has no perimeter term
	// external forces: the coefficients of their free energy and density as
	// the force classes form them (ChannelWalls.cpp:26-31, ShearFlowChannel.cpp:33-39),
	// and the last force's velocity profile; the shear rate acts after
	// equilibration (set_ext_properties)
	p.nforce = d.nforces;
	for(int f = 0; f < d.nforces; ++f) {
		const rc_ext_force &F = d.forces[f];
		p.fF[f] = F.kind == RC_FORCE_CHANNEL_WALLS ? 2.0 * F.kappa_wall / F.lambda
				: (F.kind == RC_FORCE_SHEAR_FLOW || F.kind == RC_FORCE_POIS_FLOW) ? 4.0 * F.kappa_wall / F.lambda : 0.;
		p.fD[f] = F.kind == RC_FORCE_SHEAR_FLOW ? 2.0 * F.kappa_wall / F.lambda : 0.;
		p.vext_kind = F.kind == RC_FORCE_SHEAR_FLOW ? 1 : F.kind == RC_FORCE_POIS_FLOW ? 2 : F.kind == RC_FORCE_STRETCHING ? 3 : 0;
		p.vext_lambda = F.lambda;
		p.vext_rate = active ? F.shear_rate : 0.;
	}
	p.friction_rcp = exact_rcp(p.friction);
	p.pcoef = 4 * (p.beta / p.p0);
	return p;
}

// ---------------------------------------------------------------------------
// Static staging of rc_box_field_noalloc: the four L x L fields (f_x, f_y and
// their coarse averages) of one box up to RC_BOX_FIELD_STATIC_L sides.  Each
// GPU holds its own copy from the library's load to the end of the process;
// host code reaches it only through the address hipGetSymbolAddress returns on
// that device (the array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[4 * RC_BOX_FIELD_STATIC_L * RC_BOX_FIELD_STATIC_L];
constexpr int SCRATCH_DEVICES = 64;
static double *scratch_ptr[SCRATCH_DEVICES];
static std::mutex scratch_mu;

// once per device, with that device current (called from rc_create)
static int scratch_setup(rc_handle *h) {
	HIPCHK(h, hipGetDevice(&h->device));
	HIPCHK(h, hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->device));
	if(h->device < 0 || h->device >= SCRATCH_DEVICES) return RC_OK;   // rc_box_field_noalloc refuses such a device
	std::lock_guard<std::mutex> lock(scratch_mu);
	if(!scratch_ptr[h->device]) {
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p)
			return fail(h, RC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", h->device, hipGetErrorString(e));
		scratch_ptr[h->device] = (double *) p;
	}
	h->scratch = scratch_ptr[h->device];
	return RC_OK;
}

// ---------------------------------------------------------------------------
extern "C" const char *rc_last_error(rc_handle *h) { return h ? h->err.c_str() : "null handle"; }

extern "C" int rc_create(const rc_box_desc *boxes, int nboxes, rc_handle **out) {
	if(!out || !boxes || nboxes <= 0) return RC_ERR_ARG;
	*out = nullptr;
	rc_handle *h = new rc_handle();
	h->nboxes = nboxes;
	h->boxes.assign(boxes, boxes + nboxes);
	int cells = 0, tiles = 0;
	long long sites = 0;
	for(int b = 0; b < nboxes; ++b) {
		const rc_box_desc &d = boxes[b];
		if(d.L <= 2 || d.ncells < 0) {
			*out = h;
			return fail(h, RC_ERR_ARG, "box %d: invalid side %d or cell count %d", b, d.L, d.ncells);
		}
		if(!(d.params.wall_slip > 0)) {
			*out = h;
			return fail(h, RC_ERR_ARG, "box %d: wall_slip must be > 0 in a periodic square box", b);
		}
		if(!(d.params.friction > 0)) {
			*out = h;
			return fail(h, RC_ERR_ARG, "box %d: friction must be > 0", b);
		}
		DevBox B{};
		B.L = d.L;
		B.box_off = (int) sites;
		B.tile_off = tiles;
		B.ntx = (d.L + TW - 1) / TW;
		B.nty = (d.L + TH - 1) / TH;
		B.cell_begin = cells;
		B.cell_end = cells + d.ncells;
		h->dboxes.push_back(B);
		if(d.nforces < 0 || d.nforces > RC_MAX_FORCES) {
			*out = h;
			return fail(h, RC_ERR_ARG, "box %d: %d external forces (0..%d)", b, d.nforces, RC_MAX_FORCES);
		}
		for(int f = 0; f < d.nforces; ++f)
			if(d.forces[f].kind < RC_FORCE_CHANNEL_WALLS || d.forces[f].kind > RC_FORCE_STRETCHING ||
					(d.forces[f].kind != RC_FORCE_STRETCHING && !(d.forces[f].lambda != 0))) {
				*out = h;
				return fail(h, RC_ERR_ARG, "box %d: external force %d has kind %d / lambda %g", b, f, d.forces[f].kind,
						d.forces[f].lambda);
			}
		if(d.walls || d.nforces > 0) {
			h->ext = true;
			if(d.params.J_Q != 0 && d.params.interaction == RC_INTERACTION_WETPOLAR) {
				*out = h;
				return fail(h, RC_ERR_ARG, "box %d: J_Q with walls or external forces is not supported", b);
			}
			if(d.walls && d.params.friction_cell != 0) {
				*out = h;
				return fail(h, RC_ERR_ARG, "box %d: walls with the wet solve (friction_cell != 0) are not supported", b);
			}
		}
		h->dparams.push_back(make_params(d, false));
		h->dparams.push_back(make_params(d, true));
		cells += d.ncells;
		tiles += B.ntx * B.nty;
		sites += (long long) d.L * d.L;
		if(d.params.interaction != RC_INTERACTION_WETPOLAR && d.params.interaction != RC_INTERACTION_WETNEMATIC) {
			*out = h;
			return fail(h, RC_ERR_ARG, "box %d: unknown interaction %d", b, d.params.interaction);
		}
		if(d.params.interaction != boxes[0].params.interaction) {
			*out = h;
			return fail(h, RC_ERR_ARG, "all boxes of a handle must share the interaction");
		}
		if(d.params.interaction == RC_INTERACTION_WETNEMATIC) h->nematic = true;
		if(d.params.friction_cell != 0) h->wet = true;
		if(d.params.J_Q != 0 && d.params.interaction == RC_INTERACTION_WETPOLAR) h->jq = true;
		if(d.params.omega != 0) h->win = true;
	}
	for(int b = 1; b < nboxes; ++b)
		if((boxes[b].params.friction_cell != 0) != h->wet) {
			*out = h;
			return fail(h, RC_ERR_ARG, "all boxes of a handle must share the dry/wet model");
		}
	h->split = h->wet || h->nematic;
	h->ncells = cells;
	h->ntiles = tiles;
	h->box_sites = sites;
	*out = h;
	int rc = scratch_setup(h);
	if(rc) return rc;
	HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
	HIPCHK(h, dalloc(&h->d_boxes, nboxes));
	HIPCHK(h, dalloc(&h->d_params, 2 * nboxes));
	HIPCHK(h, dalloc(&h->d_cells, cells));
	HIPCHK(h, dalloc(&h->d_geo, cells));
	HIPCHK(h, dalloc(&h->d_tile_cells, (size_t) tiles * MAXT));
	HIPCHK(h, dalloc(&h->d_tile_geo, (size_t) tiles * MAXT));
	HIPCHK(h, dalloc(&h->d_tile_count, tiles));
	HIPCHK(h, dalloc(&h->d_phi2, sites));
	HIPCHK(h, dalloc(&h->d_cover, sites));
	HIPCHK(h, dalloc(&h->d_st, 1));
	HIPCHK(h, hipHostMalloc((void **) &h->h_st, sizeof(DevStatus)));
	HIPCHK(h, hipHostMalloc((void **) &h->h_cg, sizeof(CGState)));
	HIPCHK(h, hipEventCreate(&h->cg_e0));
	HIPCHK(h, hipEventCreate(&h->cg_e1));
	std::vector<int> tile_box(tiles);
	for(int b = 0; b < nboxes; ++b)
		for(int t = 0; t < h->dboxes[b].ntx * h->dboxes[b].nty; ++t) tile_box[h->dboxes[b].tile_off + t] = b;
	HIPCHK(h, dalloc(&h->d_tile_box, tiles));
	HIPCHK(h, hipMemcpy(h->d_tile_box, tile_box.data(), sizeof(int) * tiles, hipMemcpyHostToDevice));
	HIPCHK(h, hipMemcpy(h->d_boxes, h->dboxes.data(), sizeof(DevBox) * nboxes, hipMemcpyHostToDevice));
	HIPCHK(h, hipMemcpy(h->d_params, h->dparams.data(), sizeof(DevParams) * 2 * nboxes, hipMemcpyHostToDevice));
	HIPCHK(h, hipMemset(h->d_st, 0, sizeof(DevStatus)));
	HIPCHK(h, hipMemset(h->d_phi2, 0, sizeof(double) * sites));
	if(h->nematic) {
		HIPCHK(h, dalloc(&h->d_sq00, sites));
		HIPCHK(h, dalloc(&h->d_sq01, sites));
	}
	if(h->ext) {
		// the wall field and the open bits of every site: E 1, W 2, N 4, S 8,
		// the site itself 16 (getWalls < wall_slip, BaseInteraction.h:196-263)
		std::vector<double> w(sites, 0.);
		std::vector<unsigned char> m(sites, 0);
		for(int b = 0; b < nboxes; ++b) {
			const int L = boxes[b].L, off = h->dboxes[b].box_off;
			const double ws = boxes[b].params.wall_slip;
			if(boxes[b].walls) std::copy(boxes[b].walls, boxes[b].walls + (size_t) L * L, w.begin() + off);
			for(int y = 0; y < L; ++y)
				for(int x = 0; x < L; ++x) {
					auto open = [&](int xx, int yy) { return w[off + ((yy + L) % L) * L + (xx + L) % L] < ws; };
					m[off + y * L + x] = (open(x + 1, y) ? 1 : 0) | (open(x - 1, y) ? 2 : 0) | (open(x, y + 1) ? 4 : 0) |
							(open(x, y - 1) ? 8 : 0) | (open(x, y) ? 16 : 0);
				}
		}
		HIPCHK(h, dalloc(&h->d_walls, sites));
		HIPCHK(h, dalloc(&h->d_wmask, sites));
		HIPCHK(h, hipMemcpy(h->d_walls, w.data(), sizeof(double) * sites, hipMemcpyHostToDevice));
		HIPCHK(h, hipMemcpy(h->d_wmask, m.data(), sites, hipMemcpyHostToDevice));
	}
	if(h->wet) {
		HIPCHK(h, dalloc(&h->d_diag, sites));
		HIPCHK(h, dalloc(&h->d_sp, sites));
		HIPCHK(h, dalloc(&h->d_p2, sites));
		HIPCHK(h, dalloc(&h->d_n4, sites));
		HIPCHK(h, dalloc(&h->d_part_box, 2 * ((sites + 255) / 256)));
		// partials: 6 per cell (row passes), 2 per tile (the learned stage's gather<E>)
		HIPCHK(h, dalloc(&h->d_part, (size_t) std::max(6 * std::max(cells, 1), 2 * tiles)));
		HIPCHK(h, dalloc(&h->d_cg, 1));
		HIPCHK(h, hipMemset(h->d_cg, 0, sizeof(CGState)));
		// the site-form tiles of every box (box-major) and each box's solver parameters
		std::vector<SCGBox> &cgb = h->cgb;
		cgb.assign(nboxes, SCGBox{});
		int cg_tiles = 0;
		std::vector<SCGTileGeo> cg_geo;
		for(int b = 0; b < nboxes; ++b) {
			const DevBox &B = h->dboxes[b];
			SCGBox &g = cgb[b];
			g.L = B.L;
			g.box_off = B.box_off;
			g.ntx = (B.L + SCG_TX - 1) / SCG_TX;
			g.tile_off = cg_tiles;
			g.ntiles = g.ntx * ((B.L + SCG_TY - 1) / SCG_TY);
			g.cell_begin = B.cell_begin;
			g.cell_end = B.cell_end;
			const rc_model_params &mp = boxes[b].params;
			g.tol = mp.CGtolerance > 0 ? mp.CGtolerance : 2.220446049250313e-16;
			g.max_iter = mp.cg_max_iter > 0 ? mp.cg_max_iter : 1 << 30;
			cg_tiles += g.ntiles;
			for(int lt = 0; lt < g.ntiles; ++lt) {
				SCGTileGeo G{};
				G.box = b;
				G.o = g.box_off;
				G.L = g.L;
				G.tile_off = g.tile_off;
				G.ntiles = g.ntiles;
				G.x0 = (lt % g.ntx) * SCG_TX;
				G.y0 = (lt / g.ntx) * SCG_TY;
				G.lead = lt == 0;
				G.eta[0] = h->dparams[2 * b].friction_cell;
				G.eta[1] = h->dparams[2 * b + 1].friction_cell;
				cg_geo.push_back(G);
			}
		}
		// the one-box iteration takes box 0 from its launch arguments at site 0
		if(cgb[0].box_off != 0) return fail(h, RC_ERR_STATE, "the first box does not start at site 0");
		// D by the four neighbours' row count n (friction, then friction_cell
		// added once per row, as k_scg_init adds them) and its reciprocal
		std::vector<double> dtab((size_t) nboxes * 2 * 2 * SCG_NTAB);
		for(int b = 0; b < nboxes; ++b)
			for(int set = 0; set < 2; ++set) {
				const DevParams &P = h->dparams[2 * b + set];
				double *t = dtab.data() + (size_t) (2 * b + set) * 2 * SCG_NTAB;
				double D = P.friction;
				for(int n = 0; n < SCG_NTAB; ++n) {
					if(n > 0) D += P.friction_cell * 1.0;
					t[n] = D;
					t[SCG_NTAB + n] = 1. / D;
				}
			}
		HIPCHK(h, dalloc(&h->d_dtab, dtab.size()));
		HIPCHK(h, hipMemcpy(h->d_dtab, dtab.data(), sizeof(double) * dtab.size(), hipMemcpyHostToDevice));
		HIPCHK(h, dalloc(&h->d_mn, sites));
		h->cg_ntiles = cg_tiles;
		for(double2 **f : { &h->d_sc, &h->d_ss, &h->d_sy[0], &h->d_sy[1], &h->d_sq[0], &h->d_sq[1] })
			HIPCHK(h, dalloc(f, sites));
		for(int i = 0; i < 2; ++i) HIPCHK(h, dalloc(&h->d_psum[i], 6 * (size_t) h->cg_ntiles));
		HIPCHK(h, dalloc(&h->d_scg, 2 * (size_t) nboxes));
		HIPCHK(h, hipMemset(h->d_scg, 0, 2 * nboxes * sizeof(SCGState)));
		HIPCHK(h, hipHostMalloc((void **) &h->h_scg, 2 * nboxes * sizeof(SCGState)));
		HIPCHK(h, dalloc(&h->d_cgb, nboxes));
		HIPCHK(h, hipMemcpy(h->d_cgb, cgb.data(), sizeof(SCGBox) * nboxes, hipMemcpyHostToDevice));
		HIPCHK(h, dalloc(&h->d_cg_geo, h->cg_ntiles));
		HIPCHK(h, hipMemcpy(h->d_cg_geo, cg_geo.data(), sizeof(SCGTileGeo) * h->cg_ntiles, hipMemcpyHostToDevice));
	}
	return RC_OK;
}

extern "C" void rc_destroy(rc_handle *h) {
	if(!h) return;
	if(h->stream) (void) hipStreamSynchronize(h->stream);
	dfree(h->d_boxes); dfree(h->d_tile_box); dfree(h->d_cells); dfree(h->d_geo); dfree(h->d_params);
	dfree(h->d_phiA); dfree(h->d_phiB); dfree(h->d_dphi); dfree(h->d_rep);
	dfree(h->d_v); dfree(h->d_f); dfree(h->d_r); dfree(h->d_p); dfree(h->d_trig);
	dfree(h->d_phi2); dfree(h->d_diag); dfree(h->d_sq00); dfree(h->d_sq01); dfree(h->d_walls); dfree(h->d_wmask); dfree(h->d_cover); dfree(h->d_tile_cells); dfree(h->d_tile_geo); dfree(h->d_tile_count);
	dfree(h->d_sp); dfree(h->d_p2); dfree(h->d_n4); dfree(h->d_part_box); dfree(h->d_part); dfree(h->d_cg); dfree(h->d_st); dfree(h->d_gauss);
	for(auto p : h->d_out) dfree(p);
	dfree(h->d_snap); dfree(h->d_rec); dfree(h->d_xprev); dfree(h->d_rprev);
	for(auto &g : h->groups) dfree(g.d_cells);
	if(h->graph_exec) (void) hipGraphExecDestroy(h->graph_exec);
	dfree(h->d_ggauss); dfree(h->d_step_base); dfree(h->d_bf);
	if(h->h_step_base) (void) hipHostFree(h->h_step_base); dfree(h->d_xrec); dfree(h->d_rnrec);
	if(h->h_cg) (void) hipHostFree(h->h_cg);
	dfree(h->d_sc); dfree(h->d_ss); dfree(h->d_scg); dfree(h->d_cgb); dfree(h->d_cg_geo);
	dfree(h->d_mn); dfree(h->d_dtab);
	for(int i = 0; i < 2; ++i) {
		dfree(h->d_sy[i]); dfree(h->d_sq[i]); dfree(h->d_psum[i]);
	}
	if(h->h_scg) (void) hipHostFree(h->h_scg);
	if(h->h_st) (void) hipHostFree(h->h_st);
	for(auto e : h->pool) (void) hipEventDestroy(e);
	if(h->cg_e0) (void) hipEventDestroy(h->cg_e0);
	if(h->cg_e1) (void) hipEventDestroy(h->cg_e1);
	if(h->stream) (void) hipStreamDestroy(h->stream);
	delete h;
}

// Patch trig table of MultiPhaseField::resizing (MultiPhaseField.cpp:54-65):
// cos/sin(2*PI*i/L) with the argument evaluated in float; the reference build
// evaluates each cos/sin pair of one argument with a single glibc sincos call.
static void trig_table(int L, std::vector<double> &out) {
	std::vector<double> c(L), s(L);
	for(int i = 0; i < L; i++) {
		float a = 2 * PI_F * i / L;
		sincos((double) a, &s[i], &c[i]);
	}
	out.insert(out.end(), c.begin(), c.end());
	out.insert(out.end(), s.begin(), s.end());
}

// Lays out the rows of all cells (field_start_index order), the trig pool and
// the cell table on the device.  rep: Repulsion rows to keep (J_Q != 0), or null.
// phi / rep: rows in the new layout, host arrays, or device arrays when on_device
static int upload_layout(rc_handle *h, const rc_cell *cells, const double *phi, const double *rep,
		bool on_device = false) {
	const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
	std::vector<DevCell> dc(h->ncells);
	std::vector<double> trig;
	std::map<int, int> trig_off;
	long long rows = 0;
	int maxS = 0, maxW = 0;
	std::vector<std::pair<int, int>> dims;
	std::vector<long long> box_rows(h->nboxes + 1, 0);
	for(int b = 0; b < h->nboxes; ++b) {
		const DevBox &B = h->dboxes[b];
		box_rows[b] = rows;
		for(int i = B.cell_begin; i < B.cell_end; ++i) {
			const rc_cell &c = cells[i];
			DevCell &d = dc[i];
			if(c.LsubX < 3 || c.LsubY < 3 || c.LsubX > B.L || c.LsubY > B.L)
				return fail(h, RC_ERR_ARG, "cell %d: patch %dx%d does not fit box %d", i, c.LsubX, c.LsubY, B.L);
			if(c.offset[0] < 0 || c.offset[0] >= c.LsubX || c.offset[1] < 0 || c.offset[1] >= c.LsubY)
				return fail(h, RC_ERR_ARG, "cell %d: offset out of range", i);
			if(c.corner < 0 || c.corner >= B.L * B.L) return fail(h, RC_ERR_ARG, "cell %d: corner out of range", i);
			d.LX = c.LsubX;
			d.LY = c.LsubY;
			d.S = c.LsubX * c.LsubY;
			d.start = (int) rows;
			d.offx = c.offset[0];
			d.offy = c.offset[1];
			d.corner = c.corner;
			d.corner_old = c.corner_old;
			d.x_sub_left = c.x_sub_left;
			d.y_sub_bottom = c.y_sub_bottom;
			d.box = b;
			for(int side : { c.LsubX, c.LsubY }) {
				if(!trig_off.count(side)) {
					trig_off[side] = (int) trig.size();
					trig_table(side, trig);
				}
			}
			d.trig_x = trig_off[c.LsubX];
			d.trig_y = trig_off[c.LsubY];
			d.CoMx = c.CoM[0];
			d.CoMy = c.CoM[1];
			d.CoM_oldx = c.CoM_old[0];
			d.CoM_oldy = c.CoM_old[1];
			d.unrapx = c.unrap[0];
			d.unrapy = c.unrap[1];
			d.Q00 = c.Q00;
			d.Q01 = c.Q01;
			d.thetaQ = c.thetaQ;
			d.thetaQ_old = c.thetaQ_old;
			d.area = c.area;
			d.area_old = c.area_old;
			d.sumF = c.sumF;
			d.sumF_old = c.sumF_old;
			d.perimeter = c.perimeter;
			d.S00 = c.S00;
			d.S01 = c.S01;
			d.trx = c.total_force_repulsion[0];
			d.try_ = c.total_force_repulsion[1];
			d.nem0 = c.nemQ[0];
			d.nem1 = c.nemQ[1];
			d.nem_old0 = c.nemQ_old[0];
			d.nem_old1 = c.nemQ_old[1];
			rows += d.S;
			maxS = std::max(maxS, d.S);
			maxW = std::max(maxW, (d.LX + 2) * (d.LY + 2));
			dims.emplace_back(d.LX, d.LY);
		}
	}
	if(rows >= (1ll << 31)) return fail(h, RC_ERR_ARG, "too many patch sites (%lld)", rows);
	box_rows[h->nboxes] = rows;
	h->box_rows = box_rows;
	if(h->wet) {
		// Eigen's default cap, 2 x the matrix size, per box (its own solve)
		for(int b = 0; b < h->nboxes; ++b)
			if(h->boxes[b].params.cg_max_iter <= 0)
				h->cgb[b].max_iter = (int) std::min<long long>(2 * (box_rows[b + 1] - box_rows[b]), 1ll << 30);
		HIPCHK(h, hipMemcpy(h->d_cgb, h->cgb.data(), sizeof(SCGBox) * h->nboxes, hipMemcpyHostToDevice));
	}
	if(rows != h->rows) {
		dfree(h->d_phiA); dfree(h->d_phiB); dfree(h->d_dphi); dfree(h->d_rep); dfree(h->d_v);
		dfree(h->d_f); dfree(h->d_r); dfree(h->d_p);
		for(auto &p : h->d_out) { dfree(p); p = nullptr; }
		h->d_phiA = h->d_phiB = h->d_dphi = h->d_rep = nullptr;
		h->d_v = h->d_f = h->d_r = h->d_p = nullptr;
		HIPCHK(h, dalloc(&h->d_phiA, rows));
		HIPCHK(h, dalloc(&h->d_phiB, rows));
		HIPCHK(h, dalloc(&h->d_dphi, rows));
		HIPCHK(h, dalloc(&h->d_v, rows));
		if(h->jq) HIPCHK(h, dalloc(&h->d_rep, rows));
		if(h->wet) {
			HIPCHK(h, dalloc(&h->d_f, rows));
			HIPCHK(h, dalloc(&h->d_p, rows));
		}
		if(h->split) HIPCHK(h, dalloc(&h->d_r, rows));   // CG residual; scratch of materialize
		h->rows = rows;
		if(h->d_xprev) {
			dfree(h->d_xprev);
			dfree(h->d_rprev);
			h->d_xprev = h->d_rprev = nullptr;
			HIPCHK(h, dalloc(&h->d_xprev, rows));
			HIPCHK(h, dalloc(&h->d_rprev, rows));
		}
		if(h->d_xrec) {
			dfree(h->d_xrec);
			h->d_xrec = nullptr;
			HIPCHK(h, dalloc(&h->d_xrec, (size_t) RC_MAX_STAGES * rows));
		}
	}
	if(h->jq) {
		if(rep) HIPCHK(h, hipMemcpyAsync(h->d_rep, rep, sizeof(double) * rows, kind, h->stream));
		else HIPCHK(h, hipMemsetAsync(h->d_rep, 0, sizeof(double) * rows, h->stream));
	}
	HIPCHK(h, hipMemsetAsync(h->d_v, 0, sizeof(double2) * rows, h->stream));
	dfree(h->d_trig);
	HIPCHK(h, dalloc(&h->d_trig, trig.size()));
	HIPCHK(h, hipMemcpyAsync(h->d_trig, trig.data(), sizeof(double) * trig.size(), hipMemcpyHostToDevice, h->stream));
	HIPCHK(h, hipMemcpyAsync(h->d_cells, dc.data(), sizeof(DevCell) * h->ncells, hipMemcpyHostToDevice, h->stream));
	std::vector<int4> geo(h->ncells);
	for(int i = 0; i < h->ncells; ++i)
		geo[i] = make_int4(dc[i].corner, dc[i].LX | (dc[i].LY << 16), dc[i].offx | (dc[i].offy << 16), dc[i].start);
	HIPCHK(h, hipMemcpyAsync(h->d_geo, geo.data(), sizeof(int4) * h->ncells, hipMemcpyHostToDevice, h->stream));
	h->lists_valid = false;
	HIPCHK(h, hipMemcpyAsync(h->d_phiA, phi, sizeof(double) * rows, kind, h->stream));
	HIPCHK(h, hipMemsetAsync(h->d_st, 0, sizeof(DevStatus), h->stream));
	h->max_S = maxS;
	h->max_W = maxW;
	if(std::max(maxS, maxW) >= (1 << 16))
		return fail(h, RC_ERR_ARG, "a patch of %d sites is too large", maxS);
	if(h->nematic && nem_lds_bytes(maxS, h->d_wmask != nullptr) > NEM_LDS_MAX)
		return fail(h, RC_ERR_ARG, "a patch of %d sites is too large for the nematic force kernel (at most %d)", maxS,
				(int) (NEM_LDS_MAX / nem_lds_bytes(1, h->d_wmask != nullptr)));
	{
		// group the cells by the first of the group shapes 256x7, 512x8 and
		// 512x16 (table order) that fits their patch; a group may run in any
		// shape that fits all its cells.  Fewer, larger groups: a launch of a
		// few hundred cells fills a fraction of the device (an uneven
		// 32-replica ensemble in two groups instead of four: 0.66 -> 0.61 ms
		// per dry step)
		int shapes[16][2];
		const int n = stage_shapes(shapes);
		auto group_shape = [&](int j) { return (shapes[j][0] == 256 && shapes[j][1] == 7) || (shapes[j][0] == 512 && shapes[j][1] >= 8); };
		// shapes chosen for the previous layout (kept where they still fit: no re-tuning)
		std::map<int, std::pair<int, int>> prev;
		const bool prev_tuned = h->stage_tuned && !h->groups.empty();
		for(auto &g : h->groups) prev[g.key] = { g.threads, g.rows };
		for(auto &g : h->groups) dfree(g.d_cells);
		h->groups.assign(n, rc_handle::StageGroup());
		for(int i = 0; i < h->ncells; ++i) {
			const DevCell &d = dc[i];
			int k = 0;
			while(k < n && (!group_shape(k) || stage_strip_rows(d.LX, d.LY, shapes[k][0]) > shapes[k][1])) ++k;
			if(k == n) return fail(h, RC_ERR_ARG, "cell %d: a %dx%d patch does not fit the stage kernel", i, d.LX, d.LY);
			h->groups[k].cells.push_back(i);
		}
		std::vector<rc_handle::StageGroup> used;
		bool all_kept = prev_tuned;
		for(int k = 0; k < n; ++k) {
			rc_handle::StageGroup &g = h->groups[k];
			if(g.cells.empty()) continue;
			for(int i : g.cells) {
				g.mx = std::max(g.mx, dc[i].LX);
				g.my = std::max(g.my, dc[i].LY);
			}
			for(int j = 0; j < n; ++j) {
				bool ok = stage_lds_bytes(g.mx, g.my, shapes[j][1], h->jq, h->win) <= STAGE_LDS_MAX;
				for(int i : g.cells) ok = ok && stage_strip_rows(dc[i].LX, dc[i].LY, shapes[j][0]) <= shapes[j][1];
				if(ok) g.valid.emplace_back(shapes[j][0], shapes[j][1]);
			}
			if(g.valid.empty())
				return fail(h, RC_ERR_ARG, "a %dx%d patch needs more LDS than the stage kernel has", g.mx, g.my);
			g.key = k;
			g.threads = g.valid[0].first;
			g.rows = g.valid[0].second;
			bool kept = false;
			for(const auto &v : g.valid) {
				if(v.first == h->stage_forced[0] && v.second == h->stage_forced[1]) {
					g.threads = v.first;
					g.rows = v.second;
				}
				if(prev_tuned && prev.count(k) && prev[k] == v && h->stage_forced[0] == 0) {
					g.threads = v.first;
					g.rows = v.second;
					kept = true;
				}
			}
			if(!kept) all_kept = false;
			HIPCHK(h, dalloc(&g.d_cells, g.cells.size()));
			HIPCHK(h, hipMemcpy(g.d_cells, g.cells.data(), sizeof(int) * g.cells.size(), hipMemcpyHostToDevice));
			used.push_back(std::move(g));
			h->groups[k].d_cells = nullptr;
		}
		h->groups = std::move(used);
		// the persistent form (k_stage_p): dry fused stages in periodic boxes
		// without J_Q, omega and beta, patches of at most 3 strip rows at 1024
		// threads, and LDS for the cells of one workgroup
		const bool was_persist = h->persist;
		h->persist_ok = !h->split && !h->jq && !h->win && !h->ext && h->ncu > 0 && h->nboxes <= 4096;
		for(const DevParams &p : h->dparams) h->persist_ok = h->persist_ok && p.beta == 0 && p.omega == 0;
		h->p_npt = 2;
		h->p_tile = h->p_rows = h->p_sites = 0;
		for(int i = 0; i < h->ncells && h->persist_ok; ++i) h->p_npt = std::max(h->p_npt, p_rows_of(dc[i].LX, dc[i].LY));
		if(h->p_npt > P_STRIP_ROWS) h->persist_ok = false;
		for(int i = 0; i < h->ncells && h->persist_ok; ++i) {
			h->p_tile = std::max(h->p_tile, p_tile_of(dc[i].LX, dc[i].LY, h->p_npt));
			h->p_rows = std::max(h->p_rows, dc[i].LY + h->p_npt);
			h->p_sites = std::max(h->p_sites, dc[i].LX * dc[i].LY);
		}
		if(h->persist_ok) {
			StageArgs pa{};
			pa.p_tile = h->p_tile;
			pa.p_rows = h->p_rows;
			pa.p_sites = h->p_sites;
			pa.p_nbox = h->nboxes;
			pa.store = 1;
			const int grid = std::min(h->ncells, h->ncu);
			if(stage_p_lds_bytes(pa, (h->ncells + grid - 1) / grid) > STAGE_LDS_MAX - 512) h->persist_ok = false;   // + k_stage_p's static LDS
		}
		bool tunable = h->persist_ok && h->ncells >= P_MIN_CELLS_PER_CU * h->ncu;
		for(const auto &g : h->groups) tunable = tunable || g.valid.size() > 1;
		if(h->stage_forced[0] == P_THREADS) h->persist = h->persist_ok;   // forced; the groups after a growth it no longer fits
		else if(h->stage_forced[0] != 0) h->persist = false;
		else h->persist = h->persist_ok && all_kept && was_persist;
		h->stage_tuned = !tunable || h->stage_forced[0] != 0 || all_kept;
		h->graph_valid = false;
	}
	// largest stage LDS over the groups' valid shapes
	h->lds_stage = 0;
	for(const auto &g : h->groups)
		for(const auto &v : g.valid) h->lds_stage = std::max(h->lds_stage, stage_lds_bytes(g.mx, g.my, v.second, h->jq, h->win));
	h->lds_cg = sizeof(double2) * (size_t) maxW;
	HIPCHK(h, stage_set_lds_limit(std::max<size_t>(h->lds_stage, 64 * 1024)));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	return RC_OK;
}

extern "C" int rc_upload_cells(rc_handle *h, const rc_cell *cells, const double *phi) {
	if(!h) return RC_ERR_ARG;
	if(!cells || !phi) return fail(h, RC_ERR_ARG, "null cells/phi");
	int rc = upload_layout(h, cells, phi, nullptr);
	if(rc) return rc;
	h->uploaded = true;
	h->initialized = false;
	return RC_OK;
}

// ---------------------------------------------------------------------------
static StageArgs stage_args(rc_handle *h) {
	StageArgs a{};
	a.boxes = h->d_boxes;
	a.cells = h->d_cells;
	a.params = h->d_params;
	a.trig = h->d_trig;
	a.phi2 = h->d_phi2;
	a.vrows = h->d_v;
	a.frows = h->d_f;
	a.rep = h->d_rep;
	a.jq = h->jq ? 1 : 0;
	a.win = h->win ? 1 : 0;
	a.dphi_old = h->d_dphi;
	a.geo = h->d_geo;
	a.threads = 256;
	a.npt = 7;
	a.step_base = h->capturing ? h->d_step_base : nullptr;
	a.rcp = 1;
	for(const DevParams &p : h->dparams)
		if(p.lambda_rcp == 0 || p.friction_rcp == 0) a.rcp = 0;
	a.wmask = h->d_wmask;
	a.walls = h->d_walls;
	a.st = h->d_st;
	return a;
}

// box sums of phi^2 (k_box_gather); the tile lists are rebuilt only after the
// patches moved (corrector recentering, upload, growth)
static int run_gather(rc_handle *h, const double *phi) {
	GatherArgs g{};
	g.boxes = h->d_boxes;
	g.tile_box = h->d_tile_box;
	g.geo = h->d_geo;
	g.phi = phi;
	g.phi2 = h->d_phi2;
	g.cover = h->d_cover;
	g.tile_cells = h->d_tile_cells;
	g.tile_geo = h->d_tile_geo;
	g.tile_count = h->d_tile_count;
	g.build_lists = h->lists_valid ? 0 : 1;
	g.nboxes = h->nboxes;
	g.one_L = h->dboxes[0].L;
	g.one_ntx = h->dboxes[0].ntx;
	g.one_cells = h->dboxes[0].cell_end;
	g.nem = h->nematic ? 1 : 0;
	g.cells = h->d_cells;
	g.sq00 = h->d_sq00;
	g.sq01 = h->d_sq01;
	g.st = h->d_st;
	TimeScope ts(h, RC_T_BOX_GATHER);
	HIPCHK(h, launch_gather(g, h->ntiles, h->stream));
	h->lists_valid = true;
	return RC_OK;
}

// one launch per stage group, each with its shape, cell list and LDS size
static int launch_groups(rc_handle *h, int mode, const StageArgs &a0) {
	for(const auto &g : h->groups) {
		StageArgs a = a0;
		a.threads = g.threads;
		a.npt = g.rows;
		a.cell_list = h->groups.size() > 1 ? g.d_cells : nullptr;
		HIPCHK(h, launch_stage(mode, a, (int) g.cells.size(), stage_lds_bytes(g.mx, g.my, a.npt, h->jq, h->win), h->stream));
	}
	return RC_OK;
}

// StageArgs of k_stage_p: the layout's tile, row table and patch bounds
static StageArgs persist_args(rc_handle *h, const StageArgs &a0) {
	StageArgs a = a0;
	a.npt = h->p_npt;
	a.p_tile = h->p_tile;
	a.p_rows = h->p_rows;
	a.p_sites = h->p_sites;
	a.p_nbox = h->nboxes;
	a.cell_list = nullptr;
	return a;
}

static int run_stage(rc_handle *h, int mode, const StageArgs &a) {
	const int kind = mode == MODE_CF ? RC_T_STAGE_CF
			: mode == MODE_FUSED ? (a.store ? RC_T_STAGE_FUSED_PRED : RC_T_STAGE_FUSED_CORR)
			: (a.store ? RC_T_STAGE_FS_PRED : RC_T_STAGE_FS_CORR);
	if(mode == MODE_FUSED && h->persist) {
		TimeScope ts(h, kind);
		HIPCHK(h, launch_stage_p(persist_args(h, a), h->ncells, h->ncu, h->stream));
	}
	else {
		TimeScope ts(h, kind, h->groups.size() > 1);   // several launches are timed as one
		int rc = launch_groups(h, mode, a);
		if(rc) return rc;
	}
	if(mode != MODE_CF && !a.store) h->lists_valid = false;   // the corrector recentres patches
	return RC_OK;
}

static CGArgs cg_args(rc_handle *h) {
	CGArgs c{};
	c.boxes = h->d_boxes;
	c.tile_box = h->d_tile_box;
	c.cells = h->d_cells;
	c.geo = h->d_geo;
	c.params = h->d_params;
	c.set = h->cur_set;
	c.tile_cells = h->d_tile_cells;
	c.tile_geo = h->d_tile_geo;
	c.tile_count = h->d_tile_count;
	c.cover = h->d_cover;
	c.diag = h->d_diag;
	c.sp = h->d_sp;
	c.p2 = h->d_p2;
	c.n4 = h->d_n4;
	c.part_box = h->d_part_box;
	c.x = h->d_v;
	c.r = h->d_r;
	c.p = h->d_p;
	c.b = h->d_f;
	c.part = h->d_part;
	c.cg = h->d_cg;
	c.tol = h->boxes[0].params.CGtolerance > 0 ? h->boxes[0].params.CGtolerance : 2.220446049250313e-16;
	const int mi = h->boxes[0].params.cg_max_iter;
	c.max_iter = mi > 0 ? mi : (int) std::min<long long>(2 * h->rows, 1ll << 30);
	c.ncells = h->ncells;
	c.nboxes = h->nboxes;
	c.nbox_blk = (int) ((h->box_sites + 255) / 256);
	c.st = h->d_st;
	return c;
}

static SCGArgs scg_args(rc_handle *h, const CGArgs &c) {
	SCGArgs a{};
	a.boxes = h->d_boxes;
	a.cells = h->d_cells;
	a.params = h->d_params;
	a.set = h->cur_set;
	a.cover = h->d_cover;
	a.diag = h->d_diag;
	a.mn = h->d_mn;
	a.dtab = h->d_dtab;
	a.sb = h->d_sp;
	a.sx = h->d_p2;
	a.c = h->d_sc;
	a.s = h->d_ss;
	for(int i = 0; i < 2; ++i) {
		a.y[i] = h->d_sy[i];
		a.q[i] = h->d_sq[i];
		a.psum[i] = h->d_psum[i];
	}
	a.b = h->d_f;
	a.x = h->d_v;
	a.part = h->d_part;
	a.s2 = h->d_scg;
	a.cgb = h->d_cgb;
	a.geo = h->d_cg_geo;
	a.one_L = h->cgb[0].L;
	a.one_ntx = h->cgb[0].ntx;
	a.one_eta = h->dparams[2 * 0 + h->cur_set].friction_cell;
	a.max_box_tiles = 0;
	for(const SCGBox &g : h->cgb) a.max_box_tiles = std::max(a.max_box_tiles, g.ntiles);
	a.ncells = h->ncells;
	a.ntiles = h->cg_ntiles;
	a.nboxes = h->nboxes;
	a.st = h->d_st;
	return a;
}

// initial residual |b - A x|^2 (x + y components) of the rows in c.x
static int resid_norm(rc_handle *h, const CGArgs &c, double *rn2) {
	HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
	HIPCHK(h, launch_cg_box(0, c, (int) h->box_sites, h->stream));
	HIPCHK(h, launch_cg_row(CG_ROW_RESID, c, h->stream));
	HIPCHK(h, launch_cg_reduce(0, c, h->ncells, h->stream));
	HIPCHK(h, hipMemcpyAsync(h->h_cg, h->d_cg, sizeof(CGState), hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	*rn2 = h->h_cg->rn2[0] + h->h_cg->rn2[1];
	return RC_OK;
}

// One refinement stage of the learned initial guess.  With x the current
// start and r = b - A x its residual, the network predicts the box sums of the
// correction e = A^-1 r from the box sums of r (scaled by the residual size of
// each component) and sum phi^2; rows follow from A e = r, i.e.
// e_i = (r_i + eta N4(k_i)) / D(k_i).  x moves along e by the A-norm optimal
// step alpha_c = (e.r)/(e.A e), which never raises the A-norm error, and
// r -= alpha A e.  e.A e comes from box sums like the CG's p.Ap.  The last
// reduction leaves the CG set-up of the new start.  Launches:
//   gather<FEAT>  box sums of r -> network input      conv x6  network
//   box<2>        N4 of the predicted sums
//   gather<E>     e rows, box sums of e and e^2, e.r  box<1>  N4, e.A e
//   reduce<3>     alpha                               row<STEP>  x, r; r.r, r.z
//   reduce<4>     CG set-up
static int learned_stage(rc_handle *h, const CGArgs &c0) {
	const int ns = (int) h->box_sites;
	CGArgs c = c0;
	c.act0 = net_input(h->net);
	c.phi2 = h->d_phi2;
	c.net_out = net_output(h->net);
	c.sf = h->scale[0];
	c.sp2 = h->scale[1];
	c.sv = h->scale[2];
	c.rows = (double) h->rows;
	c.nbox_blk = (ns + 255) / 256;
	HIPCHK(h, launch_cg_gather(CG_GATHER_FEAT, c, h->ntiles, h->stream));
	// the solve's active flag: stages after convergence skip the network
	const int *run = reinterpret_cast<const int *>(reinterpret_cast<const char *>(h->d_cg) + offsetof(CGState, active));
	if(net_forward_packed(h->net, h->stream, run))
		return fail(h, RC_ERR_HIP, "network forward pass failed: %s", rc_net_last_error(h->net));
	HIPCHK(h, launch_cg_box(2, c, ns, h->stream));
	HIPCHK(h, launch_cg_gather(CG_GATHER_E, c, h->ntiles, h->stream));
	HIPCHK(h, launch_cg_box(1, c, ns, h->stream));
	HIPCHK(h, launch_cg_reduce(3, c, h->ntiles, h->stream));
	HIPCHK(h, launch_cg_row(CG_ROW_STEP, c, h->stream));
	HIPCHK(h, launch_cg_reduce(4, c, h->ncells, h->stream));
	return RC_OK;
}

// Training records, [L*L][5] each: box sums of the residual r of a start x
// (divided by the residual scale sqrt(|r_c|^2 / rows)), sum phi^2, and box
// sums of the correction x_solution - x (same scale).  record_stage runs when
// the residual of a start is in place (once per solve from the previous
// solution, once per stage with the network); record_end after the solve.
static int record_stage(rc_handle *h, const CGArgs &c0) {
	if(h->rec_n + h->rec_open >= h->rec_cap || h->rec_open >= RC_MAX_STAGES) return RC_OK;
	const int ns = (int) h->box_sites, s = h->rec_open;
	const double rows = (double) h->rows;
	CGArgs c = c0;
	float *dst = h->d_rec + (size_t) (h->rec_n + s) * 5 * ns;
	HIPCHK(h, hipMemcpyAsync(h->d_xrec + (size_t) s * h->rows, h->d_v, sizeof(double2) * h->rows, hipMemcpyDeviceToDevice,
			h->stream));
	HIPCHK(h, hipMemcpyAsync(h->d_rnrec + 2 * s, reinterpret_cast<char *>(h->d_cg) + offsetof(CGState, rn2),
			2 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
	c.x = h->d_r;
	HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
	HIPCHK(h, launch_box_snapshot(c, h->d_phi2, dst, 0, h->d_rnrec + 2 * s, rows, ns, h->stream));
	HIPCHK(h, launch_box_snapshot(c, h->d_phi2, dst, 2, nullptr, rows, ns, h->stream));
	h->rec_open += 1;
	return RC_OK;
}
static int record_end(rc_handle *h) {
	const int ns = (int) h->box_sites;
	const double rows = (double) h->rows;
	CGArgs c = cg_args(h);
	c.x = h->d_v;
	HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
	for(int s = 0; s < h->rec_open; ++s)
		HIPCHK(h, launch_box_snapshot(c, h->d_phi2, h->d_rec + (size_t) (h->rec_n + s) * 5 * ns, 3, h->d_rnrec + 2 * s,
				rows, ns, h->stream));
	for(int s = 0; s < h->rec_open; ++s) {
		c.x = h->d_xrec + (size_t) s * h->rows;
		HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
		HIPCHK(h, launch_box_snapshot(c, h->d_phi2, h->d_rec + (size_t) (h->rec_n + s) * 5 * ns, 5, h->d_rnrec + 2 * s,
				rows, ns, h->stream));
	}
	h->rec_n += h->rec_open;
	h->rec_open = 0;
	return RC_OK;
}

// Site form of the iteration (SCGState), from the start rows in d_v: the box
// sums of b and x0, the set-up, then iterations in host batches, each ended
// by a stop test the host reads; the solution rows replace the start in d_v
// and the last state is in h_scg.  The first batch is two iterations longer
// than the last solve of the same kind: a launch after convergence returns at
// entry (about 1 us), far less than another host round trip.
static int scg_iterations(rc_handle *h, const CGArgs &c, bool cold) {
	const SCGArgs a = scg_args(h, c);
	{ TimeScope t(h, RC_T_SCG_GATHER); HIPCHK(h, launch_cg_gather(CG_GATHER_BX, c, h->ntiles, h->stream)); }
	{ TimeScope t(h, RC_T_SCG_INIT); HIPCHK(h, launch_scg(SCG_INIT, a, 0, h->stream)); }
	{ TimeScope t(h, RC_T_SCG_ROW_INIT); HIPCHK(h, launch_scg(SCG_ROW_INIT, a, 0, h->stream)); }
	{ TimeScope t(h, RC_T_SCG_INIT_W); HIPCHK(h, launch_scg(SCG_INIT_W, a, 0, h->stream)); }
	int batch = std::max(2, h->last_iters[cold ? 1 : 0] + 2), next = 4;
	int j = 0, cap = 0;
	for(const SCGBox &g : h->cgb) cap = std::max(cap, g.max_iter);
	for(;;) {
		for(int it = 0; it < batch; ++it, ++j) {
			TimeScope t(h, RC_T_SCG_ITER);
			HIPCHK(h, launch_scg(SCG_ITER, a, j, h->stream));
		}
		{ TimeScope t(h, RC_T_SCG_REDUCE); HIPCHK(h, launch_scg(SCG_CHECK, a, j, h->stream)); }
		HIPCHK(h, hipMemcpyAsync(h->h_scg, h->d_scg, 2 * h->nboxes * sizeof(SCGState), hipMemcpyDeviceToHost, h->stream));
		HIPCHK(h, hipMemcpyAsync(h->h_st, h->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, h->stream));
		HIPCHK(h, hipStreamSynchronize(h->stream));
		if(h->h_st->flags) break;
		bool active = false;
		for(int b = 0; b < h->nboxes; ++b) active = active || h->h_scg[2 * b + (j & 1)].active_now;
		if(!active) break;
		if(j >= cap + 2) break;
		batch = next;
		next = std::min(2 * next, 64);
	}
	h->scg_parity = j & 1;
	{ TimeScope t(h, RC_T_SCG_FINAL); HIPCHK(h, launch_scg(SCG_FINAL, a, j, h->stream)); }
	return RC_OK;
}

// Row form of the iteration (CGState): p = z fused with the box sums of p,
// then iterations in host batches; the solution rows in d_v.
static int cg_row_iterations(rc_handle *h, CGArgs c, bool cold) {
	{ TimeScope t(h, RC_T_CG_GATHER_P_FIRST); HIPCHK(h, launch_cg_gather(CG_GATHER_P_FIRST, c, h->ntiles, h->stream)); }
	// First batch: one iteration short of the last solve of the same kind
	// (warm or cold: a cold solve needs several times the iterations);
	// then batches of 2, 4, 8, ... so that a poor guess costs few host
	// round trips.  Iterations past convergence return at entry.
	int batch = std::max(2, h->last_iters[cold ? 1 : 0] - 1), next = 2;
	long long done_iters = 0;
	for(;;) {
		// an iteration's state update is written by the next iteration's
		// k_cg_box<1> (its beta is derived in k_cg_gather), the batch's last
		// by k_cg_reduce<2>
		for(int it = 0; it < batch; ++it) {
			c.apply_update = it > 0 ? 1 : 0;
			{ TimeScope t(h, RC_T_CG_BOX_PAP); HIPCHK(h, launch_cg_box(1, c, (int) h->box_sites, h->stream)); }
			{ TimeScope t(h, RC_T_CG_ROW_UPDATE); HIPCHK(h, launch_cg_row(CG_ROW_UPDATE, c, h->stream)); }
			{ TimeScope t(h, RC_T_CG_GATHER_P); HIPCHK(h, launch_cg_gather(CG_GATHER_P_UPDATE, c, h->ntiles, h->stream)); }
		}
		c.apply_update = 0;
		{ TimeScope t(h, RC_T_CG_REDUCE); HIPCHK(h, launch_cg_reduce(2, c, h->ncells, h->stream)); }
		done_iters += batch;
		HIPCHK(h, hipMemcpyAsync(h->h_cg, h->d_cg, sizeof(CGState), hipMemcpyDeviceToHost, h->stream));
		HIPCHK(h, hipMemcpyAsync(h->h_st, h->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, h->stream));
		HIPCHK(h, hipStreamSynchronize(h->stream));
		if(h->h_st->flags) break;
		if(!h->h_cg->active) break;
		if(done_iters >= c.max_iter + 2) break;
		batch = next;
		next = std::min(2 * next, 64);
	}
	const CGState &s = *h->h_cg;
	int mask = (s.zero_rhs[0] ? 1 : 0) | (s.zero_rhs[1] ? 2 : 0);
	if(mask) HIPCHK(h, launch_zero_rows(h->d_v, h->rows, mask, h->stream));
	return RC_OK;
}

// The time of the last solve, when its end event was recorded without a wait
// (the site form needs no host read after its last batch): read before the
// events are recorded again and when the statistics are asked for.
static void fold_cg_ms(rc_handle *h) {
	if(!h->cg_ms_pending) return;
	h->cg_ms_pending = false;
	if(hipEventSynchronize(h->cg_e1) != hipSuccess) return;
	float ms = 0.f;
	(void) hipEventElapsedTime(&ms, h->cg_e0, h->cg_e1);
	h->cgs.last_solve_ms = ms;
	h->cgs.total_solve_ms += ms;
}

// Eigen ConjugateGradient solve (solve() when cold, solveWithGuess() otherwise)
static int cg_solve(rc_handle *h, bool cold) {
	fold_cg_ms(h);
	CGArgs c = cg_args(h);
	const bool sites = h->cg_form == RC_CG_SITES;
	hipEvent_t e0 = h->cg_e0, e1 = h->cg_e1;
	HIPCHK(h, hipEventRecord(e0, h->stream));
	{
		TimeScope ts(h, RC_T_CG_SOLVE);
		// solve() of Eigen (x0 = 0) in the boxes restarting cold, solveWithGuess() in the others
		for(int b = 0; b < h->nboxes; ++b)
			if(h->cold_box.empty() ? cold : h->cold_box[b])
				HIPCHK(h, hipMemsetAsync(h->d_v + h->box_rows[b], 0, sizeof(double2) * (h->box_rows[b + 1] - h->box_rows[b]),
						h->stream));
		const bool learned = h->warm_mode == RC_WARM_LEARNED && h->net;
		// the row form's start (residual rows and their sums): the row
		// iteration, the learned stages and the training records work on it
		// (the site form's set-up computes D itself)
		if(!sites || learned || h->d_rec) {
			if(h->nboxes != 1) return fail(h, RC_ERR_ARG, "the row form of the CG runs one box per handle");
			{ TimeScope t(h, RC_T_CG_DIAG); HIPCHK(h, launch_cg_diag(c, (int) h->box_sites, h->stream)); }
			{ TimeScope t(h, RC_T_CG_GATHER_X); HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream)); }
			{ TimeScope t(h, RC_T_CG_BOX_N4); HIPCHK(h, launch_cg_box(0, c, (int) h->box_sites, h->stream)); }
			{ TimeScope t(h, RC_T_CG_ROW_RESID); HIPCHK(h, launch_cg_row(CG_ROW_RESID, c, h->stream)); }
			{ TimeScope t(h, RC_T_CG_REDUCE); HIPCHK(h, launch_cg_reduce(0, c, h->ncells, h->stream)); }
		}
		if(h->d_rec) {
			int rc = record_stage(h, c);
			if(rc) return rc;
		}
		if(learned) {
			TimeScope t(h, RC_T_CG_LEARNED);
			// keep the previous-solution start for the fallback
			HIPCHK(h, hipMemcpyAsync(h->d_xprev, h->d_v, sizeof(double2) * h->rows, hipMemcpyDeviceToDevice, h->stream));
			HIPCHK(h, hipMemcpyAsync(h->d_rprev, h->d_r, sizeof(double2) * h->rows, hipMemcpyDeviceToDevice, h->stream));
			HIPCHK(h, hipMemcpyAsync(reinterpret_cast<char *>(h->d_cg) + offsetof(CGState, rn2_start),
					reinterpret_cast<char *>(h->d_cg) + offsetof(CGState, rn2), 2 * sizeof(double), hipMemcpyDeviceToDevice,
					h->stream));
			for(int st = 0; st < h->stages; ++st) {
				int rc = learned_stage(h, c);
				if(rc) return rc;
				if(h->d_rec && st + 1 < h->stages) {
					rc = record_stage(h, c);
					if(rc) return rc;
				}
			}
			// back to the previous solution in a component whose residual got larger
			CGArgs f = c;
			f.xprev = h->d_xprev;
			f.rprev = h->d_rprev;
			HIPCHK(h, launch_cg_row(CG_ROW_REVERT, f, h->stream));
			HIPCHK(h, launch_cg_reduce(5, f, h->ncells, h->stream));
			h->cgs.learned_tried += 1;
			h->pending_learned_stat = true;
		}
		if(sites) {
			int rc = scg_iterations(h, c, cold);
			if(rc) return rc;
			// the learned start's fallback flags, for the statistics below
			if(h->pending_learned_stat) HIPCHK(h, hipMemcpyAsync(h->h_cg, h->d_cg, sizeof(CGState), hipMemcpyDeviceToHost, h->stream));
		}
		else {
			int rc = cg_row_iterations(h, c, cold);
			if(rc) return rc;
		}
	}
	// a halted device ran no iteration: no statistics or training records
	// for this solve
	if(h->h_st->flags) {
		h->pending_learned_stat = false;
		h->rec_open = 0;
		return RC_OK;
	}
	HIPCHK(h, hipEventRecord(e1, h->stream));
	// the row form, the learned start and the training records read device
	// state below; the site form's state is already on the host, so its end
	// is not waited for here (the stream runs on into the next stage)
	const bool wait = !sites || h->pending_learned_stat || h->rec_open;
	float ms = 0.f;
	if(wait) {
		HIPCHK(h, hipEventSynchronize(e1));
		(void) hipEventElapsedTime(&ms, e0, e1);
	}
	else h->cg_ms_pending = true;
	if(h->pending_learned_stat) {
		if(!h->h_cg->reverted[0] && !h->h_cg->reverted[1]) h->cgs.learned_used += 1;
		h->pending_learned_stat = false;
	}
	if(h->rec_open) {
		int rc = record_end(h);
		if(rc) return rc;
	}
	// iterations, final |r|^2 and |b|^2 of each component of each box, from
	// either form's state: sums over the boxes' solves, the largest of the
	// last iteration counts and relative residuals
	int last_it[2] = { 0, 0 }, sweeps = 0;
	bool capped = false;
	for(int k = 0; k < 2; ++k) h->cgs.last_relres[k] = 0.;
	for(int b = 0; b < (sites ? h->nboxes : 1); ++b) {
		const SCGState *S = sites ? h->h_scg + 2 * b + h->scg_parity : nullptr;
		const int cap = sites ? S->max_iter : c.max_iter;
		for(int k = 0; k < 2; ++k) {
			const int it = sites ? S->iters[k] : h->h_cg->iters[k];
			const double rn2 = sites ? S->rn2_now[k] : h->h_cg->rn2[k], bb = sites ? S->bb[k] : h->h_cg->bb[k];
			const double rel = bb > 0 ? std::sqrt(rn2 / bb) : 0.;
			h->cgs.solves += 1;
			h->cgs.iters_total += it;
			last_it[k] = std::max(last_it[k], it);
			h->cgs.last_relres[k] = std::max(h->cgs.last_relres[k], rel);
			h->cgs.worst_relres = std::max(h->cgs.worst_relres, rel);
			capped = capped || it >= cap;
		}
		sweeps = std::max(sweeps, sites ? std::max(S->iters[0], S->iters[1]) : std::max(h->h_cg->iters[0], h->h_cg->iters[1]));
	}
	h->cgs.sweeps_total += sweeps;
	h->cgs.last_iters[0] = last_it[0];
	h->cgs.last_iters[1] = last_it[1];
	h->cgs.last_cold = cold ? 1 : 0;
	if(wait) {
		h->cgs.last_solve_ms = ms;
		h->cgs.total_solve_ms += ms;
	}
	h->last_iters[cold ? 1 : 0] = std::max(last_it[0], last_it[1]);
	if(capped) h->cgs.last_cold |= 2;
	return RC_OK;
}

// WetModel::begin_energy_computation(fields) (WetModel.cpp:194-233) up to the
// solve: the box sums of phi Q come from the gather of phi^2 (run_gather, in
// the same pass), then the forces of every cell (k_nem_force): the CG
// right-hand side, or the dry velocity
static int nem_forces(rc_handle *h, const double *phi) {
	NemArgs n{};
	n.boxes = h->d_boxes;
	n.cells = h->d_cells;
	n.params = h->d_params;
	n.set = h->cur_set;
	n.phi = phi;
	n.phi2 = h->d_phi2;
	n.sq00 = h->d_sq00;
	n.sq01 = h->d_sq01;
	n.frows = h->d_f;
	n.vrows = h->d_v;
	n.wmask = h->d_wmask;
	n.walls = h->d_walls;
	n.st = h->d_st;
	TimeScope ts(h, RC_T_NEM_FORCE);
	HIPCHK(h, launch_nem_force(n, h->ncells, nem_lds_bytes(h->max_S, h->d_wmask != nullptr), h->stream));
	return RC_OK;
}

// WetPolarModel::begin_energy_computation(fields) on the wet path: box sums,
// right-hand side (or dry velocity) and, for the wet solve, the CG.
static int nem_forces(rc_handle *h, const double *phi);

static int force_eval(rc_handle *h, double *phi, bool pure) {
	int rc;
	if((rc = run_gather(h, phi))) return rc;
	if(h->nematic) {
		if((rc = nem_forces(h, phi))) return rc;
	}
	else {
		StageArgs a = stage_args(h);
		a.cf_set = h->cur_set;
		a.fs_set = h->cur_set;
		a.phi_c = phi;
		a.pure = pure ? 1 : 0;
		if((rc = run_stage(h, MODE_CF, a))) return rc;
	}
	// restart bookkeeping of WetPolarModel.cpp:151-157: a cold solve after 10
	// warm ones, after 100 in WetModel (WetModel.cpp:196-201)
	// (each box is a simulation of its own: its own rows and counter)
	const int restart_after = h->nematic ? 100 : 10;
	const int nb = h->nboxes;
	if((int) h->restart_solver.size() != nb) {
		h->restart_solver.assign(nb, 0);
		h->size_rows.assign(nb, 0);
		h->size_rows_old.assign(nb, 0);
	}
	const std::vector<int> saved_restart = h->restart_solver;
	const std::vector<long long> saved_rows = h->size_rows, saved_rows_old = h->size_rows_old;
	h->cold_box.assign(nb, 0);
	bool any_cold = false;
	for(int b = 0; b < nb; ++b) {
		if(h->restart_solver[b] < restart_after || h->warm_mode != RC_WARM_REFERENCE) h->size_rows_old[b] = h->size_rows[b];
		else {
			h->size_rows_old[b] = 0;
			h->restart_solver[b] = 0;
		}
		h->size_rows[b] = h->box_rows[b + 1] - h->box_rows[b];
		h->cold_box[b] = h->size_rows[b] != h->size_rows_old[b];
		any_cold = any_cold || h->cold_box[b];
	}
	if(h->dparams[2 * 0 + h->cur_set].friction_cell != 0) {
		for(int b = 0; b < nb; ++b) {
			if(h->cold_box[b]) h->restart_solver[b] = 0;
			else h->restart_solver[b] += 1;
		}
		if((rc = cg_solve(h, any_cold))) return rc;
		// a halted device (a growth request of the corrector before this
		// evaluation, or an error) ran no solve: the counters stay as they
		// were, for the evaluation made again after the growth
		if(h->h_st->flags) {
			h->restart_solver = saved_restart;
			h->size_rows = saved_rows;
			h->size_rows_old = saved_rows_old;
		}
	}
	h->cf_set = h->cur_set;
	return RC_OK;
}

static int check_status(rc_handle *h) {
	HIPCHK(h, hipMemcpyAsync(h->h_st, h->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	const unsigned f = h->h_st->flags;
	if(f & ST_OVERLAP) return fail(h, RC_ERR_OVERLAP, "Too many field patches overlap at box site %d", h->h_st->site);
	if(f & ST_TILE) return fail(h, RC_ERR_TILE, "more than %d patches intersect one 32x8 box tile", MAXT);
	if(f & ST_NAN) {
		if(h->h_st->first_cell >= 0) return fail(h, RC_ERR_NAN, "the field of cell %d became NaN", h->h_st->first_cell);
		return fail(h, RC_ERR_NAN, "the wet solve's right-hand side or residual became NaN");
	}
	if(f & ST_GROWTH)
		return fail(h, RC_ERR_GROWTH, "patch of cell %d must grow (MultiPhaseField::set_positions growth branch)", h->h_st->first_cell);
	return RC_OK;
}

extern "C" int rc_init_forces(rc_handle *h) {
	if(!h) return RC_ERR_ARG;
	if(!h->uploaded) return fail(h, RC_ERR_STATE, "rc_init_forces before rc_upload_cells");
	h->cur_set = 0;
	h->cf_set = 0;
	h->restart_solver.assign(h->nboxes, 0);
	h->size_rows.assign(h->nboxes, 0);
	h->size_rows_old.assign(h->nboxes, 0);
	if(h->split) {
		int rc = force_eval(h, h->d_phiA, false);
		if(rc) return rc;
	}
	h->initialized = true;
	return check_status(h);
}

extern "C" int rc_apply_changes_after_equilibration(rc_handle *h) {
	if(!h) return RC_ERR_ARG;
	h->cur_set = 1;
	return RC_OK;
}

extern "C" int rc_set_cg_form(rc_handle *h, int form) {
	if(!h) return RC_ERR_ARG;
	if(form != RC_CG_SITES && form != RC_CG_ROWS) return fail(h, RC_ERR_ARG, "unknown CG form %d", form);
	if(form == RC_CG_ROWS && h->nboxes != 1) return fail(h, RC_ERR_ARG, "the row form of the CG runs one box per handle");
	h->cg_form = form;
	return RC_OK;
}

extern "C" int rc_set_warm_start(rc_handle *h, int mode) {
	if(!h) return RC_ERR_ARG;
	if(mode != RC_WARM_REFERENCE && mode != RC_WARM_PREVIOUS && mode != RC_WARM_LEARNED)
		return fail(h, RC_ERR_ARG, "unknown warm-start mode %d", mode);
	if(mode == RC_WARM_LEARNED && !h->net) return fail(h, RC_ERR_STATE, "RC_WARM_LEARNED needs rc_attach_net first");
	h->warm_mode = mode;
	return RC_OK;
}

static int step_dry(rc_handle *h, const double *gauss_dev, int step) {
	int rc;
	// predictor
	if((rc = run_gather(h, h->d_phiA))) return rc;
	StageArgs a = stage_args(h);
	a.cf_set = h->cf_set;
	a.fs_set = h->cur_set;
	a.store = 1;
	a.step = step;
	a.phi_c = h->d_phiA;
	a.phi_out = h->d_phiB;
	a.gauss = gauss_dev;
	if((rc = run_stage(h, MODE_FUSED, a))) return rc;
	// corrector
	if((rc = run_gather(h, h->d_phiB))) return rc;
	a.cf_set = h->cur_set;
	a.store = 0;
	a.phi_c = h->d_phiB;
	a.phi_old = h->d_phiA;
	a.phi_out = h->d_phiA;
	a.gauss = nullptr;
	if((rc = run_stage(h, MODE_FUSED, a))) return rc;
	h->cf_set = h->cur_set;
	return RC_OK;
}

// the split schedule (the wet model, every nematic box): field update, then the
// force evaluation, per stage.  Returns RC_OK, an error, or -1 when the device
// halted (status flag set)
static int step_split(rc_handle *h, const double *gauss_dev, int step) {
	int rc;
	StageArgs a = stage_args(h);
	a.cf_set = h->cf_set;
	a.fs_set = h->cur_set;
	a.store = 1;
	a.step = step;
	a.phi_c = h->d_phiA;
	a.phi_out = h->d_phiB;
	a.gauss = gauss_dev;
	if((rc = run_stage(h, MODE_FS, a))) return rc;
	if((rc = force_eval(h, h->d_phiB, false))) return rc;
	if(h->h_st->flags) return -1;
	a.cf_set = h->cf_set;
	a.fs_set = h->cur_set;
	a.store = 0;
	a.phi_c = h->d_phiB;
	a.phi_old = h->d_phiA;
	a.phi_out = h->d_phiA;
	a.gauss = nullptr;
	if((rc = run_stage(h, MODE_FS, a))) return rc;
	// A growth request of this corrector must be served before the force
	// evaluation that follows it.  That evaluation is enqueued without a wait:
	// on a request its launches return at entry, the solve's first host read
	// sees the request and the evaluation leaves no trace (force_eval), and
	// the step ends here for the host to serve the growth and evaluate again.
	if((rc = force_eval(h, h->d_phiA, false))) return rc;
	if(h->h_st->flags) return -1;
	return RC_OK;
}

static void to_rc_cell(const DevCell &d, rc_cell &c) {
	c.LsubX = d.LX;
	c.LsubY = d.LY;
	c.offset[0] = d.offx;
	c.offset[1] = d.offy;
	c.corner = d.corner;
	c.corner_old = d.corner_old;
	c.x_sub_left = d.x_sub_left;
	c.y_sub_bottom = d.y_sub_bottom;
	c.CoM[0] = d.CoMx;
	c.CoM[1] = d.CoMy;
	c.CoM_old[0] = d.CoM_oldx;
	c.CoM_old[1] = d.CoM_oldy;
	c.unrap[0] = d.unrapx;
	c.unrap[1] = d.unrapy;
	c.Q00 = d.Q00;
	c.Q01 = d.Q01;
	c.thetaQ = d.thetaQ;
	c.thetaQ_old = d.thetaQ_old;
	c.area = d.area;
	c.area_old = d.area_old;
	c.sumF = d.sumF;
	c.sumF_old = d.sumF_old;
	c.perimeter = d.perimeter;
	c.S00 = d.S00;
	c.S01 = d.S01;
	c.total_force_repulsion[0] = d.trx;
	c.total_force_repulsion[1] = d.try_;
	c.nemQ[0] = d.nem0;
	c.nemQ[1] = d.nem1;
	c.nemQ_old[0] = d.nem_old0;
	c.nemQ_old[1] = d.nem_old1;
}

static int download_raw(rc_handle *h, std::vector<rc_cell> &cells, std::vector<double> &phi, std::vector<double> *rep) {
	std::vector<DevCell> dc(h->ncells);
	cells.resize(h->ncells);
	phi.resize(h->rows);
	HIPCHK(h, hipMemcpyAsync(dc.data(), h->d_cells, sizeof(DevCell) * h->ncells, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipMemcpyAsync(phi.data(), h->d_phiA, sizeof(double) * h->rows, hipMemcpyDeviceToHost, h->stream));
	if(rep && h->jq) {
		rep->resize(h->rows);
		HIPCHK(h, hipMemcpyAsync(rep->data(), h->d_rep, sizeof(double) * h->rows, hipMemcpyDeviceToHost, h->stream));
	}
	HIPCHK(h, hipStreamSynchronize(h->stream));
	for(int i = 0; i < h->ncells; ++i) to_rc_cell(dc[i], cells[i]);
	return RC_OK;
}

// MultiPhaseField::set_positions, growth branch (MultiPhaseField.cpp:248-310):
// the geometry of the re-cut around the interface bounding box (border 4
// sites), on the host from the cell record alone; the field is re-cut on the
// device (k_grow_rows) from the entry this fills.  The area (sum phi^2 of the
// copied sites) comes back from the device.
static int grow_geometry(rc_handle *h, rc_cell &c, int side, long long old_start, long long new_start, int4 *ent) {
	auto bx = [side](int site, int d) {
		int x = (site - ((site / side) * side)) + d;
		while(x < 0) x += side;
		while(x >= side) x -= side;
		return x;
	};
	auto by = [side](int site, int d) {
		int y = (site / side) + d;
		while(y < 0) y += side;
		while(y >= side) y -= side;
		return y;
	};
	const int LX = c.LsubX, LY = c.LsubY;
	const int nLX = LX - (2 * c.x_sub_left) + (2 * BORDER);
	const int nLY = LY - (2 * c.y_sub_bottom) + (2 * BORDER);
	if(nLX < 3 || nLY < 3 || nLX > side || nLY > side)
		return fail(h, RC_ERR_GROWTH, "patch growth to %dx%d does not fit box %d", nLX, nLY, side);
	if(sizeof(int) * (size_t) nLX * nLY > GROW_LDS_MAX)
		return fail(h, RC_ERR_GROWTH, "patch growth to %dx%d exceeds the device re-cut (%d sites)", nLX, nLY,
				(int) (GROW_LDS_MAX / sizeof(int)));
	const int ncorner = bx(c.corner, -(-(2 * c.x_sub_left) + (2 * BORDER)) / 2)
			+ by(c.corner, -(-(2 * c.y_sub_bottom) + (2 * BORDER)) / 2) * side;
	int sy, sx, dimy, dimx;
	if(nLY < LY) { sy = (LY - nLY) / 2; dimy = nLY; } else { sy = 0; dimy = LY; }
	if(nLX < LX) { sx = (LX - nLX) / 2; dimx = nLX; } else { sx = 0; dimx = LX; }
	ent[0] = make_int4((int) old_start, (int) new_start, LX, LY);
	ent[1] = make_int4(c.offset[0], c.offset[1], c.corner, side);
	ent[2] = make_int4(nLX, nLY, ncorner, sx);
	ent[3] = make_int4(sy, dimx, dimy, 0);
	const double c0x = bx(c.corner, 0), c0y = by(c.corner, 0);
	const double c1x = bx(ncorner, 0), c1y = by(ncorner, 0);
	c.unrap[0] += c1x - c0x - std::rint((c1x - c0x) / side) * side;
	c.unrap[1] += c1y - c0y - std::rint((c1y - c0y) / side) * side;
	c.LsubX = nLX;
	c.LsubY = nLY;
	c.corner_old = c.corner;
	c.corner = ncorner;
	c.offset[0] = 0;
	c.offset[1] = 0;
	const int site = bx(ncorner, nLX / 2) + by(ncorner, nLY / 2) * side;
	c.CoM[1] = (int) (site / side);
	c.CoM[0] = (int) (site - (int) (site / side) * side);
	return RC_OK;
}

// serves a growth request raised by the device: re-cuts the flagged patches
// on the device, re-lays out the rows and clears the halt
static int serve_growth(rc_handle *h) {
	// cell state (geometry only crosses to the host); every other cell's rows
	// move to their new offsets (k_repack_rows), the growing ones are re-cut
	// (k_grow_rows)
	std::vector<DevCell> dc(h->ncells);
	HIPCHK(h, hipMemcpyAsync(dc.data(), h->d_cells, sizeof(DevCell) * h->ncells, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	std::vector<rc_cell> cells(h->ncells);
	for(int i = 0; i < h->ncells; ++i) to_rc_cell(dc[i], cells[i]);
	std::vector<int4> map, ent;
	std::vector<int> grown;
	long long nrows = 0;
	int max_sites = 0, rc;
	for(int b = 0; b < h->nboxes; ++b) {
		const DevBox &B = h->dboxes[b];
		for(int i = B.cell_begin; i < B.cell_end; ++i) {
			rc_cell &c = cells[i];
			const int S = c.LsubX * c.LsubY;
			const long long old = dc[i].start;
			if(c.x_sub_left >= BORDER && c.y_sub_bottom >= BORDER) {
				map.push_back(make_int4((int) old, (int) nrows, S, 0));
				nrows += S;
				continue;
			}
			int4 e[4];
			if((rc = grow_geometry(h, c, B.L, old, nrows, e))) return rc;
			ent.insert(ent.end(), e, e + 4);
			grown.push_back(i);
			const int nS = c.LsubX * c.LsubY;
			max_sites = std::max(max_sites, nS);
			nrows += nS;
		}
	}
	double *d_nphi = nullptr, *d_nrep = nullptr, *d_area = nullptr;
	int4 *d_map = nullptr, *d_ent = nullptr;
	const int ng = (int) grown.size();
	std::vector<double> area(ng);
	HIPCHK(h, dalloc(&d_nphi, nrows));
	if(h->jq) HIPCHK(h, dalloc(&d_nrep, nrows));
	HIPCHK(h, dalloc(&d_map, std::max<size_t>(map.size(), 1)));
	HIPCHK(h, dalloc(&d_ent, std::max<size_t>(ent.size(), 1)));
	HIPCHK(h, dalloc(&d_area, std::max(ng, 1)));
	if(!map.empty())
		HIPCHK(h, hipMemcpyAsync(d_map, map.data(), sizeof(int4) * map.size(), hipMemcpyHostToDevice, h->stream));
	if(!ent.empty())
		HIPCHK(h, hipMemcpyAsync(d_ent, ent.data(), sizeof(int4) * ent.size(), hipMemcpyHostToDevice, h->stream));
	HIPCHK(h, launch_repack_rows(h->d_phiA, d_nphi, h->jq ? h->d_rep : nullptr, d_nrep, d_map, (int) map.size(),
			h->stream));
	HIPCHK(h, launch_grow_rows(h->d_phiA, d_nphi, h->jq ? h->d_rep : nullptr, d_nrep, d_ent, ng, max_sites, d_area,
			h->stream));
	if(ng) HIPCHK(h, hipMemcpyAsync(area.data(), d_area, sizeof(double) * ng, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	rc = RC_OK;
	for(int k = 0; k < ng; ++k) {
		if(!(area[k] == area[k])) {
			rc = fail(h, RC_ERR_GROWTH, "patch growth of cell %d writes outside the new patch", grown[k]);
			break;
		}
		cells[grown[k]].area = area[k];
	}
	if(!rc) rc = upload_layout(h, cells.data(), d_nphi, d_nrep, true);
	(void) hipStreamSynchronize(h->stream);
	dfree(d_nphi);
	dfree(d_nrep);
	dfree(d_map);
	dfree(d_ent);
	dfree(d_area);
	if(rc) return rc;
	h->h_st->flags = 0;
	return RC_OK;
}

// Stage-shape autotuner: for each group, times the predictor stage of its
// cells (FUSED dry / FS wet) on the current fields with every shape that fits
// the group, 3 launches each after a warm-up, and keeps the fastest.  The
// timing launches are side-effect free: they write phi* and dphi rows (the
// predictor's own outputs, recomputed by the next step) and no cell state or
// Repulsion.  Runs before the first step after a layout change.
static int autotune_stage(rc_handle *h) {
	h->stage_tuned = true;
	fold_cg_ms(h);   // its timing reuses the solve's events
	int rc;
	if((rc = run_gather(h, h->d_phiA))) return rc;
	const bool timing = h->timing;
	h->timing = false;
	hipEvent_t e0 = h->cg_e0, e1 = h->cg_e1;
	for(auto &g : h->groups) {
		if(g.valid.size() < 2) continue;
		double best = 1e300;
		std::pair<int, int> pick = g.valid[0];
		for(size_t i = 0; i < g.valid.size(); ++i) {
			StageArgs a = stage_args(h);
			a.cf_set = h->cf_set;
			a.fs_set = h->cur_set;
			a.store = 1;
			a.pure = 1;
			a.write_v = 0;
			a.phi_c = h->d_phiA;
			a.phi_out = h->d_phiB;
			a.gauss = nullptr;
			a.threads = g.valid[i].first;
			a.npt = g.valid[i].second;
			a.cell_list = h->groups.size() > 1 ? g.d_cells : nullptr;
			const int n = (int) g.cells.size();
			const int mode = h->split ? MODE_FS : MODE_FUSED;
			const size_t lds = stage_lds_bytes(g.mx, g.my, a.npt, h->jq, h->win);
			HIPCHK(h, launch_stage(mode, a, n, lds, h->stream));
			HIPCHK(h, hipEventRecord(e0, h->stream));
			for(int r = 0; r < 3; ++r) HIPCHK(h, launch_stage(mode, a, n, lds, h->stream));
			HIPCHK(h, hipEventRecord(e1, h->stream));
			HIPCHK(h, hipEventSynchronize(e1));
			float ms = 0.f;
			HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
			const double us = 1e3 * ms / 3;
			if(i < 8) g.tune_us[i] = us;
			if(us < best) {
				best = us;
				pick = g.valid[i];
			}
		}
		g.threads = pick.first;
		g.rows = pick.second;
	}
	// the persistent form against the groups with their chosen shapes
	h->persist = false;
	h->persist_us = h->groups_us = 0.;
	if(h->persist_ok && h->ncells >= P_MIN_CELLS_PER_CU * h->ncu) {
		StageArgs a = stage_args(h);
		a.cf_set = h->cf_set;
		a.fs_set = h->cur_set;
		a.store = 1;
		a.pure = 1;
		a.write_v = 0;
		a.phi_c = h->d_phiA;
		a.phi_out = h->d_phiB;
		a.gauss = nullptr;
		for(int form = 0; form < 2; ++form) {
			auto launch = [&]() -> int {
				if(form == 1) {
					HIPCHK(h, launch_stage_p(persist_args(h, a), h->ncells, h->ncu, h->stream));
					return RC_OK;
				}
				return launch_groups(h, MODE_FUSED, a);
			};
			if((rc = launch())) return rc;
			HIPCHK(h, hipEventRecord(e0, h->stream));
			for(int r = 0; r < 3; ++r)
				if((rc = launch())) return rc;
			HIPCHK(h, hipEventRecord(e1, h->stream));
			HIPCHK(h, hipEventSynchronize(e1));
			float ms = 0.f;
			HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
			(form == 1 ? h->persist_us : h->groups_us) = 1e3 * ms / 3;
		}
		h->persist = h->ncells >= P_MIN_CELLS_PER_CU * h->ncu && h->persist_us < h->groups_us;
	}
	h->timing = timing;
	h->graph_valid = false;
	return check_status(h);
}

extern "C" int rc_set_stage_shape(rc_handle *h, int threads, int rows) {
	if(!h) return RC_ERR_ARG;
	if(threads == 0 && rows == 0) {
		h->stage_forced[0] = h->stage_forced[1] = 0;
		h->stage_tuned = false;
		h->graph_valid = false;
		return RC_OK;
	}
	if(threads == P_THREADS) {
		if(!h->persist_ok || (rows != 0 && rows != h->p_npt))
			return fail(h, RC_ERR_ARG, "the persistent stage kernel does not fit this layout or its parameters");
		h->persist = true;
		h->stage_forced[0] = threads;
		h->stage_forced[1] = rows;
		h->stage_tuned = true;
		h->graph_valid = false;
		return RC_OK;
	}
	bool any = false;
	for(auto &g : h->groups)
		for(const auto &v : g.valid)
			if(v.first == threads && v.second == rows) {
				g.threads = threads;
				g.rows = rows;
				any = true;
			}
	if(!any) return fail(h, RC_ERR_ARG, "stage shape %d threads x %d rows fits no patch group", threads, rows);
	h->persist = false;
	h->stage_forced[0] = threads;
	h->stage_forced[1] = rows;
	h->stage_tuned = true;
	h->graph_valid = false;
	return RC_OK;
}

extern "C" int rc_stage_groups(rc_handle *h, int max_groups, int *ngroups, int (*shape)[3], int (*valid)[8][2],
		double (*tune_us)[8]) {
	if(!h || !ngroups) return RC_ERR_ARG;
	*ngroups = (int) h->groups.size();
	for(int k = 0; k < (int) h->groups.size() && k < max_groups; ++k) {
		const auto &g = h->groups[k];
		if(shape) {
			shape[k][0] = g.threads;
			shape[k][1] = g.rows;
			shape[k][2] = (int) g.cells.size();
		}
		for(int i = 0; i < 8; ++i) {
			const bool in = i < (int) g.valid.size();
			if(valid) {
				valid[k][i][0] = in ? g.valid[i].first : 0;
				valid[k][i][1] = in ? g.valid[i].second : 0;
			}
			if(tune_us) tune_us[k][i] = in ? g.tune_us[i] : 0.;
		}
	}
	return RC_OK;
}

extern "C" int rc_stage_persistent(rc_handle *h, int *eligible, int *used, double *tune_us_persistent,
		double *tune_us_groups) {
	if(!h) return RC_ERR_ARG;
	if(eligible) *eligible = h->persist_ok ? 1 : 0;
	if(used) *used = h->persist ? 1 : 0;
	if(tune_us_persistent) *tune_us_persistent = h->persist_us;
	if(tune_us_groups) *tune_us_groups = h->groups_us;
	return RC_OK;
}

// Dry steps as a HIP graph: GRAPH_STEPS steps (gather, predictor groups,
// gather, corrector groups each) captured once per layout, stage shapes and
// parameter set, replayed per window.  The window's noise is copied into
// d_ggauss and its first step index into d_step_base before each replay, so a
// growth request raised inside a replay reports its absolute step.  The
// first gather of the window always rebuilds the tile lists.
constexpr int GRAPH_STEPS = 16;

static int capture_graph(rc_handle *h, bool with_gauss) {
	if(h->graph_exec) {
		HIPCHK(h, hipGraphExecDestroy(h->graph_exec));
		h->graph_exec = nullptr;
	}
	h->graph_valid = false;
	if(!h->d_ggauss) HIPCHK(h, dalloc(&h->d_ggauss, (size_t) GRAPH_STEPS * std::max(h->ncells, 1)));
	if(!h->d_step_base) HIPCHK(h, dalloc(&h->d_step_base, 1));
	if(!h->h_step_base) HIPCHK(h, hipHostMalloc((void **) &h->h_step_base, sizeof(int) * 64, hipHostMallocDefault));
	h->lists_valid = false;
	h->capturing = true;
	hipGraph_t graph = nullptr;
	hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
	int rc = RC_OK;
	if(e == hipSuccess) {
		for(int s = 0; s < GRAPH_STEPS && rc == RC_OK; ++s)
			rc = step_dry(h, with_gauss ? h->d_ggauss + (size_t) s * h->ncells : nullptr, s);
		e = hipStreamEndCapture(h->stream, &graph);
	}
	h->capturing = false;
	if(rc) return rc;
	if(e != hipSuccess) return fail(h, RC_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
	e = hipGraphInstantiateWithFlags(&h->graph_exec, graph, 0);
	(void) hipGraphDestroy(graph);
	if(e != hipSuccess) return fail(h, RC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
	h->graph_valid = true;
	h->graph_gauss = with_gauss;
	h->graph_set = h->cur_set;
	return RC_OK;
}

extern "C" int rc_set_graphs(rc_handle *h, int enable) {
	if(!h) return RC_ERR_ARG;
	h->use_graphs = enable != 0;
	return RC_OK;
}

extern "C" int rc_graph_stats(rc_handle *h, long long *graph_launches, long long *eager_steps) {
	if(!h) return RC_ERR_ARG;
	if(graph_launches) *graph_launches = h->graph_launches;
	if(eager_steps) *eager_steps = h->eager_steps;
	return RC_OK;
}

extern "C" int rc_growth_stats(rc_handle *h, long long *growths, double *host_ms) {
	if(!h) return RC_ERR_ARG;
	if(growths) *growths = h->growths;
	if(host_ms) *host_ms = h->growth_ms;
	return RC_OK;
}

extern "C" int rc_step(rc_handle *h, int nsteps, const double *gauss) {
	if(!h) return RC_ERR_ARG;
	if(!h->initialized) return fail(h, RC_ERR_STATE, "rc_step before rc_init_forces");
	if(nsteps < 0) return fail(h, RC_ERR_ARG, "negative step count");
	if(!h->stage_tuned) {
		int rt = autotune_stage(h);
		if(rt) return rt;
	}
	const int chunk = 256;
	hipEvent_t t0 = nullptr, t1 = nullptr;
	if(h->timing) {
		t0 = ev_get(h);
		t1 = ev_get(h);
		HIPCHK(h, hipEventRecord(t0, h->stream));
	}
	int rc = RC_OK;
	for(int s0 = 0; s0 < nsteps && rc == RC_OK; s0 += chunk) {
		const int ns = std::min(chunk, nsteps - s0);
		if(gauss) {
			const long long need = (long long) ns * h->ncells;
			if(need > h->gauss_cap) {
				HIPCHK(h, hipStreamSynchronize(h->stream));
				dfree(h->d_gauss);
				h->d_gauss = nullptr;
				HIPCHK(h, dalloc(&h->d_gauss, need));
				h->gauss_cap = need;
			}
			HIPCHK(h, hipMemcpyAsync(h->d_gauss, gauss + (size_t) s0 * h->ncells, sizeof(double) * need,
					hipMemcpyHostToDevice, h->stream));
		}
		int s = 0;
		bool grown = false;
		while(s < ns) {
			bool halted = false;
			int window = 0;
			for(; s < ns; ++s) {
				// graph replay of a whole window (dry model, steady parameter set, no timing)
				// after a growth (new layout: the graph is captured again) only when
				// at least two windows remain, so that the capture pays
				if(h->use_graphs && !h->split && !h->timing && ns - s >= (grown ? 2 : 1) * GRAPH_STEPS && h->cf_set == h->cur_set &&
						window < 64) {
					if(!h->graph_valid || h->graph_gauss != (gauss != nullptr) || h->graph_set != h->cur_set) {
						HIPCHK(h, hipStreamSynchronize(h->stream));
						if((rc = capture_graph(h, gauss != nullptr))) return rc;
					}
					h->h_step_base[window] = s;
					HIPCHK(h, hipMemcpyAsync(h->d_step_base, h->h_step_base + window, sizeof(int), hipMemcpyHostToDevice,
							h->stream));
					if(gauss)
						HIPCHK(h, hipMemcpyAsync(h->d_ggauss, h->d_gauss + (size_t) s * h->ncells,
								sizeof(double) * GRAPH_STEPS * h->ncells, hipMemcpyDeviceToDevice, h->stream));
					HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
					h->graph_launches += 1;
					h->lists_valid = false;
					h->cf_set = h->cur_set;
					++window;
					s += GRAPH_STEPS - 1;
					continue;
				}
				const double *g = gauss ? h->d_gauss + (size_t) s * h->ncells : nullptr;
				rc = h->split ? step_split(h, g, s) : step_dry(h, g, s);
				h->eager_steps += 1;
				if(rc == -1) {
					halted = true;
					rc = RC_OK;
					break;
				}
				if(rc) return rc;
			}
			HIPCHK(h, hipMemcpyAsync(h->h_st, h->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, h->stream));
			HIPCHK(h, hipStreamSynchronize(h->stream));
			const DevStatus st = *h->h_st;
			if((st.flags & ST_KIND_MASK) == ST_GROWTH) {
				// the device halted after the corrector of step st.step
				{
					const auto g0 = std::chrono::steady_clock::now();
					if((rc = serve_growth(h))) return rc;
					h->growths += 1;
					h->growth_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
				}
				if(h->split) {
					if((rc = force_eval(h, h->d_phiA, false))) return rc;
				}
				else
					h->cf_set = h->cur_set;
				s = st.step + 1;
				grown = true;
				continue;
			}
			if(st.flags) break;
			(void) halted;
		}
		if(gauss && s0 + chunk < nsteps) HIPCHK(h, hipStreamSynchronize(h->stream));  // d_gauss is reused
	}
	if(h->timing) HIPCHK(h, hipEventRecord(t1, h->stream));
	rc = check_status(h);
	if(h->timing) {
		float ms = 0.f;
		(void) hipEventElapsedTime(&ms, t0, t1);
		h->t_ms[RC_T_STEP] += ms;
		h->t_n[RC_T_STEP] += 1;
		collect_timing(h);
	}
	return rc;
}

// the force evaluation of the current state without side effects (dry model)
static int materialize(rc_handle *h, bool site_outputs) {
	if(!h->initialized) return fail(h, RC_ERR_STATE, "no force evaluation before rc_init_forces");
	if(h->split && !site_outputs) return RC_OK;
	if(site_outputs) {
		for(auto &p : h->d_out)
			if(!p) HIPCHK(h, dalloc(&p, h->rows));
	}
	int rc;
	if((rc = run_gather(h, h->d_phiA))) return rc;
	StageArgs a = stage_args(h);
	a.cf_set = h->cf_set;
	a.fs_set = h->cf_set;
	a.phi_c = h->d_phiA;
	a.pure = 1;
	if(h->split) {
		// keep the evaluated velocity: divert the force rows to scratch
		// (mu, dx, dy and the Laplacian of the nematic model are the polar
		// model's with beta = 0, which this evaluation takes)
		a.frows = h->d_r;
		a.vrows = h->d_r;
	}
	if(site_outputs) {
		a.out_mu = h->d_out[0];
		a.out_dx = h->d_out[1];
		a.out_dy = h->d_out[2];
		a.out_lap = h->d_out[3];
	}
	if((rc = run_stage(h, MODE_CF, a))) return rc;
	return check_status(h);
}

extern "C" int rc_download_cells(rc_handle *h, rc_cell *cells, double *phi) {
	if(!h) return RC_ERR_ARG;
	if(!h->uploaded) return fail(h, RC_ERR_STATE, "no cells uploaded");
	int rc;
	if(h->initialized && !h->split)
		if((rc = materialize(h, false))) return rc;
	std::vector<rc_cell> c;
	std::vector<double> p;
	if((rc = download_raw(h, c, p, nullptr))) return rc;
	if(cells) std::copy(c.begin(), c.end(), cells);
	if(phi) std::copy(p.begin(), p.end(), phi);
	return RC_OK;
}

extern "C" int rc_get_rows(rc_handle *h, int which, double *out) {
	if(!h) return RC_ERR_ARG;
	if(!out || which < 0 || which > 3) return fail(h, RC_ERR_ARG, "bad rows request %d", which);
	if(which >= 2 && !h->wet) return fail(h, RC_ERR_ARG, "force rows exist only for the wet model");
	int rc;
	if(!h->split && (rc = materialize(h, false))) return rc;
	std::vector<double2> tmp(h->rows);
	const double2 *src = which < 2 ? h->d_v : h->d_f;
	HIPCHK(h, hipMemcpyAsync(tmp.data(), src, sizeof(double2) * h->rows, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	for(long long i = 0; i < h->rows; ++i) out[i] = (which & 1) ? tmp[i].y : tmp[i].x;
	return RC_OK;
}

extern "C" int rc_get_site_array(rc_handle *h, int which, double *out) {
	if(!h) return RC_ERR_ARG;
	if(!out || which < 0 || which > 4) return fail(h, RC_ERR_ARG, "bad site array %d", which);
	int rc;
	if(which == 0) {
		HIPCHK(h, hipMemcpyAsync(out, h->d_phiA, sizeof(double) * h->rows, hipMemcpyDeviceToHost, h->stream));
		HIPCHK(h, hipStreamSynchronize(h->stream));
		return RC_OK;
	}
	if((rc = materialize(h, true))) return rc;
	HIPCHK(h, hipMemcpyAsync(out, h->d_out[which - 1], sizeof(double) * h->rows, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	return RC_OK;
}

// VelocityField / ShapeField box fields of one box, the raw sums and their
// size_grid x size_grid coarse average (VelocityField.cpp:69-128,
// ShapeField.cpp:69-128), computed on the device from the current fields into
// work (f_x, f_y, coarse_x, coarse_y: four L x L arrays), then the requested
// outputs copied out of it
static int box_field_into(rc_handle *h, int box, int which, int size_grid, double *work, double *const outs[4]) {
	int rc;
	// the velocity rows of the current fields (the dry model does not keep them)
	if((rc = materialize(h, false))) return rc;
	if(!h->lists_valid) {
		if((rc = run_gather(h, h->d_phiA))) return rc;
		// a list overflow or an overlap raised by this gather leaves truncated lists
		if((rc = check_status(h))) return rc;
	}
	const DevBox &B = h->dboxes[box];
	const size_t n = (size_t) B.L * B.L;
	FieldArgs a{};
	a.boxes = h->d_boxes;
	a.tile_box = h->d_tile_box;
	a.tile_geo = h->d_tile_geo;
	a.tile_count = h->d_tile_count;
	a.tile_cells = h->d_tile_cells;
	a.cells = h->d_cells;
	a.phi = h->d_phiA;
	a.v = h->d_v;
	a.which = which;
	a.tile0 = B.tile_off;
	a.fx = work;
	a.fy = work + n;
	hipError_t e = launch_box_field(a, B.ntx * B.nty, h->stream);
	if(e == hipSuccess) e = launch_box_coarse(work, work + n, work + 2 * n, work + 3 * n, B.L, size_grid, h->stream);
	for(int i = 0; i < 4 && e == hipSuccess; ++i)
		if(outs[i]) e = hipMemcpyAsync(outs[i], work + i * n, sizeof(double) * n, hipMemcpyDefault, h->stream);
	if(e == hipSuccess) e = hipStreamSynchronize(h->stream);
	if(e != hipSuccess) return fail(h, RC_ERR_HIP, "box field: %s", hipGetErrorString(e));
	return RC_OK;
}

static int box_field_check(rc_handle *h, int box, int which, int size_grid) {
	if(box < 0 || box >= h->nboxes) return fail(h, RC_ERR_ARG, "box %d out of range", box);
	if(which != 0 && which != 1) return fail(h, RC_ERR_ARG, "unknown box field %d (0 velocity, 1 shape)", which);
	const int L = h->dboxes[box].L;
	if(size_grid < 1 || size_grid > L) return fail(h, RC_ERR_ARG, "size_grid %d outside 1..%d", size_grid, L);
	return RC_OK;
}

extern "C" int rc_box_field(rc_handle *h, int box, int which, int size_grid, double *fx, double *fy, double *cx,
		double *cy) {
	if(!h) return RC_ERR_ARG;
	int rc;
	if((rc = box_field_check(h, box, which, size_grid))) return rc;
	if(!h->d_bf) {
		size_t big = 0;
		for(const DevBox &o : h->dboxes) big = std::max(big, (size_t) o.L * o.L);
		HIPCHK(h, dalloc(&h->d_bf, 4 * big));
	}
	double *const outs[4] = { fx, fy, cx, cy };
	return box_field_into(h, box, which, size_grid, h->d_bf, outs);
}

// rc_box_field staged in this device's static array: no allocation on any path
extern "C" int rc_box_field_noalloc(rc_handle *h, int box, int which, int size_grid, double *fx, double *fy,
		double *cx, double *cy) {
	if(!h) return RC_ERR_ARG;
	int rc;
	if((rc = box_field_check(h, box, which, size_grid))) return rc;
	const int L = h->dboxes[box].L;
	if(L > RC_BOX_FIELD_STATIC_L)
		return fail(h, RC_ERR_ARG, "box side %d exceeds the static staging limit RC_BOX_FIELD_STATIC_L = %d", L,
				RC_BOX_FIELD_STATIC_L);
	if(!h->scratch) return fail(h, RC_ERR_ARG, "device %d has no static staging array (devices 0..%d)", h->device,
				SCRATCH_DEVICES - 1);
	double *const outs[4] = { fx, fy, cx, cy };
	return box_field_into(h, box, which, size_grid, h->scratch, outs);
}

extern "C" int rc_total_rows(rc_handle *h, long long *rows) {
	if(!h || !rows) return RC_ERR_ARG;
	*rows = h->rows;
	return RC_OK;
}

extern "C" int rc_cg_stats(rc_handle *h, rc_cg_stats_t *out) {
	if(!h || !out) return RC_ERR_ARG;
	fold_cg_ms(h);
	*out = h->cgs;
	return RC_OK;
}

extern "C" int rc_enable_timing(rc_handle *h, int on) {
	if(!h) return RC_ERR_ARG;
	h->timing = on != 0;
	for(int k = 0; k < RC_NKINDS; ++k) {
		h->t_ms[k] = 0;
		h->t_n[k] = 0;
	}
	return RC_OK;
}

extern "C" int rc_timing(rc_handle *h, int kind, double *ms, long long *launches) {
	if(!h || kind < 0 || kind >= RC_NKINDS) return RC_ERR_ARG;
	if(ms) *ms = h->t_ms[kind];
	if(launches) *launches = h->t_n[kind];
	return RC_OK;
}

extern "C" int rc_attach_net(rc_handle *h, rc_net *n, const float scale[3]) {
	if(!h) return RC_ERR_ARG;
	if(!h->wet || h->nboxes != 1) return fail(h, RC_ERR_ARG, "the learned warm start needs the wet model on one box");
	if(!n) {
		h->net = nullptr;
		if(h->warm_mode == RC_WARM_LEARNED) h->warm_mode = RC_WARM_PREVIOUS;
		return RC_OK;
	}
	int N, H, W;
	net_dims(n, &N, &H, &W);
	if(N != 1 || H != h->boxes[0].L || W != h->boxes[0].L)
		return fail(h, RC_ERR_ARG, "network shape %dx%dx%d does not match the %d^2 box", N, H, W, h->boxes[0].L);
	if(!h->d_xprev) {
		HIPCHK(h, dalloc(&h->d_xprev, std::max<long long>(h->rows, 1)));
		HIPCHK(h, dalloc(&h->d_rprev, std::max<long long>(h->rows, 1)));
	}
	h->net = n;
	for(int k = 0; k < 3; ++k) h->scale[k] = scale ? scale[k] : 1.f;
	return RC_OK;
}

extern "C" int rc_snapshot_box(rc_handle *h, float *out) {
	if(!h || !out) return RC_ERR_ARG;
	if(!h->wet || h->nboxes != 1) return fail(h, RC_ERR_ARG, "snapshots exist for the wet model on one box");
	if(!h->initialized) return fail(h, RC_ERR_STATE, "no force evaluation yet");
	const int ns = (int) h->box_sites;
	if(!h->d_snap) HIPCHK(h, dalloc(&h->d_snap, 5 * (size_t) ns));
	CGArgs c = cg_args(h);
	c.x = h->d_f;
	HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
	HIPCHK(h, launch_box_snapshot(c, h->d_phi2, h->d_snap, 0, nullptr, 1., ns, h->stream));
	HIPCHK(h, launch_box_snapshot(c, h->d_phi2, h->d_snap, 2, nullptr, 1., ns, h->stream));
	c.x = h->d_v;
	HIPCHK(h, launch_cg_gather(CG_GATHER_X, c, h->ntiles, h->stream));
	HIPCHK(h, launch_box_snapshot(c, h->d_phi2, h->d_snap, 3, nullptr, 1., ns, h->stream));
	HIPCHK(h, hipMemcpyAsync(out, h->d_snap, sizeof(float) * 5 * (size_t) ns, hipMemcpyDeviceToHost, h->stream));
	HIPCHK(h, hipStreamSynchronize(h->stream));
	return RC_OK;
}

extern "C" int rc_record_solves(rc_handle *h, int capacity) {
	if(!h || capacity < 0) return RC_ERR_ARG;
	if(!h->wet || h->nboxes != 1) return fail(h, RC_ERR_ARG, "recording needs the wet model on one box");
	dfree(h->d_rec);
	dfree(h->d_xrec);
	dfree(h->d_rnrec);
	h->d_rec = nullptr;
	h->d_xrec = nullptr;
	h->d_rnrec = nullptr;
	h->rec_cap = h->rec_n = h->rec_open = 0;
	if(capacity == 0) return RC_OK;
	HIPCHK(h, dalloc(&h->d_rec, (size_t) capacity * 5 * h->box_sites));
	HIPCHK(h, dalloc(&h->d_xrec, (size_t) RC_MAX_STAGES * std::max<long long>(h->rows, 1)));
	HIPCHK(h, dalloc(&h->d_rnrec, 2 * RC_MAX_STAGES));
	h->rec_cap = capacity;
	return RC_OK;
}

extern "C" int rc_set_learned_stages(rc_handle *h, int stages) {
	if(!h) return RC_ERR_ARG;
	if(stages < 1 || stages > RC_MAX_STAGES)
		return fail(h, RC_ERR_ARG, "learned stages %d outside 1..%d", stages, RC_MAX_STAGES);
	h->stages = stages;
	return RC_OK;
}

extern "C" int rc_get_recorded(rc_handle *h, float *out, int *count) {
	if(!h || !count) return RC_ERR_ARG;
	*count = h->rec_n;
	if(out && h->rec_n)
		HIPCHK(h, hipMemcpy(out, h->d_rec, sizeof(float) * (size_t) h->rec_n * 5 * h->box_sites, hipMemcpyDeviceToHost));
	return RC_OK;
}
