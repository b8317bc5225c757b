/*
 * rheocell.h -- C ABI of the MI355X (gfx950) RheoCell per-step hot path.
 *
 * The library owns device buffers behind an opaque handle.  Host callers
 * (the C++ FD_HIPBackend in rheocell_amd/host, or ctypes) keep their own
 * BaseField-like host copies and refresh them through the sync calls.
 *
 * Reference interfaces replaced (files of deppinto/RheoCell):
 *   rc_create            SimBackend::init + WetPolarModel::init/set_box
 *                        (src/Backends/SimBackend.cpp:171-218,
 *                         src/Interactions/WetPolarModel.cpp:49-57,95-106)
 *   rc_upload_cells      SimBackend::read_next_configuration per-cell state
 *                        (src/Backends/SimBackend.cpp:268-299) and
 *                        SimBackend::apply_changes_to_simulation_data
 *                        (src/Backends/SimBackend.h:148)
 *   rc_init_forces       FD_CPUBackend::init force evaluation
 *                        (src/Backends/FD_CPUBackend.cpp:19-25)
 *   rc_apply_changes_after_equilibration
 *                        WetPolarModel::apply_changes_after_equilibration
 *                        (src/Interactions/WetPolarModel.cpp:84-93)
 *   rc_step              SimBackend::sim_step = FD_CPUBackend::sim_step
 *                        (src/Backends/SimBackend.h:133,
 *                         src/Backends/FD_CPUBackend.cpp:134-160)
 *   rc_download_cells    SimBackend::apply_simulation_data_changes
 *                        (src/Backends/SimBackend.h:138) for Configuration
 *                        (src/Observables/Configuration.cpp:46-55)
 *   rc_get_rows          WetPolarModel::get_velocity_x/y and the force rows
 *                        (src/Interactions/WetPolarModel.h:101-106)
 *   rc_get_site_array    BaseField public site vectors (src/Fields/BaseField.h:188-195)
 *   rc_cg_stats          ConjugateGradient::iterations()/error() of the wet solve
 *                        (src/Interactions/WetPolarModel.cpp:220-233)
 *   rc_set_warm_start    warm-start rule of the wet solve
 *                        (src/Interactions/WetPolarModel.cpp:152-157,223-232)
 *   rc_last_error        RCexception message (src/Utilities/RCexception.h)
 *   rc_destroy           SimBackend::~SimBackend (src/Backends/SimBackend.cpp:37-40)
 *
 * Every call returns RC_OK (0) or an RC_ERR_* status; rc_last_error() gives
 * the message.  A handle is not re-entrant: one host thread at a time.
 */
#ifndef RHEOCELL_H_
#define RHEOCELL_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RC_OK 0
#define RC_ERR_ARG 1          /* invalid argument / unsupported option */
#define RC_ERR_HIP 2          /* HIP runtime error */
#define RC_ERR_OVERLAP 3      /* > 20 patches cover one box site (WetPolarModel.cpp:250-255) */
#define RC_ERR_GROWTH 4       /* a patch must grow (MultiPhaseField.cpp:248-310) */
#define RC_ERR_CG 5           /* reserved: a solve that reaches its iteration cap is reported in
                                 rc_cg_stats_t.last_cold (bit 1) and the step goes on, as the
                                 reference's unchecked Eigen solve does (WetPolarModel.cpp:223-232) */
#define RC_ERR_STATE 6        /* call out of order (e.g. step before upload) */
#define RC_ERR_TILE 7         /* too many patches intersect one box tile */
#define RC_ERR_NAN 8          /* a cell's field or the wet solve's residual became NaN (no counterpart:
                                 the reference carries NaN on); the device halts at the first one */

/* Model parameters of one box, as read from the input file: the *_active
 * values of WetPolarModel::get_settings (WetPolarModel.cpp:28-46) plus the
 * backend keys J0 (FD_CPUBackend.cpp:15) and dt (FDBackend.cpp:34). */
typedef struct rc_model_params {
	double gamma, lambda, omega, mu, kappa;
	int R;
	int anchoring;
	double zetaQ_self, zetaQ_inter, J_Q;
	double friction, friction_cell, CGtolerance, wall_slip, passive_alpha, beta;
	double J0, dt;
	int cg_max_iter;      /* <= 0: Eigen's default of 2 * rows */
	int interaction;      /* RC_INTERACTION_WETPOLAR (interaction_type = wetpolarmodel) or
	                         RC_INTERACTION_WETNEMATIC (interaction_type = wetmodel,
	                         src/Interactions/WetModel.cpp; beta is not one of its keys and
	                         is ignored); every box of a handle has the same one */
} rc_model_params;

#define RC_INTERACTION_WETPOLAR 0
#define RC_INTERACTION_WETNEMATIC 1

/* An external force acting on every cell (ForceFactory::add_force,
 * src/Forces/ForceFactory.cpp:31-46, with field = -1).  kind:
 *   RC_FORCE_CHANNEL_WALLS  types channel_walls, square_walls, circle_walls,
 *                           triangle_walls, moving_walls (ChannelWalls.cpp):
 *                           free energy 2 kappa_wall / lambda phi w^2
 *   RC_FORCE_SHEAR_FLOW     shear_flow_channel (ShearFlowChannel.cpp): free energy
 *                           4 kappa_wall / lambda phi w^2, density 2 kappa_wall /
 *                           lambda phi^2 w^2, a linear shear profile between the walls
 *   RC_FORCE_POIS_FLOW      pois_flow_channel (PoiseuilleFlowChannel.cpp): free
 *                           energy as the shear flow, a parabolic profile
 *   RC_FORCE_STRETCHING     stretching (Stretching.cpp): the cells of index 0 and
 *                           9 mod 10 pulled apart along x
 * w is the box's wall field; shear_rate acts after equilibration
 * (rc_apply_changes_after_equilibration), 0 before. */
#define RC_FORCE_CHANNEL_WALLS 0
#define RC_FORCE_SHEAR_FLOW 1
#define RC_FORCE_POIS_FLOW 2
#define RC_FORCE_STRETCHING 3
#define RC_MAX_FORCES 4
typedef struct rc_ext_force {
	int kind;
	double kappa_wall, lambda, shear_rate;
} rc_ext_force;

/* One square box (SquareBox, src/Boxes/SquareBox.cpp:21-29) with its own
 * cells and parameters.  Several boxes = independent replicas advanced
 * together.  walls: the getWalls field of a wall box of the reference
 * (SquareWalls / CircleWalls / TriangleWalls, L*L values, site x + L y),
 * read by rc_create, or NULL for the periodic box; sites with walls >=
 * wall_slip mask the stencils (BaseInteraction.h:196-263) and carry no force.
 * Walls are supported on the dry path (friction_cell = 0); external forces
 * on both, without J_Q. */
typedef struct rc_box_desc {
	int L;                /* box side */
	int ncells;
	rc_model_params params;
	const double *walls;
	int nforces;          /* 0 .. RC_MAX_FORCES, in the external_forces_file's order */
	rc_ext_force forces[RC_MAX_FORCES];
} rc_box_desc;

/* Per-cell scalar state (public members of BaseField, src/Fields/BaseField.h:122-225). */
typedef struct rc_cell {
	int LsubX, LsubY;
	int offset[2];
	int corner, corner_old;          /* sub_corner_bottom_left(_old) */
	int x_sub_left, y_sub_bottom;
	double CoM[2], CoM_old[2];
	double unrap[2];                 /* unrap_sub_corner_bottom_left_x/y */
	double Q00, Q01, thetaQ, thetaQ_old;
	double area, area_old, sumF, sumF_old, perimeter, S00, S01;
	double total_force_repulsion[2];
	double nemQ[2], nemQ_old[2];     /* nematic director and its stored value (WetModel); the
	                                    reference sets them in MultiPhaseField::init from the
	                                    cell's drand48 draw (MultiPhaseField.cpp:79-95) */
} rc_cell;

typedef struct rc_cg_stats_t {
	long long solves;                /* one per component */
	long long iters_total;           /* SpMV iterations summed over solves */
	int last_iters[2];               /* x, y of the last solve */
	double last_relres[2];
	int last_cold;
	double last_solve_ms;            /* wall time of the last x+y solve (HIP events) */
	double total_solve_ms;
	long long sweeps_total;          /* launches of the iteration kernels that did work:
	                                    sum over solves of max(iters_x, iters_y) */
	long long learned_used;          /* solves that kept the network's start */
	long long learned_tried;         /* solves that ran the network stages (the rest fell
	                                    back to the previous solution in some component) */
	double worst_relres;             /* largest final |r| / |b| of any component of any solve */
} rc_cg_stats_t;

#define RC_WARM_REFERENCE 0   /* previous solution, cold restart every 11th solve (reference rule) */
#define RC_WARM_PREVIOUS 1    /* previous solution, never a periodic cold restart */
#define RC_WARM_LEARNED 2     /* previous solution refined by the attached network
                                 (needs rc_attach_net; see rc_set_learned_stages) */

typedef struct rc_handle rc_handle;

int rc_create(const rc_box_desc *boxes, int nboxes, rc_handle **out);
/* cells: sum(ncells) entries, box-major; phi: the patches concatenated in
 * cell order, storage order inside a patch (rows = field_start_index order). */
int rc_upload_cells(rc_handle *h, const rc_cell *cells, const double *phi);
int rc_init_forces(rc_handle *h);
int rc_apply_changes_after_equilibration(rc_handle *h);
/* gauss: nsteps * total_cells standard normal draws, step-major, cell order
 * (the Utils::gaussian() calls of the predictor, WetPolarModel.cpp:365).
 * NULL means zero noise. */
int rc_step(rc_handle *h, int nsteps, const double *gauss);
int rc_download_cells(rc_handle *h, rc_cell *cells, double *phi);
/* which: 0 vx, 1 vy, 2 fx, 3 fy -- rows in field_start_index order */
int rc_get_rows(rc_handle *h, int which, double *out);
/* which: 0 phi, 1 freeEnergy, 2 fieldDX, 3 fieldDY, 4 laplacianPhi -- all cells concatenated */
int rc_get_site_array(rc_handle *h, int which, double *out);
int rc_total_rows(rc_handle *h, long long *rows);
/* Box fields of the observables VelocityField (which 0; VelocityField.cpp:69-128)
 * and ShapeField (which 1; ShapeField.cpp:69-128) for box `box`, on the device
 * from the current fields: f(k) = sum over the patches covering box site k (cell
 * order) of w * phi, w = the velocity row (vx, vy) or the cell's (S00, S01);
 * c = the size_grid x size_grid periodic average of f (f_field_coarse).  Each
 * output is L*L doubles, site x + L y; any may be NULL. */
int rc_box_field(rc_handle *h, int box, int which, int size_grid, double *fx, double *fy, double *cx, double *cy);
/* rc_box_field without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The four
 * L x L staging fields live in a static device array of the library, one per
 * GPU, held from the library's load to the end of the process
 * (4 * RC_BOX_FIELD_STATIC_L^2 doubles = 8 MiB); boxes with a side above
 * RC_BOX_FIELD_STATIC_L are refused with RC_ERR_ARG.  Every output that is
 * not NULL receives L*L doubles and may be a device pointer or a host
 * pointer (pinned host memory keeps the copy free of runtime staging).  Every
 * handle on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (several host threads, or handles
 * on different streams) race.  Results equal rc_box_field's bit for bit. */
#define RC_BOX_FIELD_STATIC_L 512
int rc_box_field_noalloc(rc_handle *h, int box, int which, int size_grid, double *fx, double *fy, double *cx,
		double *cy);
int rc_cg_stats(rc_handle *h, rc_cg_stats_t *out);
int rc_set_warm_start(rc_handle *h, int mode);
/* Form of the wet solve's Jacobi-PCG (no counterpart in the reference, which
 * runs Eigen's ConjugateGradient on the assembled rows, WetPolarModel.cpp:200-233).
 * Both forms run Eigen's iteration from the same start with the same stop
 * test; they differ in rounding only.  RC_CG_SITES (default): the iterates
 * are carried as scalar multiples of the first residual's part that sums to
 * zero at every box site plus L x L box fields, since every row's diagonal
 * depends on its box site only; an iteration moves O(L^2) data.
 * RC_CG_ROWS: every iterate as rows; an iteration moves O(rows) data.  The
 * learned start's refinement stages and rc_record_solves run on rows in
 * either form. */
#define RC_CG_SITES 0
#define RC_CG_ROWS 1
int rc_set_cg_form(rc_handle *h, int form);
/* Kernel timing with HIP events on the library's stream (enabled by
 * rc_enable_timing).  Single-kernel kinds record their events at kernel
 * start and end (hipExtLaunchKernel), so they time the kernel alone;
 * RC_T_CG_SOLVE and RC_T_STEP bracket a whole group of launches.  rc_timing
 * returns the summed device time (ms) and launch count of one kind since
 * timing was enabled. */
#define RC_T_BOX_GATHER 0          /* k_box_gather: box sums of phi^2 */
#define RC_T_STAGE_FUSED_PRED 1    /* dry: force evaluation + predictor update */
#define RC_T_STAGE_FUSED_CORR 2    /* dry: force evaluation + corrector update */
#define RC_T_STAGE_CF 3            /* force evaluation only (wet RHS / dry velocity) */
#define RC_T_STAGE_FS_PRED 4       /* wet: predictor field update */
#define RC_T_STAGE_FS_CORR 5       /* wet: corrector field update */
#define RC_T_CG_DIAG 6
#define RC_T_CG_GATHER_X 7
#define RC_T_CG_GATHER_P 8         /* p = z + beta p fused with the box gather of p */
#define RC_T_CG_ROW_RESID 9
#define RC_T_CG_BOX_PAP 10         /* box pass: N4 of p and p.Ap -> alpha */
#define RC_T_CG_ROW_UPDATE 11      /* x += alpha p, r -= alpha Ap, r.r, r.z */
#define RC_T_CG_BOX_N4 12          /* box pass: N4 of the initial guess */
#define RC_T_CG_SOLVE 13           /* a whole x+y solve */
#define RC_T_STEP 14               /* a whole rc_step call */
#define RC_T_CG_GATHER_P_FIRST 15  /* p = z fused with the box gather of p */
#define RC_T_CG_REDUCE 16          /* one-workgroup reduction of CG partial sums */
#define RC_T_CG_LEARNED 17         /* the learned refinement stages of one solve (group) */
#define RC_T_NEM_SUMS 19           /* (unused: the nematic box sums of phi Q00, phi Q01 come from k_box_gather) */
#define RC_T_NEM_FORCE 18          /* k_nem_force: the nematic model's force evaluation */
#define RC_T_SCG_GATHER 20         /* site form: box sums of b and of the start x0 */
#define RC_T_SCG_INIT 21           /* site form: per-site set-up (site means, first residual's site part) */
#define RC_T_SCG_ROW_INIT 22       /* site form: ker-P residual sums over the rows */
#define RC_T_SCG_ITER 23           /* site form: one iteration (one pass over the box sites) */
#define RC_T_SCG_FINAL 24          /* site form: the solution rows */
#define RC_T_SCG_REDUCE 25         /* site form: set-up reduction and the stop test after a batch */
#define RC_T_SCG_INIT_W 26         /* site form: set-up, the first residual's A z and the sums */
#define RC_NKINDS 27
int rc_timing(rc_handle *h, int kind, double *ms, long long *launches);
int rc_enable_timing(rc_handle *h, int on);
const char *rc_last_error(rc_handle *h);
void rc_destroy(rc_handle *h);

/* ---- Stage kernel shape (no counterpart in the reference) ---------------
 * The per-cell field-update kernel covers a patch with `threads` threads per
 * cell (256 or 512), each thread taking one patch column and a strip of up to
 * `rows` rows.  Cells are grouped by the smallest shape that fits their patch
 * (uneven patch sizes: one launch per group, each with its own LDS size).
 * Before the first rc_step after an upload or a patch growth the library
 * times every shape that fits a group on the current fields (a
 * side-effect-free force evaluation) and keeps the fastest, unless
 * rc_set_stage_shape forced one.  Results agree between shapes to rounding of
 * the per-cell sums. */
/* force a shape on every group it fits (an error if it fits none); 0, 0: autotune */
int rc_set_stage_shape(rc_handle *h, int threads, int rows);
/* the groups (up to max_groups): shape[g] = {threads, rows, cells}; valid[g] =
 * the shapes that fit it ({0, 0} padded); tune_us[g] = the tuner's mean time per
 * valid shape in microseconds (0 before tuning).  Any output may be null. */
int rc_stage_groups(rc_handle *h, int max_groups, int *ngroups, int (*shape)[3], int (*valid)[8][2],
		double (*tune_us)[8]);
/* The dry fused stage (no J_Q, omega, beta; periodic boxes; patch sides of at
 * most 3 strip rows at 1024 threads) can instead run as one persistent launch
 * over every cell of every box, one 1024-thread workgroup per CU with the
 * next cell's loads in flight.  The tuner times it against the groups and
 * keeps the faster; rc_set_stage_shape(h, 1024, 0) forces it, any other shape
 * the groups.  Reports whether the layout allows it, whether it is used, and
 * the tuner's times of both (predictor stage, us; 0 if not timed). */
int rc_stage_persistent(rc_handle *h, int *eligible, int *used, double *tune_us_persistent, double *tune_us_groups);

/* ---- HIP graphs (no counterpart in the reference) -----------------------
 * Dry-model steps are replayed from a HIP graph of 16 steps (captured once per
 * layout, stage shapes and parameter set); remainders, timing runs and the
 * steps after a patch growth run as ordinary launches.  Results are identical
 * either way.  Default: enabled. */
int rc_set_graphs(rc_handle *h, int enable);
int rc_graph_stats(rc_handle *h, long long *graph_launches, long long *eager_steps);
/* patch growths served inside rc_step (MultiPhaseField.cpp:248-310: geometry
 * and re-layout on the host, the re-cut of the fields on the device) and the
 * host time they took */
int rc_growth_stats(rc_handle *h, long long *growths, double *host_ms);

/* ---- Learned CG warm start (no counterpart in the reference) ----------
 * A 6-layer 3x3 convolutional network (32 channels, circular padding, ReLU,
 * BF16 weights and activations, FP32 accumulation on MFMA) maps box fields --
 * the residual r = f - A x of a start x (summed over the patches covering a
 * site, divided per component by sqrt(|r_c|^2 / rows)) and sum phi^2 -- to
 * the box sums of the correction e = A^-1 r in the same units.  Rows follow
 * from e_i = (r_i + eta N4(k_i)) / D(k_i); x moves along e by the A-norm
 * optimal step (e.r)/(e.A e), which never raises the A-norm error.  A solve
 * starts from the previous solution and applies the stage rc_set_learned_stages
 * times; a component whose residual ends larger than the previous solution's
 * falls back to the previous solution.  The reference's CG then runs to its
 * tolerance from the result.
 * Tensors are host float arrays, pixel-major: x [N][H][W][3], y [N][H][W][2].
 * Weights: w [6][9][32][32] (layer, tap = 3*(dy+1)+(dx+1), out ch, in ch),
 * b [6][32]; entries of absent channels (inputs >= 3 of layer 0, outputs >= 2
 * of layer 5) are kept at zero. */
typedef struct rc_net rc_net;
int rc_net_create(int N, int H, int W, rc_net **out);
int rc_net_set_params(rc_net *n, const float *w, const float *b);
int rc_net_get_params(rc_net *n, float *w, float *b);
int rc_net_forward(rc_net *n, const float *x, float *out);
/* forward + MSE loss + backward, no update (gradient checks) */
int rc_net_grads(rc_net *n, const float *x, const float *y, float *gw, float *gb, double *loss);
/* forward + backward + one Adam step (beta 0.9 / 0.999) on the GPU */
int rc_net_train_step(rc_net *n, const float *x, const float *y, float lr, double *loss);
const char *rc_net_last_error(rc_net *n);
void rc_net_destroy(rc_net *n);
/* Use net (created with N = 1, H = W = box side) for RC_WARM_LEARNED.
 * scale[0]: multiplier of the scaled residual sums, scale[1]: of sum phi^2,
 * scale[2]: multiplier of the network output (-> scaled correction sums). */
int rc_attach_net(rc_handle *h, rc_net *n, const float scale[3]);
#define RC_MAX_STAGES 8
/* network refinement stages per solve, 1..RC_MAX_STAGES (default 1) */
int rc_set_learned_stages(rc_handle *h, int stages);
/* Box fields of the last force evaluation, [L*L][5]: sum fx, sum fy,
 * sum phi^2, sum vx, sum vy (wet model, one box). */
int rc_snapshot_box(rc_handle *h, float *out);
/* Record training data for the network from every following wet solve, up to
 * capacity records: one per start the solve passes through (the previous
 * solution, then the result of each network stage but the last).
 * rc_get_recorded copies [count][L*L][5]: box sums of the start's residual
 * (x, y), sum phi^2, box sums of the correction to the converged solution
 * (x, y), both divided per component by sqrt(|r_c|^2 / rows) of that start. */
int rc_record_solves(rc_handle *h, int capacity);
int rc_get_recorded(rc_handle *h, float *out, int *count);

#ifdef __cplusplus
}
#endif
#endif
