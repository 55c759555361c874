/*
 * dim2d_hip.h -- C ABI of the MI355X (gfx950) DIM time-integration library.
 *
 * The reference solver (Fortran, jdesmarais/dim2D_finite_volume) binds these
 * entry points through ISO_C_BINDING in place of its td_operators /
 * td_integrator / bc_operators modules for the serial hot path.  Every entry
 * point returns 0 on success and a negative DIM2D_E* code otherwise; the
 * message of the last error on the calling thread is available from
 * dim2d_last_error().  The reference prints and stops on errors
 * (errors_module.f:69,106); its Fortran shim keeps that convention by
 * stopping on a non-zero status.
 *
 * Arrays exchanged with the host have the reference's layout nodes(nx,ny,ne):
 * Fortran column-major, i (x) fastest, then j (y), then the variable
 * k = 1 mass density, 2 momentum_x, 3 momentum_y, 4 total energy; nx and ny
 * include the bc_size = 2 ghost layers on each side.
 *
 * One context owns one simulation's device buffers and one HIP stream; calls on
 * one context must not overlap (one host thread per context).
 */
#ifndef DIM2D_HIP_H
#define DIM2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DIM2D_OK 0
#define DIM2D_EINVAL (-1)   /* invalid argument or unsupported configuration */
#define DIM2D_EDEVICE (-2)  /* HIP runtime error */
#define DIM2D_ENOMEM (-3)   /* device allocation failed */
#define DIM2D_ESTATE (-4)   /* call out of order (e.g. step before upload) */
#define DIM2D_ENONFINITE (-5) /* non-finite values found by a diagnostic */

/* boundary kinds per edge (bc_operators_gen_class.f dispatch choices) */
#define DIM2D_BC_REFLECTION 0 /* reflection_x / reflection_y */
#define DIM2D_BC_PERIODIC 1   /* periodic_xy */
#define DIM2D_BC_WALL 2       /* wall_xy (equilibrium wall, constant heat flux) */
#define DIM2D_BC_OPEN 3       /* hedstrom_xy: characteristic open boundary, all four edges at
                                 once; the ghost frame is integrated with its own time
                                 derivative (bc_operators_hedstrom_xy_class.f:182-496) */

/* time integrators */
#define DIM2D_INT_RK3TVD 0 /* the reference's RK3-TVD (rk3tvd_steps_module.f) */
#define DIM2D_INT_LSRK3 1  /* low-storage RK3, Williamson (1980) 2N-storage; no counterpart
                              in the reference (its only integrator is RK3-TVD); single
                              contexts only; FP64: three FP64 state-sized buffers instead of
                              four; mixed precision: two FP64 states (stage states stay
                              FP64) and the register q in FP32 */

/* interior face stencils (the reference's compile-time sd_choice,
 * config/code_config/makefile_sd.mk) */
#define DIM2D_OPS_MATTSSON 0      /* mattsson_operators (sd_operators_class.f:185-1039) */
#define DIM2D_OPS_COCKBURN_GAU 1  /* cg_operators (cg_operators/sd_operators_class.f): FP64
                                     or mixed precision, RK3-TVD, reflection and periodic
                                     edges, single contexts */

/* arithmetic of the stage kernels */
#define DIM2D_PREC_FP64 0  /* everything in FP64 */
#define DIM2D_PREC_MIXED 1 /* FP64 u^n and RK accumulation, FP32 stage increments and stencils */

typedef struct dim2d_params {
  int32_t nx, ny;          /* grid points including ghosts (parameters_input.f nx, ny) */
  int32_t bc_size;         /* must be 2 (Mattsson operators, sd_operators_class.f get_bc_size) */
  int32_t bc_w, bc_e, bc_s, bc_n; /* DIM2D_BC_* per edge */
  int32_t precision;       /* DIM2D_PREC_* */
  int32_t gravity_ac;      /* body force switch (parameters_input.f gravity_ac) */
  double x_min, x_max, y_min, y_max; /* coordinates of the first/last interior points */
  double re, we, pr, viscous_r, cv_r; /* dim2d_parameters.f */
  double gravity;          /* gravity_amp */
  double wall_contact_angle; /* degrees (wall_micro_contact_angle) */
  double wall_heat_flux;   /* constant wall heat flux, 0 = no_heat_source */
  int32_t integrator;      /* DIM2D_INT_* (ti_choice of the reference's makefile_ti.mk) */
  int32_t operators;       /* DIM2D_OPS_* (sd_choice); 0 = Mattsson */
} dim2d_params;

typedef struct dim2d_ctx dim2d_ctx;

/* Text of the last error on the calling thread ("" if none). */
const char* dim2d_last_error(void);

/* Library version string. */
const char* dim2d_version(void);

/* Create a context for one simulation on the current HIP device.
 * Replaces the set-up half of field_abstract%ini (field_abstract_class.f:148-198):
 * ini_coordinates (:286-315) and ini_for_timeInt (:244-270). */
int dim2d_ctx_create(const dim2d_params* params, dim2d_ctx** out);

/* Release a context and its device buffers. */
int dim2d_ctx_destroy(dim2d_ctx* ctx);

/* Copy the host field nodes(nx,ny,4) into the device state u^n (blocking).
 * Replaces the host ownership of field_abstract%nodes (field_abstract_class.f:87). */
int dim2d_upload(dim2d_ctx* ctx, const double* nodes);

/* Copy the device state back to nodes(nx,ny,4), ghosts filled as after
 * bc_operators_gen%apply_bc_on_nodes (blocking). */
int dim2d_download(dim2d_ctx* ctx, double* nodes);

/* Same two calls on device memory owned by the caller (stream-ordered on the
 * context's stream, no host synchronisation). */
int dim2d_upload_device(dim2d_ctx* ctx, const double* d_nodes);
int dim2d_download_device(dim2d_ctx* ctx, double* d_nodes);

/* time_dev(nx,ny,4) of the current state, interior cells only; ghost cells are
 * set to the reference's debug_real (NF90_FILL_DOUBLE).  Replaces
 * td_operators%compute_time_dev (td_operators/finitevolume/td_operators_class.f:119-235)
 * including bc_operators_gen%apply_bc_on_fluxes for wall edges (:171-177). */
int dim2d_time_dev(dim2d_ctx* ctx, double t, double* time_dev);

/* dim2d_time_dev without a context and without allocating: the time
 * derivative of the planar device array d_nodes(nx,ny,4) into the planar
 * device array d_time_dev(nx,ny,4) (ghost cells set to debug_real unless the
 * edges are open), FP64 whatever params->precision says, stream-ordered on
 * `stream` (a hipStream_t; NULL for the default stream) with no host
 * synchronisation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory, such as code under HIP graph capture and
 * tight serving loops.  Its working memory is a 64 MiB device array each GPU
 * holds for the life of the process, enough for about 1 million cells of
 * pitch x ny (pitch = nx rounded up to 32), e.g. 1020 x 1020 points;
 * larger grids return DIM2D_EINVAL.  Calls on one GPU share that array:
 * calls that overlap on one GPU (other streams, other host threads) race. */
int dim2d_time_dev_noalloc(const dim2d_params* params, const double* d_nodes, double* d_time_dev,
                           void* stream);

/* Fill the ghost cells of the device state (bc_operators_gen%apply_bc_on_nodes,
 * bc_operators_gen_class.f:238-381).  Reflection and periodic ghosts are also
 * resolved on the fly inside the stage kernels. */
int dim2d_apply_bc(dim2d_ctx* ctx, double t);

/* One step of the context's integrator with time step dt: three fused
 * stages, each computing the time derivative, the stage update and the wall
 * ghost cells.  Replaces td_integrator%integrate
 * (td_integrator/rungekutta3rdtvd/td_integrator_class.f:103-161) with
 * compute_{1st,2nd,3rd}_step (rk3tvd_steps_module.f:52-100,205-269,424-490)
 * for DIM2D_INT_RK3TVD. */
int dim2d_rk3_step(dim2d_ctx* ctx, double t, double dt);

/* nsteps steps of the context's integrator without host synchronisation
 * (steps are replayed from captured HIP graphs of 8 steps, then of 2 for the
 * remainder, one graph per step parity, unless profiling is on; with open
 * edges the frame kernels are branches of the graph beside the stage
 * kernels).
 * cfl > 0: dt of each step from the CFL rule below, evaluated on the device
 *          (fused into the third stage of the previous step); dt_fixed ignored.
 *          In DIM2D_PREC_MIXED that evaluation takes c and |u| in FP32 after
 *          an FP64 temperature, so its dt is within ~1e-7 relative of the rule.
 * cfl <= 0: every step uses dt_fixed.
 * Replaces the time loop of sim_dim2d.f:81-92 for the steps between outputs. */
int dim2d_integrate(dim2d_ctx* ctx, double t0, double dt_fixed, double cfl, int32_t nsteps);

/* Time step rule (no counterpart in the reference, whose dt is a compile-time
 * constant, parameters_input.f:63):
 *   dt = cfl * min( h / max(|u|+c), Re*min(rho)*h^2, sqrt(We/max(rho))*h^2 ),
 * h = min(dx,dy), c the speed of sound (dim2d_prim_module.f:1179-1215) in the
 * reduced form c^2 = 3 (e + (P + 3 rho^2)/cv) / (rho (3 - rho)), e = P + rho^2 (2 rho - 3),
 * with c = 0 where c^2 <= 0 (spinodal part of an interface), over interior cells.  out[0]=dt, out[1]=max(|u|+c), out[2]=min rho, out[3]=max rho. */
int dim2d_cfl_dt(dim2d_ctx* ctx, double cfl, double out[4]);

/* Interior sums and extrema of the current state:
 * out[0..3] = sum of rho, rho*u, rho*v, E; out[4] = min rho; out[5] = max rho;
 * out[6] = number of non-finite values.  (No counterpart in the reference.) */
int dim2d_diagnostics(dim2d_ctx* ctx, double out[7]);

/* field%check_steady_state (field_abstract_class.f:629-667), used by the time
 * loop of steady-state runs (sim_dim2d.f:75-101, steady_state_limit = 1e-12 in
 * parameters_input.f:913) (blocking).  Computes the time derivative of the
 * current state and its max |.| per variable over the cells that carry one:
 * the interior, and the ghost frame when the edges are open.  out[0..3] (may
 * be NULL) receives the maxima, non-finite values counting as +inf;
 * *steady = 1 when every maximum is <= limit.  The reference scans every
 * cell of time_dev, whose ghost cells hold debug_real (td_operators_class.f:
 * 150-154) unless the edges are open, and it passes NaN (abs(NaN) > limit is
 * false); this checks the cells with a derivative and fails on NaN. */
int dim2d_check_steady_state(dim2d_ctx* ctx, double limit, double out[4], int32_t* steady);

/* Simulated time and the dt of the last step after dim2d_integrate
 * (blocking): out[0] = t, out[1] = dt. */
int dim2d_time(dim2d_ctx* ctx, double out[2]);

/* Block until the context's stream is idle. */
int dim2d_synchronize(dim2d_ctx* ctx);

/* Kernel tile selection: lanes per workgroup along x (64: one wave per strip
 * of 60 output columns; 128: two neighbouring strips per workgroup) and rows
 * per workgroup along y; 0 keeps the current value.  The defaults depend on
 * the grid: 16 rows from about 1500^2 interior points up, fewer on smaller
 * grids (8, 4, 2), which would leave most of the GPU idle with 16-row tiles;
 * two strips per workgroup on large grids only.  Results do not depend on
 * the tiling. */
int dim2d_set_tiles(dim2d_ctx* ctx, int32_t block_x, int32_t rows_per_block);

/* Stage fusion (no counterpart in the reference, whose stages each write a
 * full field, rk3tvd_steps_module.f:52-490): fuse = 1 computes stages 1 and 2
 * in one launch whose stage-1 state stays on chip (RK3-TVD: 64 fewer HBM
 * bytes per cell and step in FP64, 32 in mixed precision; low-storage RK3:
 * 128 fewer in FP64, 96 in mixed precision, u1 and q1 on chip); fuse = 0
 * launches the stages apart.
 * rows_per_block: tile height of the fused launch (0 keeps the current
 * value, default from the grid size).  On by default where it applies: the
 * Mattsson operators, reflection, periodic, S/N wall or (RK3-TVD only) open
 * edges, at least 8 interior points per axis; DIM2D_EINVAL otherwise.  With
 * open edges the ghost frame's stage 1 runs before the fused launch and its
 * stage 2 after it.  Results equal those of the separate launches to
 * rounding (the launch shapes contract products into FMAs differently; the
 * low-storage RK3's wall root for the ghosts of u2 starts from u0, u1 not
 * being stored). */
int dim2d_set_stage_fusion(dim2d_ctx* ctx, int32_t fuse, int32_t rows_per_block);

/* Kernel timing with HIP events recorded on the context's stream around the
 * kernel launches of the following steps whose kind is in the bit mask
 * `enable` (1 stage 1, 2 stage 2, 4 stage 3, 8 boundary kernels: the wall
 * ghost kernel, and with open edges and stage fusion the frame kernels of
 * stages 1 and 2; 15 all),
 * or off (enable = 0, which also discards what was recorded).  Every event
 * pair costs a few microseconds of step time, and while profiling,
 * dim2d_integrate launches the steps one by one instead of replaying them
 * from a graph.  No counterpart in the reference, which only reports CPU_TIME
 * totals (sim_dim2d.f:46,64,106). */
int dim2d_profile_enable(dim2d_ctx* ctx, int32_t enable);

/* Mean duration (ms) and launch count per kernel since enabling (blocking):
 * out[0],out[1] stage 1, or the fused stages 1 and 2 with stage fusion on
 * (out[2], out[3] then stay 0); out[2],out[3] stage 2; out[4],out[5] stage 3;
 * out[6],out[7] boundary kernels (as in dim2d_profile_enable). */
int dim2d_profile_read(dim2d_ctx* ctx, double out[8]);

/* ------------------------------------------------------------------ */
/* Domain extension at open edges (host-driven).                        */
/* ------------------------------------------------------------------ */

/* The context's current parameters: nx, ny and the coordinate range change
 * when the domain is extended. */
int dim2d_get_params(dim2d_ctx* ctx, dim2d_params* out);

/* Open edges only (blocking).  out[0..3] = number of points, among the
 * `distance` interior columns (W, E) or rows (S, N) nearest each edge, where
 * the open boundary conditions are undermined (are_openbc_undermined,
 * pmodel_eq_class.f:1991-2070): the speed of sound is imaginary,
 * P + rho^2 (2 rho - 3) < 0, or, when md_threshold >= 0 (the reference's
 * bf_openbc_md_threshold_ac = .true. with bf_openbc_md_threshold =
 * md_threshold, parameters_input.f:833-846; its default is off: pass a
 * negative value), the density lies in the two-phase band
 * [rho_vap + eps (rho_mid - rho_vap), rho_liq - eps (rho_liq - rho_mid)] of the
 * saturation densities at the point's temperature 3/(8 cv) temperature_eff.
 * The reference checks the points of its increase detectors,
 * obc_dct_distance = 5 points inside each edge (parameters_input.f:679,
 * parameters_bf_layer.f:138-142); this checks every point of the band between
 * the edge and that line. */
int dim2d_openbc_undermined(dim2d_ctx* ctx, int32_t distance, double md_threshold,
                            int32_t out[4]);

/* Open edges only (blocking).  The conservative far-field state [rho, rho u,
 * rho v, E] that the new points of a domain extension take their incoming
 * characteristics from: the initial condition's get_far_field
 * (pmodel_eq_class.f:2097-2109), uniform here as for the reference's
 * transported-bubble conditions.  Setting it switches dim2d_extend to the
 * reference's new-point rule and makes every step first copy the two
 * outermost rows and columns of u^n into a snapshot (the state at t - dt of
 * that rule; 4 (nx + ny - 4) cells); NULL switches both off.  No counterpart
 * in the reference's interface: there the far field comes from the initial
 * condition module compiled into the solver. */
int dim2d_set_far_field(dim2d_ctx* ctx, const double far_field[4]);

/* Open edges only (blocking).  Grow the grid by add_w/add_e columns and
 * add_s/add_n rows, keeping dx and dy: x_min moves add_w*dx west, x_max
 * add_e*dx east, and so on.  The state moves to its place in the larger
 * grid.  Time, dt and the step count carry over; the ctx handle stays valid
 * and its device buffers are reallocated.
 * New points, with a far field set (dim2d_set_far_field): each add_* is 0 or
 * 1 and at least one step must have run since the last upload or extension
 * (DIM2D_ESTATE otherwise).  Every new point is extrapolated in time along
 * the characteristics from the last step's u^n (t - dt) and the current
 * state (t) by the reference's procedure for its configuration of old
 * points (bf_newgrdpt_class.f:183-2400 compute_newgrdpt_x/_y/_xy with
 * bf_newgrdpt_procedure_module.f get_config_procedures): compute_newgrdpt_x
 * on new columns, _y on new rows (centered transverse gradient, one-sided at
 * their ends), _xy along the diagonal at new corners; outgoing
 * characteristics from the grid, incoming ones from the far field.  The
 * reference adds these points around each new interior point of its buffer
 * layers (bf_interface_icr_class.f:251-386); here whole edges move, one layer
 * per call.
 * Without a far field every new point takes the value of the nearest point
 * of the old grid (the old edge point repeated along the edge normal, the old
 * corner point in the new corners), any number of layers at once (no
 * counterpart in the reference). */
int dim2d_extend(dim2d_ctx* ctx, int32_t add_w, int32_t add_e, int32_t add_s, int32_t add_n);

/* Open edges only.  Columns/rows added per edge (W, E, S, N) by dim2d_extend
 * since the context was created, less those removed by dim2d_shrink: the
 * grid the context was created on plays the reference's interior domain, the
 * added points its buffer layers. */
int dim2d_get_growth(dim2d_ctx* ctx, int32_t out[4]);

/* Open edges only (blocking).  Remove del_w/del_e columns and del_s/del_n
 * rows added earlier (0 <= del_k <= dim2d_get_growth's out[k]): the grid is
 * cropped, dx and dy kept, every remaining point keeps its state (the new
 * ghost frame takes the state at its points, which open edges integrate as
 * state).  Time, dt and the step count carry over; the handle stays valid.
 * The reference removes whole buffer layers (bf_interface_dcr, the
 * remove_inactivated_bf_layers of bf_interface_dcr_class.f:101-216). */
int dim2d_shrink(dim2d_ctx* ctx, int32_t del_w, int32_t del_e, int32_t del_s, int32_t del_n);

/* Open edges only (blocking).  The decrease check of the reference
 * (check_if_bf_layer_remains, bf_decrease_module.f:90-600): out[0..3] =
 * number of undermined points (criterion as in dim2d_openbc_undermined) in
 * the check band of each edge that has grown since creation: `distance` rows
 * (S, N) or columns (W, E) on each side of the edge's place at creation (the
 * reference's search_dcr = obc_dct_distance), along the created grid's
 * extent and distance - 1 points beyond it (the W/E bands one more point to
 * the south, as get_check_line_param), clipped to the points whose 3x3
 * neighbourhood lies in the grid; 0 for an edge that has not grown. */
int dim2d_openbc_undermined_decrease(dim2d_ctx* ctx, int32_t distance, double md_threshold,
                                     int32_t out[4]);

/* field_extended%adapt_domain (field_extended_class.f:427-452,
 * bf_interface_dcr_class.f:243-294), called by the host between steps.
 * Bits 4..7 of adapt_mask (16 W, 32 E, 64 S, 128 N): the decrease first,
 * every such edge that has grown and whose check band
 * (dim2d_openbc_undermined_decrease) has no undermined point goes back to its
 * place at creation (dim2d_shrink).  Bits 0..3 (1 W, 2 E, 4 S, 8 N: the
 * reference's adapt_*_choice, parameters_input.f:796-817): then the increase,
 * each such edge where dim2d_openbc_undermined(distance, md_threshold) finds
 * points moves out by `grow` points (dim2d_extend; grow must be 1 with a far
 * field, and with a far field an edge cropped in this call lets no edge grow
 * before the next step).  out[0..3] = points added (negative: removed) per
 * edge. */
int dim2d_adapt_domain(dim2d_ctx* ctx, int32_t adapt_mask, int32_t distance, double md_threshold,
                       int32_t grow, int32_t out[4]);

/* ------------------------------------------------------------------ */
/* Batches of independent simulations (uneven grid sizes).              */
/* ------------------------------------------------------------------ */
typedef struct dim2d_batch dim2d_batch;

/* Create a batch of n simulations; every simulation has its own params
 * (grid size, physics, boundaries; all must share the precision mode).  The
 * stage kernels walk one work list over all simulations in one persistent
 * launch per RK stage, and dim2d_batch_integrate replays those launches from a
 * captured HIP graph. */
int dim2d_batch_create(const dim2d_params* params, int32_t n, dim2d_batch** out);
int dim2d_batch_destroy(dim2d_batch* b);
int dim2d_batch_upload(dim2d_batch* b, int32_t sim, const double* nodes);
int dim2d_batch_download(dim2d_batch* b, int32_t sim, double* nodes);
/* nsteps RK3 steps of every simulation, each with its own fixed dt (dts[n]). */
int dim2d_batch_integrate(dim2d_batch* b, const double* dts, int32_t nsteps);
/* Tile selection for the batch's separate stage kernels (stage 3, or all
 * three without stage fusion): block_x 64 (one strip per workgroup; the
 * persistent kernel does not pair strips) or 0, rows per tile at most (each
 * simulation's tiles share the smallest height that keeps their number), and
 * the number of persistent workgroups (0: one per work item, capped by
 * residency). */
int dim2d_batch_set_tiles(dim2d_batch* b, int32_t block_x, int32_t rows_per_block,
                          int32_t workgroups);
int dim2d_batch_synchronize(dim2d_batch* b);

/* RK3-TVD stages 1 and 2 of every simulation in one persistent launch (the
 * stage pair of dim2d_set_stage_fusion: the stage-1 state stays on chip),
 * then stage 3 as before.  On by default when every simulation has
 * reflection, periodic or S/N wall edges and at least 8 interior points per
 * axis; fuse = 1 on any other batch returns DIM2D_EINVAL.  rows_per_block
 * (at most, balanced per simulation as in dim2d_batch_set_tiles) and
 * workgroups shape the pair's tiles (56-column strips) and persistent grid
 * (0: 32 rows; one workgroup per tile, capped by residency).
 * fuse = 2 pipelines the batch: the simulations form two groups of about
 * equal cell counts, the second half a step behind the first, and each step
 * is two persistent launches that run one group's stage pair beside the
 * other group's stage 3 (instruction-bound and HBM-bound work side by side);
 * workgroups then sizes those launches.  Results are those of fuse = 1, bit
 * for bit.  The default is 2 for FP64 batches of two or more simulations
 * that allow fusion, 1 for mixed-precision ones (whose stage pair needs more
 * waves per SIMD than a mixed launch can hold).  No counterpart in the
 * reference. */
int dim2d_batch_set_stage_fusion(dim2d_batch* b, int32_t fuse, int32_t rows_per_block,
                                 int32_t workgroups);
/* The launch shapes in use: out = {stage fusion (0, 1 or 2 as set), pair
 * rows, pair (pipelined: mixed-launch) workgroups launched, block_x, rows,
 * workgroups launched}. */
int dim2d_batch_get_tiles(dim2d_batch* b, int32_t out[6]);

/* Tile autotuner of a batch (no counterpart in the reference): times nsteps
 * (>= 2) steps with the given dts for each candidate, keeps the fastest, and
 * restores the state it started from.  With stage fusion it first tries the
 * pair's tile heights (16, 24, 32, 48, 60, 96, 128 rows at most: a
 * simulation's tiles share the smallest height that keeps their number) and
 * persistent workgroup
 * counts (1x and 2x the pair workgroups that fit the device at once), then
 * the other stage kernels' tile heights (8, 12, 16, 24, 32 rows) and
 * workgroup counts.  out[0] = rows, out[1] = workgroups of the latter,
 * out[2] = ms per step of the fastest candidate; dim2d_batch_get_tiles gives
 * the pair's. */
int dim2d_batch_autotune(dim2d_batch* b, const double* dts, int32_t nsteps, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
