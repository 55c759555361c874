/*
 * dim2d_oracle.h -- CPU restatement of the reference DIM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links or calls it.
 *
 * It restates, function by function and in the same evaluation order, the
 * reference's serial Fortran:
 *   physical_models/dim2d/dim2d_prim_module.f       (primitives, EOS)
 *   physical_models/dim2d/dim2d_fluxes_module.f     (26 face fluxes)
 *   sd_operators/mattsson_operators/sd_operators_class.f, sd_operators_fd_module.f
 *   physical_models/dim2d/pmodel_eq_class.f:1073-1214 (flux loops), :1840-1918 (body forces)
 *   td_operators/finitevolume/td_operators_class.f:119-235 (compute_time_dev)
 *   td_integrator/rungekutta3rdtvd/{td_integrator_class.f:103-161, rk3tvd_steps_module.f}
 *   boundary_conditions/bc_operators_gen_class.f:95-102,238-381 (sections, order)
 *   boundary_conditions/reflection_xy/, periodic_xy/bc_operators_class.f:156-210
 *   boundary_conditions/wall_xy/{bc_operators_wall_xy_class.f, wall_xy_equilibrium_module.f,
 *                                ridders_method_module.f}
 *   physical_models/dim2d/dim2d_state_eq_module.f
 *
 * Arrays use the reference's Fortran layout nodes(nx,ny,ne): i fastest, then j,
 * then the variable k (1: mass, 2: momentum_x, 3: momentum_y, 4: total energy).
 * Indices in this file are 1-based like the reference.
 */
#ifndef DIM2D_ORACLE_H
#define DIM2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary kinds per edge */
enum { ORC_BC_REFLECTION = 0, ORC_BC_PERIODIC = 1, ORC_BC_WALL = 2, ORC_BC_OPEN = 3 };
/* ORC_BC_OPEN: hedstrom_xy, only on all four edges at once */

typedef struct orc_model {
  int nx, ny;            /* total grid points including bc_size=2 ghosts */
  double re, we, pr, viscous_r, cv_r;
  double gravity;        /* gravity_amp; body force active when gravity_ac != 0 */
  int gravity_ac;
  int bc_w, bc_e, bc_s, bc_n;
  /* wall parameters (wall_xy_equilibrium_module.f, parameters_input.f:524-611) */
  double wall_contact_angle;   /* micro contact angle in degrees */
  double wall_heat_flux;       /* constant heat flux (0: no_heat_source) */
  int integrator;              /* ORC_INT_RK3TVD or ORC_INT_LSRK3 */
  int operators;               /* ORC_OPS_MATTSSON or ORC_OPS_COCKBURN_GAU (interior faces) */
} orc_model;

/* interior face stencils (sd_choice): Mattsson, or Cockburn and Gau */
enum { ORC_OPS_MATTSSON = 0, ORC_OPS_COCKBURN_GAU = 1 };

/* time integrators: the reference's RK3-TVD, and a low-storage RK3 that the
 * reference does not have (see orc_lsrk3_step) */
enum { ORC_INT_RK3TVD = 0, ORC_INT_LSRK3 = 1 };

/* ---- primitives (point functions) ---- */
double orc_mass_density(const orc_model*, const double*, int, int);
double orc_velocity_x(const orc_model*, const double*, int, int);
double orc_velocity_y(const orc_model*, const double*, int, int);
double orc_classical_pressure(const orc_model*, const double*, int, int);
double orc_classical_temperature_eff(const orc_model*, const double*, int, int);
double orc_speed_of_sound_local(const orc_model*, const double nodes4[4]);

/* ---- Mattsson stencil values of the mass density (pins) ----
 * which: 0 f, 1 dfdx, 2 dfdy, 3 d2fdx2, 4 d2fdy2, 5 d2fdxdy, 6 gradient_x,
 *        7 g, 8 dgdx, 9 dgdy, 10 d2gdx2, 11 d2gdy2, 12 d2gdxdy, 13 gradient_y,
 *        14 dfdx_nl, 15 dgdy_nl (of gradient_x + gradient_y of the density),
 *        16-27 the one-sided stencils of the open boundaries
 *        (sd_operators_{y,x}_oneside_{L0,R0}_class.f): 16 dfdy, 17 d2fdy2,
 *        18 dfdx_nl of y_L0; 19-21 of y_R0; 22 dgdx, 23 d2gdx2, 24 dgdy_nl of
 *        x_L0; 25-27 of x_R0  */
double orc_mattsson_density_stencil(const orc_model*, const double* nodes, int which,
                                    int i, int j, double dx, double dy);

/* ---- the 26 DIM flux functions, in the order of the reference's flux test ---- */
double orc_flux_term(const orc_model*, const double* nodes, int which, int i, int j,
                     double dx, double dy);

/* ---- finite-volume time derivative (compute_time_dev) ----
 * time_dev(nx,ny,4); ghosts are set to the reference's debug_real. */
void orc_compute_time_dev(const orc_model*, double t, const double* nodes,
                          const double* x_map, const double* y_map, double* time_dev);

/* ---- ghost cells (apply_bc_on_nodes) ---- */
void orc_apply_bc_on_nodes(const orc_model*, double t, const double* x_map,
                           const double* y_map, const double* nodes_tmp, double* nodes);

/* ---- one RK3-TVD step (td_integrator%integrate); work: 2*nx*ny*4 doubles ---- */
void orc_rk3_step(const orc_model*, double t, double dt, const double* x_map,
                  const double* y_map, double* nodes, double* work);

/* one step of the 2N-storage low-storage RK3 of Williamson (1980, J. Comput.
 * Phys. 35:48), no counterpart in the reference: for stages k = 1..3,
 *   q = A_k q + dt L(u),  u = u + B_k q   on the interior cells,
 *   A = (0, -5/9, -153/128), B = (1/3, 15/16, 8/15),
 * then apply_bc_on_nodes with the stage's input state as the wall root guess.
 * work: 3*nx*ny*4 doubles. */
void orc_lsrk3_step(const orc_model*, double t, double dt, const double* x_map,
                    const double* y_map, double* nodes, double* work);
/* the same stage recurrence on u' = lambda u from u = 1: the amplification
 * factor R(z), z = lambda dt (for the pins: 1 + z + z^2/2 + z^3/6) */
double orc_lsrk3_amplification(double z);

/* nsteps steps of the model's integrator with a given dt sequence
 * (dts[nsteps]); t advances by dt.
 * Returns 0, or the 1-based step at which a non-finite interior value appeared. */
int orc_run(const orc_model*, double t0, int nsteps, const double* dts,
            const double* x_map, const double* y_map, double* nodes);

/* characteristic time derivative of an open boundary point
 * (hedstrom_xy_module.f:445-558): direction ydir (0: x, 1: y), side right
 * (0: left, 1: right), one-sided gradient grad_type (0: L0, 1: R0) with
 * spacing dn; eigen-quantities [rho, ux, uy, P, c] from the point itself
 * (obc_eigenqties_bc) when pe is NULL, else the given ones
 * (obc_eigenqties_lin). */
void orc_openbc_timedev(const orc_model*, const double* nodes, int i, int j, int ydir, int right,
                        int grad_type, double dn, const double pe[5], double out[4]);

/* ---- new grid points of the buffer layers (bf_layer_newgrdpt/bf_newgrdpt_class.f) ----
 * A grid as the reference passes it to the new-point procedures: extents,
 * alignment bf_align(1,1), bf_align(2,1) (only differences between the two
 * grids matter), coordinate maps and nodes(nx,ny,4).  Grid 0 holds the state
 * at t - dt, grid 1 the state at t and the new point (i1, j1) (1-based).
 * ff is the conservative far-field state (the initial condition's
 * get_far_field, uniform here); eigen_lin selects obc_eigenqties_lin instead
 * of the default obc_eigenqties_bc (parameters_input.f:649). */
typedef struct orc_bf_grid {
  int nx, ny;
  int align_x, align_y;
  const double* x_map;
  const double* y_map;
  const double* nodes;
} orc_bf_grid;
/* gradient procedures: interior (centered), one-sided L0, one-sided R0 */
enum { ORC_GRAD_I = 0, ORC_GRAD_L0 = 1, ORC_GRAD_R0 = 2 };
/* diagonal directions of compute_newgrdpt_xy */
enum { ORC_N1 = 1, ORC_N2 = 2 };
/* procedure types (parameters_constant.f *_edge_type, *_corner_type) */
enum {
  ORC_NG_N_EDGE = 0, ORC_NG_S_EDGE = 1, ORC_NG_E_EDGE = 2, ORC_NG_W_EDGE = 3,
  ORC_NG_NE_CORNER = 4, ORC_NG_NW_CORNER = 5, ORC_NG_SE_CORNER = 6, ORC_NG_SW_CORNER = 7
};
/* helpers: nodes/coeff as [point][variable] (2 or 3 points x 4 variables) */
void orc_interp_coeff_1d(const double x_map[2], const double nodes[8], double coeff[8]);
void orc_interpolate_1d(double x, const double coeff[8], double out[4]);
void orc_interp_coeff_2d(const double x_map[3], const double y_map[3], const double nodes[12],
                         double coeff[12]);
void orc_interpolate_2d(double x, double y, const double coeff[12], double out[4]);
void orc_newton_cotes(const double d0[4], const double d1[4], double dt, double out[4]);
/* compute_newgrdpt_x / _y: side_right 1 (right) or 0 (left); grad_* the
 * transverse gradient procedure (ORC_GRAD_*) */
void orc_newgrdpt_x(const orc_model*, const double ff[4], int eigen_lin, double dt,
                    const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int side_right,
                    int grad_y, double out[4]);
void orc_newgrdpt_y(const orc_model*, const double ff[4], int eigen_lin, double dt,
                    const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int side_right,
                    int grad_x, double out[4]);
/* compute_newgrdpt_xy: grads[6] = x and y gradient procedures of the three
 * interpolation points; eigen_idx[2] and inter_idx1[6] are (i, j) pairs in grid 1 */
void orc_newgrdpt_xy(const orc_model*, const double ff[4], int eigen_lin, double dt,
                     const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int n_dir,
                     int side_right, const int grads[6], const int eigen_idx[2],
                     const int inter_idx1[6], double out[4]);
/* compute_newgrdpt_local for a procedure ORC_NG_* and gradient ORC_GRAD_* */
void orc_newgrdpt_local(const orc_model*, const double ff[4], int eigen_lin, double dt,
                        const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1,
                        int procedure, int grad_type, double out[4]);
/* get_config_procedures for the configurations of a rectangle grown by one
 * layer; returns 0, or -1 for a configuration outside that set */
int orc_newgrdpt_config_procedure(int config_id, int* procedure, int* grad_type);
/* a rectangular grid grown by one layer at the edges with add[e] = 1
 * (W, E, S, N; each 0 or 1); the old state at t (nodes1) keeps its place,
 * new points from the state at t - dt (nodes0) by the characteristic
 * extrapolation.  out: the new grid's nodes.  Returns 0 or -1. */
int orc_extend_newgrdpt(const orc_model*, const double ff[4], double dt, const double* nodes0,
                        const double* nodes1, const double* x_map, const double* y_map,
                        const int add[4], double* out);

/* coordinates (field_abstract_class.f:286-315) */
void orc_coordinates(int nx, int ny, double x_min, double x_max, double y_min,
                     double y_max, double* x_map, double* y_map);

/* time step rule used by the build (not in the reference): dt = cfl*min(
 * dx/max(|u|+c), Re*min(rho)*dx^2, sqrt(We/max(rho))*dx^2 ), dx = min(dx,dy),
 * over interior cells; c = sqrt(max(c^2, 0)) with c^2 the reduced form of
 * speed_of_sound (0 inside the spinodal region).  out[0]=dt, out[1]=max(|u|+c), out[2]=min rho, out[3]=max rho */
void orc_cfl_dt(const orc_model*, const double* nodes, double dx, double dy,
                double cfl, double out[4]);

/* are_openbc_undermined (pmodel_eq_class.f:1991-2070), default criterion:
 * 1 if P + rho^2 (2 rho - 3) < 0 (imaginary speed of sound) at the point */
int orc_openbc_undermined_point(const orc_model*, const double v[4]);
/* the criterion at interior point (i, j) (1-based) with neighbours; with
 * md_thr >= 0 also the mass-density band test (bf_openbc_md_threshold) */
int orc_openbc_undermined_at(const orc_model*, const double* nodes, int i, int j, double dx,
                             double dy, double md_thr);
/* undermined points within `dist` interior columns/rows of each edge,
 * out[0..3] = W, E, S, N */
void orc_openbc_undermined(const orc_model*, const double* nodes, int dist, double dx, double dy,
                           double md_thr, int out[4]);

/* conserved sums over interior cells: out[0]=sum rho, [1]=sum qx, [2]=sum qy, [3]=sum E */
void orc_sums(const orc_model*, const double* nodes, double out[4]);

/* wall ghost density root solve, exposed for pins:
 * returns the Ridders root of the wall equilibrium function
 * (wall_xy_equilibrium_module.f:1747-1821) for one ghost cell. */
double orc_wall_md_ghost(const orc_model*, double T_guess, double md_guess,
                         const double md_x[2], const double md_y[2], double velocity_x,
                         double velocity_y, double Ed, double delta_x, double delta_y);
double orc_wall_root_f(const orc_model*, double md, const double md_x[2],
                       const double md_y[2], double velocity_x, double velocity_y,
                       double Ed, double delta_x, double delta_y);
double orc_mass_density_liquid(double T);
/* Cockburn and Gau stencils of the mass density (cases as orc_mattsson_density_stencil) */
double orc_cg_density_stencil(const orc_model*, const double* nodes, int which, int i, int j,
                              double dx, double dy);
double orc_mass_density_vapor(double T);
double orc_surface_tension(const orc_model*, double T);
double orc_dwall_internal_energy_dmd(const orc_model*, double md, double temperature,
                                     double micro_angle);
double orc_wall_temperature(const orc_model*, double md, double md_grad_squared, double ux,
                            double uy, double Ed);
/* generic bracket search and Ridders root for pins with the reference's test functions */
int orc_root_brackets(double (*f)(void*, double), void* ctx, double T_guess, double md_guess,
                      double out[2]);
double orc_ridders(double (*f)(void*, double), void* ctx, double x1, double x2, double xacc,
                   int* status);
/* wall face flux (compute_wall_flux_x / _y) at face (i,j); dir_y selects y */
void orc_wall_flux(const orc_model*, const double* nodes, int dir_y, int i, int j, double dx,
                   double dy, int side_right, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
