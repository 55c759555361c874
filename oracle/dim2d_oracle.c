/*
 * dim2d_oracle.c -- CPU restatement of the reference DIM time-derivative and
 * RK3-TVD path.  TEST INFRASTRUCTURE ONLY (see dim2d_oracle.h): the checker,
 * never the thing measured or shipped.
 *
 * Every function follows the reference function named beside it, keeping the
 * reference's structure (a stencil re-evaluates its point function at every
 * stencil point) and its left-to-right evaluation order.  Compile without
 * floating-point contraction (-ffp-contract=off).
 */
#include "dim2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Planted mistakes in the step composition, for the test that shows the
 * composition pins fail on each of them (tests/test_oracle_composition.py
 * compiles copies with -DORC_PLANT=1..9).  The default build plants none:
 *   1 no ghost fill after stage 1          2 stage 2 blends the stage-1 state
 *   3 stage-3 weight 1/4 for 1/3           4 periodic x / walls y: wall rows
 *   5 periodic x / walls y: ghost rows       before the x-periodic fill
 *     not x-periodic (stale corners)       6 reflection: corners before W/E
 *   7 stages 2, 3 take the derivative      8 walls W/E, periodic y: ghost
 *     of u^n                                 columns not y-periodic
 *   9 periodic x and y: the y fill over the interior columns only */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

/* NF90_FILL_DOUBLE, the reference's debug_real (parameters_input.f:892) */
#define DEBUG_REAL 9.9692099683868690e+36
#define BC_SIZE 2

/* nodes(i,j,k), 1-based, Fortran layout */
#define ND(a, i, j, k) \
  (a)[((size_t)((k) - 1) * (size_t)m->ny + (size_t)((j) - 1)) * (size_t)m->nx + (size_t)((i) - 1)]

typedef double (*prim_fn)(const orc_model*, const double*, int, int);
typedef double (*grad_fn)(const orc_model*, const double*, int, int, prim_fn, double);
typedef double (*sec_fn)(const orc_model*, const double*, int, int, double, double, grad_fn,
                         grad_fn);

static double sq(double x) { return x * x; }

/* ------------------------------------------------------------------ */
/* primitives: dim2d_prim_module.f                                      */
/* ------------------------------------------------------------------ */

/* dim2d_prim_module.f:99-110 */
double orc_mass_density(const orc_model* m, const double* n, int i, int j) { return ND(n, i, j, 1); }
/* :135-146 */
static double momentum_x(const orc_model* m, const double* n, int i, int j) { return ND(n, i, j, 2); }
/* :171-182 */
static double momentum_y(const orc_model* m, const double* n, int i, int j) { return ND(n, i, j, 3); }
/* :242-253 */
double orc_velocity_x(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 2) / ND(n, i, j, 1);
}
/* :278-289 */
double orc_velocity_y(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 3) / ND(n, i, j, 1);
}

/* classical_pressure_local, :425-453 (rkind=8 branch) */
static double classical_pressure_local(const orc_model* m, const double* v) {
  return 3.0 / ((3.0 - v[0]) * m->cv_r) *
             (v[3] - 0.5 * v[0] * (sq(v[1] / v[0]) + sq(v[2] / v[0])) + 3.0 * sq(v[0])) -
         3.0 * sq(v[0]);
}

/* classical_pressure, :392-403 */
double orc_classical_pressure(const orc_model* m, const double* n, int i, int j) {
  double v[4] = {ND(n, i, j, 1), ND(n, i, j, 2), ND(n, i, j, 3), ND(n, i, j, 4)};
  return classical_pressure_local(m, v);
}

/* pressure(temperature, mass_density), :476-488 */
static double pressure_T_md(double temperature, double md) {
  return 8.0 * md * temperature / (3.0 - md) - 3.0 * sq(md);
}

/* classical_temperature_eff, :515-544 */
double orc_classical_temperature_eff(const orc_model* m, const double* n, int i, int j) {
  const double r = ND(n, i, j, 1);
  return 1.0 / r *
         (ND(n, i, j, 4) - 0.5 * r * (sq(ND(n, i, j, 2) / r) + sq(ND(n, i, j, 3) / r)) + 3.0 * sq(r));
}

/* capillarity_temperature_eff, :582-613 */
static double capillarity_temperature_eff(const orc_model* m, const double* n, int i, int j,
                                          double dx, double dy, grad_fn gx, grad_fn gy) {
  return 0.5 / ND(n, i, j, 1) *
         (sq(gx(m, n, i, j, orc_mass_density, dx)) + sq(gy(m, n, i, j, orc_mass_density, dy)));
}

/* temperature_eff, :652-691 */
static double temperature_eff(const orc_model* m, const double* n, int i, int j, double dx,
                              double dy, grad_fn gx, grad_fn gy) {
  const double r = ND(n, i, j, 1);
  return 1.0 / r *
         (ND(n, i, j, 4) - 0.5 * r * (sq(ND(n, i, j, 2) / r) + sq(ND(n, i, j, 3) / r)) -
          0.5 / m->we *
              (sq(gx(m, n, i, j, orc_mass_density, dx)) + sq(gy(m, n, i, j, orc_mass_density, dy))) +
          3.0 * sq(r));
}

/* classical_pressure_xwork / ywork, :718-806 */
static double classical_pressure_xwork(const orc_model* m, const double* n, int i, int j) {
  const double r = ND(n, i, j, 1);
  return (3.0 / ((3.0 - r) * m->cv_r) *
              (ND(n, i, j, 4) - 0.5 * r * (sq(ND(n, i, j, 2) / r) + sq(ND(n, i, j, 3) / r)) +
               3.0 * sq(r)) -
          3.0 * sq(r)) *
         ND(n, i, j, 2) / r;
}
static double classical_pressure_ywork(const orc_model* m, const double* n, int i, int j) {
  const double r = ND(n, i, j, 1);
  return (3.0 / ((3.0 - r) * m->cv_r) *
              (ND(n, i, j, 4) - 0.5 * r * (sq(ND(n, i, j, 2) / r) + sq(ND(n, i, j, 3) / r)) +
               3.0 * sq(r)) -
          3.0 * sq(r)) *
         ND(n, i, j, 3) / r;
}

/* transport products, :830-1000 */
static double qx_transport_x(const orc_model* m, const double* n, int i, int j) {
  return sq(ND(n, i, j, 2)) / ND(n, i, j, 1);
}
static double qy_transport_x(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 3) * ND(n, i, j, 2) / ND(n, i, j, 1);
}
static double qx_transport_y(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 2) * ND(n, i, j, 3) / ND(n, i, j, 1);
}
static double qy_transport_y(const orc_model* m, const double* n, int i, int j) {
  return sq(ND(n, i, j, 3)) / ND(n, i, j, 1);
}
static double energy_transport_x(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 4) * ND(n, i, j, 2) / ND(n, i, j, 1);
}
static double energy_transport_y(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 4) * ND(n, i, j, 3) / ND(n, i, j, 1);
}
/* capillarity_pressure{,_xwork,_ywork}, :1030-1150 */
static double capillarity_pressure(const orc_model* m, const double* n, int i, int j) {
  return 1.0 / (3.0 - ND(n, i, j, 1));
}
static double capillarity_pressure_xwork(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 2) / (ND(n, i, j, 1) * (3.0 - ND(n, i, j, 1)));
}
static double capillarity_pressure_ywork(const orc_model* m, const double* n, int i, int j) {
  return ND(n, i, j, 3) / (ND(n, i, j, 1) * (3.0 - ND(n, i, j, 1)));
}

/* speed_of_sound, :1179-1215 */
double orc_speed_of_sound_local(const orc_model* m, const double v[4]) {
  const double P = classical_pressure_local(m, v);
  const double e = P + sq(v[0]) * (-3.0 + 2.0 * v[0]);
  const double b = sqrt((P + 3.0 * sq(v[0])) / (m->cv_r * e));
  const double a = 1.0 / sqrt(1.0 + sq(b));
  const double d = sqrt(v[0] * v[0] * v[0]) * sqrt((3.0 - v[0]) / e);
  return sqrt(3.0) * v[0] / (a * d);
}

/* ------------------------------------------------------------------ */
/* Mattsson stencils: sd_operators_class.f, sd_operators_fd_module.f    */
/* ------------------------------------------------------------------ */

/* gradient_x_interior / gradient_y_interior, sd_operators_fd_module.f:71-110 */
static double gradient_x_interior(const orc_model* m, const double* n, int i, int j, prim_fn p,
                                  double dn) {
  return 0.5 / dn * (-p(m, n, i - 1, j) + p(m, n, i + 1, j));
}
static double gradient_y_interior(const orc_model* m, const double* n, int i, int j, prim_fn p,
                                  double dn) {
  return 0.5 / dn * (-p(m, n, i, j - 1) + p(m, n, i, j + 1));
}
/* f_mattsson :185 */
static double s_f(const orc_model* m, const double* n, int i, int j, prim_fn p) {
  return 0.5 * (+p(m, n, i - 1, j) + p(m, n, i, j));
}
/* dfdx_mattsson :239 */
static double s_dfdx(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / dx * (-p(m, n, i - 1, j) + p(m, n, i, j));
}
/* dfdx_mattsson_nl :300 */
static double s_dfdx_nl(const orc_model* m, const double* n, int i, int j, sec_fn p, double dx,
                        double dy) {
  return 1.0 / dx *
         (-p(m, n, i - 1, j, dx, dy, gradient_x_interior, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior));
}
/* dfdy_mattsson :359 */
static double s_dfdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / (2.0 * dy) *
         (-0.5 * p(m, n, i - 1, j - 1) - 0.5 * p(m, n, i, j - 1) + 0.5 * p(m, n, i - 1, j + 1) +
          0.5 * p(m, n, i, j + 1));
}
/* d2fdx2_mattsson :421 */
static double s_d2fdx2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 0.5 / (dx * dx) *
         (+p(m, n, i - 2, j) - p(m, n, i - 1, j) - p(m, n, i, j) + p(m, n, i + 1, j));
}
/* d2fdy2_mattsson :485 */
static double s_d2fdy2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return (1.0 / (dy * dy)) *
         (+0.5 * p(m, n, i - 1, j - 1) + 0.5 * p(m, n, i, j - 1) -
          2.0 * (+0.5 * p(m, n, i - 1, j) + 0.5 * p(m, n, i, j)) + 0.5 * p(m, n, i - 1, j + 1) +
          0.5 * p(m, n, i, j + 1));
}
/* d2fdxdy_mattsson :549 */
static double s_d2fdxdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx,
                        double dy) {
  return (p(m, n, i - 1, j - 1) - p(m, n, i, j - 1) - p(m, n, i - 1, j + 1) + p(m, n, i, j + 1)) *
         0.5 / (dy * dx);
}
/* g_mattsson :622 */
static double s_g(const orc_model* m, const double* n, int i, int j, prim_fn p) {
  return 0.5 * (+p(m, n, i, j - 1) + p(m, n, i, j));
}
/* dgdx_mattsson :681 */
static double s_dgdx(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 0.25 / dx *
         (-p(m, n, i - 1, j - 1) - p(m, n, i - 1, j) + p(m, n, i + 1, j - 1) + p(m, n, i + 1, j));
}
/* dgdy_mattsson :744 */
static double s_dgdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / dy * (-p(m, n, i, j - 1) + p(m, n, i, j));
}
/* dgdy_mattsson_nl :806 */
static double s_dgdy_nl(const orc_model* m, const double* n, int i, int j, sec_fn p, double dx,
                        double dy) {
  return 1.0 / dy *
         (-p(m, n, i, j - 1, dx, dy, gradient_x_interior, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior));
}
/* d2gdx2_mattsson :868 */
static double s_d2gdx2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 0.5 / (dx * dx) *
         (+p(m, n, i - 1, j - 1) + p(m, n, i - 1, j) - 2.0 * (+p(m, n, i, j - 1) + p(m, n, i, j)) +
          p(m, n, i + 1, j - 1) + p(m, n, i + 1, j));
}
/* d2gdy2_mattsson :940 */
static double s_d2gdy2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 0.5 / (dy * dy) *
         (+p(m, n, i, j - 2) - p(m, n, i, j - 1) - p(m, n, i, j) + p(m, n, i, j + 1));
}
/* d2gdxdy_mattsson :1003 */
static double s_d2gdxdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx,
                        double dy) {
  return (p(m, n, i - 1, j - 1) - p(m, n, i - 1, j) - p(m, n, i + 1, j - 1) + p(m, n, i + 1, j)) *
         0.5 / (dx * dy);
}

/* An sd_operators object: the face stencils a flux evaluation uses.  The
 * interior (Mattsson) set and the one-sided sets of the open boundaries
 * (sd_operators_{x,y}_oneside_{L0,R0}_class.f) differ in a few members; the
 * L1/R1 classes override only the other face family, so for the fluxes they
 * are the interior set. */
typedef double (*st0_fn)(const orc_model*, const double*, int, int, prim_fn);
typedef double (*st1_fn)(const orc_model*, const double*, int, int, prim_fn, double);
typedef double (*stnl_fn)(const orc_model*, const double*, int, int, sec_fn, double, double);
typedef struct sd_ops {
  st0_fn f;
  st1_fn dfdx;
  stnl_fn dfdx_nl;
  st1_fn dfdy, d2fdx2, d2fdy2;
  st0_fn g;
  st1_fn dgdx, dgdy;
  stnl_fn dgdy_nl;
  st1_fn d2gdx2, d2gdy2;
} sd_ops;
static const sd_ops MATTSSON = {s_f,    s_dfdx, s_dfdx_nl, s_dfdy,    s_d2fdx2, s_d2fdy2,
                                s_g,    s_dgdx, s_dgdy,    s_dgdy_nl, s_d2gdx2, s_d2gdy2};

/* Cockburn and Gau operators (sd_operators/cg_operators/sd_operators_class.f:
 * f :151, dfdx :209, dfdx_nl :270, dfdy :329, d2fdx2 :399, d2fdy2, d2fdxdy,
 * g, dgdx, dgdy, dgdy_nl, d2gdx2, d2gdy2, d2gdxdy; bc_size 2, :118-121) */
static double cg_f(const orc_model* m, const double* n, int i, int j, prim_fn p) {
  return 1.0 / 12.0 *
         (-p(m, n, i - 2, j) + 7.0 * p(m, n, i - 1, j) + 7.0 * p(m, n, i, j) - p(m, n, i + 1, j));
}
static double cg_dfdx(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / dx * (-p(m, n, i - 1, j) + p(m, n, i, j));
}
static double cg_dfdx_nl(const orc_model* m, const double* n, int i, int j, sec_fn p, double dx,
                         double dy) {
  return 1.0 / dx *
         (-p(m, n, i - 1, j, dx, dy, gradient_x_interior, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior));
}
static double cg_dfdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / (24.0 * dy) *
         (+p(m, n, i - 2, j - 1) - 7.0 * p(m, n, i - 1, j - 1) - 7.0 * p(m, n, i, j - 1) +
          p(m, n, i + 1, j - 1) - p(m, n, i - 2, j + 1) + 7.0 * p(m, n, i - 1, j + 1) +
          7.0 * p(m, n, i, j + 1) - p(m, n, i + 1, j + 1));
}
static double cg_d2fdx2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 0.5 / (dx * dx) *
         (+p(m, n, i - 2, j) - p(m, n, i - 1, j) - p(m, n, i, j) + p(m, n, i + 1, j));
}
static double cg_d2fdy2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return (1.0 / (12.0 * dy * dy)) *
         ((-p(m, n, i - 2, j - 1) + 7.0 * p(m, n, i - 1, j - 1) + 7.0 * p(m, n, i, j - 1) -
           p(m, n, i + 1, j - 1)) -
          2.0 * (-p(m, n, i - 2, j) + 7.0 * p(m, n, i - 1, j) + 7.0 * p(m, n, i, j) -
                 p(m, n, i + 1, j)) +
          (-p(m, n, i - 2, j + 1) + 7.0 * p(m, n, i - 1, j + 1) + 7.0 * p(m, n, i, j + 1) -
           p(m, n, i + 1, j + 1)));
}
static double cg_d2fdxdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx,
                         double dy) {
  return (p(m, n, i - 1, j - 1) - p(m, n, i, j - 1) - p(m, n, i - 1, j + 1) + p(m, n, i, j + 1)) *
         0.5 / (dy * dx);
}
static double cg_g(const orc_model* m, const double* n, int i, int j, prim_fn p) {
  return 1.0 / 12.0 *
         (-p(m, n, i, j - 2) + 7.0 * p(m, n, i, j - 1) + 7.0 * p(m, n, i, j) - p(m, n, i, j + 1));
}
static double cg_dgdx(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / (24.0 * dx) *
         (+p(m, n, i - 1, j - 2) - 7.0 * p(m, n, i - 1, j - 1) - 7.0 * p(m, n, i - 1, j) +
          p(m, n, i - 1, j + 1) - p(m, n, i + 1, j - 2) + 7.0 * p(m, n, i + 1, j - 1) +
          7.0 * p(m, n, i + 1, j) - p(m, n, i + 1, j + 1));
}
static double cg_dgdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / dy * (-p(m, n, i, j - 1) + p(m, n, i, j));
}
static double cg_dgdy_nl(const orc_model* m, const double* n, int i, int j, sec_fn p, double dx,
                         double dy) {
  return 1.0 / dy *
         (-p(m, n, i, j - 1, dx, dy, gradient_x_interior, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior));
}
static double cg_d2gdx2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / (12.0 * dx * dx) *
         (-p(m, n, i - 1, j - 2) + 7.0 * p(m, n, i - 1, j - 1) + 7.0 * p(m, n, i - 1, j) -
          p(m, n, i - 1, j + 1) -
          2.0 * (-p(m, n, i, j - 2) + 7.0 * p(m, n, i, j - 1) + 7.0 * p(m, n, i, j) -
                 p(m, n, i, j + 1)) -
          p(m, n, i + 1, j - 2) + 7.0 * p(m, n, i + 1, j - 1) + 7.0 * p(m, n, i + 1, j) -
          p(m, n, i + 1, j + 1));
}
static double cg_d2gdy2(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 0.5 / (dy * dy) *
         (+p(m, n, i, j - 2) - p(m, n, i, j - 1) - p(m, n, i, j) + p(m, n, i, j + 1));
}
static double cg_d2gdxdy(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx,
                         double dy) {
  return (p(m, n, i - 1, j - 1) - p(m, n, i - 1, j) - p(m, n, i + 1, j - 1) + p(m, n, i + 1, j)) *
         0.5 / (dx * dy);
}
static const sd_ops COCKBURN_GAU = {cg_f, cg_dfdx, cg_dfdx_nl, cg_dfdy,    cg_d2fdx2, cg_d2fdy2,
                                    cg_g, cg_dgdx, cg_dgdy,    cg_dgdy_nl, cg_d2gdx2, cg_d2gdy2};

/* the interior operator set of a model (sd_choice: mattsson or cockburnandgau) */
static const sd_ops* interior_ops(const orc_model* m) {
  return m->operators == ORC_OPS_COCKBURN_GAU ? &COCKBURN_GAU : &MATTSSON;
}

/* gradient_av of the reference stencil test (test_mattsson_operators.f:1219-1238) */
static double gradient_av(const orc_model* m, const double* n, int i, int j, double dx, double dy,
                          grad_fn gx, grad_fn gy) {
  return gx(m, n, i, j, orc_mass_density, dx) + gy(m, n, i, j, orc_mass_density, dy);
}

double orc_oneside_density_stencil(const orc_model* m, const double* n, int which, int i, int j,
                                   double dx, double dy);

double orc_mattsson_density_stencil(const orc_model* m, const double* n, int which, int i, int j,
                                    double dx, double dy) {
  if (which >= 16) return orc_oneside_density_stencil(m, n, which, i, j, dx, dy);
  prim_fn p = orc_mass_density;
  switch (which) {
    case 0: return s_f(m, n, i, j, p);
    case 1: return s_dfdx(m, n, i, j, p, dx);
    case 2: return s_dfdy(m, n, i, j, p, dy);
    case 3: return s_d2fdx2(m, n, i, j, p, dx);
    case 4: return s_d2fdy2(m, n, i, j, p, dy);
    case 5: return s_d2fdxdy(m, n, i, j, p, dx, dy);
    case 6: return gradient_x_interior(m, n, i, j, p, dx);
    case 7: return s_g(m, n, i, j, p);
    case 8: return s_dgdx(m, n, i, j, p, dx);
    case 9: return s_dgdy(m, n, i, j, p, dy);
    case 10: return s_d2gdx2(m, n, i, j, p, dx);
    case 11: return s_d2gdy2(m, n, i, j, p, dy);
    case 12: return s_d2gdxdy(m, n, i, j, p, dx, dy);
    case 13: return gradient_y_interior(m, n, i, j, p, dy);
    case 14: return s_dfdx_nl(m, n, i, j, gradient_av, dx, dy);
    case 15: return s_dgdy_nl(m, n, i, j, gradient_av, dx, dy);
    default: return NAN;
  }
}

/* the Cockburn and Gau stencils of the mass density at (i, j), numbered as
 * orc_mattsson_density_stencil's cases 0-15 */
double orc_cg_density_stencil(const orc_model* m, const double* n, int which, int i, int j,
                              double dx, double dy) {
  prim_fn p = orc_mass_density;
  switch (which) {
    case 0: return cg_f(m, n, i, j, p);
    case 1: return cg_dfdx(m, n, i, j, p, dx);
    case 2: return cg_dfdy(m, n, i, j, p, dy);
    case 3: return cg_d2fdx2(m, n, i, j, p, dx);
    case 4: return cg_d2fdy2(m, n, i, j, p, dy);
    case 5: return cg_d2fdxdy(m, n, i, j, p, dx, dy);
    case 6: return gradient_x_interior(m, n, i, j, p, dx);
    case 7: return cg_g(m, n, i, j, p);
    case 8: return cg_dgdx(m, n, i, j, p, dx);
    case 9: return cg_dgdy(m, n, i, j, p, dy);
    case 10: return cg_d2gdx2(m, n, i, j, p, dx);
    case 11: return cg_d2gdy2(m, n, i, j, p, dy);
    case 12: return cg_d2gdxdy(m, n, i, j, p, dx, dy);
    case 13: return gradient_y_interior(m, n, i, j, p, dy);
    case 14: return cg_dfdx_nl(m, n, i, j, gradient_av, dx, dy);
    case 15: return cg_dgdy_nl(m, n, i, j, gradient_av, dx, dy);
    default: return NAN;
  }
}

/* ------------------------------------------------------------------ */
/* fluxes: dim2d_fluxes_module.f                                       */
/* ------------------------------------------------------------------ */

static double flux_x_mass_density(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->f(m, n, i, j, momentum_x);
}
static double flux_y_mass_density(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->g(m, n, i, j, momentum_y);
}

/* momentum_x */
static double flux_x_inviscid_momentum_x(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->f(m, n, i, j, qx_transport_x) + S->f(m, n, i, j, orc_classical_pressure);
}
static double flux_x_viscid_momentum_x(const orc_model* m, const double* n, int i, int j, double dx,
                                       double dy, const sd_ops* S) {
  return (2.0 + m->viscous_r) * S->dfdx(m, n, i, j, orc_velocity_x, dx) +
         m->viscous_r * S->dfdy(m, n, i, j, orc_velocity_y, dy);
}
static double flux_x_capillarity_momentum_x(const orc_model* m, const double* n, int i, int j,
                                            double dx, double dy, const sd_ops* S) {
  prim_fn r = orc_mass_density;
  return 1.5 / m->cv_r * S->f(m, n, i, j, capillarity_pressure) *
             (sq(S->dfdx(m, n, i, j, r, dx)) + sq(S->dfdy(m, n, i, j, r, dy))) +
         S->f(m, n, i, j, r) * (S->d2fdx2(m, n, i, j, r, dx) + S->d2fdy2(m, n, i, j, r, dy)) +
         0.5 * (-sq(S->dfdx(m, n, i, j, r, dx)) + sq(S->dfdy(m, n, i, j, r, dy)));
}
static double flux_y_inviscid_momentum_x(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->g(m, n, i, j, qx_transport_y);
}
static double flux_y_viscid_momentum_x(const orc_model* m, const double* n, int i, int j, double dx,
                                       double dy, const sd_ops* S) {
  return S->dgdy(m, n, i, j, orc_velocity_x, dy) + S->dgdx(m, n, i, j, orc_velocity_y, dx);
}
static double flux_y_capillarity_momentum_x(const orc_model* m, const double* n, int i, int j,
                                            double dx, double dy, const sd_ops* S) {
  return -S->dgdx(m, n, i, j, orc_mass_density, dx) * S->dgdy(m, n, i, j, orc_mass_density, dy);
}

/* momentum_y */
static double flux_x_inviscid_momentum_y(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->f(m, n, i, j, qy_transport_x);
}
static double flux_x_viscid_momentum_y(const orc_model* m, const double* n, int i, int j, double dx,
                                       double dy, const sd_ops* S) {
  return S->dfdy(m, n, i, j, orc_velocity_x, dy) + S->dfdx(m, n, i, j, orc_velocity_y, dx);
}
static double flux_x_capillarity_momentum_y(const orc_model* m, const double* n, int i, int j,
                                            double dx, double dy, const sd_ops* S) {
  return -S->dfdx(m, n, i, j, orc_mass_density, dx) * S->dfdy(m, n, i, j, orc_mass_density, dy);
}
static double flux_y_inviscid_momentum_y(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->g(m, n, i, j, qy_transport_y) + S->g(m, n, i, j, orc_classical_pressure);
}
static double flux_y_viscid_momentum_y(const orc_model* m, const double* n, int i, int j, double dx,
                                       double dy, const sd_ops* S) {
  return m->viscous_r * S->dgdx(m, n, i, j, orc_velocity_x, dx) +
         (2.0 + m->viscous_r) * S->dgdy(m, n, i, j, orc_velocity_y, dy);
}
static double flux_y_capillarity_momentum_y(const orc_model* m, const double* n, int i, int j,
                                            double dx, double dy, const sd_ops* S) {
  prim_fn r = orc_mass_density;
  return 1.5 / m->cv_r * S->g(m, n, i, j, capillarity_pressure) *
             (sq(S->dgdx(m, n, i, j, r, dx)) + sq(S->dgdy(m, n, i, j, r, dy))) +
         S->g(m, n, i, j, r) * (S->d2gdx2(m, n, i, j, r, dx) + S->d2gdy2(m, n, i, j, r, dy)) +
         0.5 * (sq(S->dgdx(m, n, i, j, r, dx)) - sq(S->dgdy(m, n, i, j, r, dy)));
}

/* total_energy */
static double flux_x_inviscid_total_energy(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->f(m, n, i, j, energy_transport_x) + S->f(m, n, i, j, classical_pressure_xwork);
}
static double flux_x_viscid_total_energy(const orc_model* m, const double* n, int i, int j,
                                         double dx, double dy, const sd_ops* S) {
  const double ux = S->f(m, n, i, j, orc_velocity_x);
  const double uy = S->f(m, n, i, j, orc_velocity_y);
  return ux * ((2.0 + m->viscous_r) * S->dfdx(m, n, i, j, orc_velocity_x, dx) +
               m->viscous_r * S->dfdy(m, n, i, j, orc_velocity_y, dy)) +
         uy * (S->dfdy(m, n, i, j, orc_velocity_x, dy) + S->dfdx(m, n, i, j, orc_velocity_y, dx)) +
         1.0 / m->pr * S->dfdx(m, n, i, j, orc_classical_temperature_eff, dx);
}
static double flux_x_capillarity_total_energy(const orc_model* m, const double* n, int i, int j,
                                              double dx, double dy, const sd_ops* S) {
  prim_fn r = orc_mass_density;
  const double ux = S->f(m, n, i, j, orc_velocity_x);
  const double uy = S->f(m, n, i, j, orc_velocity_y);
  const double duxdx = S->dfdx(m, n, i, j, orc_velocity_x, dx);
  const double duydy = S->dfdy(m, n, i, j, orc_velocity_y, dy);
  const double drhodx = S->dfdx(m, n, i, j, r, dx);
  const double drhody = S->dfdy(m, n, i, j, r, dy);
  return 1.5 / m->cv_r * S->f(m, n, i, j, capillarity_pressure_xwork) *
             (sq(drhodx) + sq(drhody)) +
         S->f(m, n, i, j, momentum_x) * (S->d2fdx2(m, n, i, j, r, dx) + S->d2fdy2(m, n, i, j, r, dy)) +
         0.5 * ux * (-sq(drhodx) + sq(drhody)) - drhodx * drhody * uy -
         S->f(m, n, i, j, r) * drhodx * (duxdx + duydy) -
         1.0 / (m->re * m->pr) * S->dfdx_nl(m, n, i, j, capillarity_temperature_eff, dx, dy);
}
static double flux_y_inviscid_total_energy(const orc_model* m, const double* n, int i, int j, const sd_ops* S) {
  return S->g(m, n, i, j, energy_transport_y) + S->g(m, n, i, j, classical_pressure_ywork);
}
static double flux_y_viscid_total_energy(const orc_model* m, const double* n, int i, int j,
                                         double dx, double dy, const sd_ops* S) {
  const double ux = S->g(m, n, i, j, orc_velocity_x);
  const double uy = S->g(m, n, i, j, orc_velocity_y);
  const double duxdx = S->dgdx(m, n, i, j, orc_velocity_x, dx);
  const double duydy = S->dgdy(m, n, i, j, orc_velocity_y, dy);
  return ux * (S->dgdy(m, n, i, j, orc_velocity_x, dy) + S->dgdx(m, n, i, j, orc_velocity_y, dx)) +
         uy * (m->viscous_r * duxdx + (2.0 + m->viscous_r) * duydy) +
         1.0 / m->pr * S->dgdy(m, n, i, j, orc_classical_temperature_eff, dy);
}
static double flux_y_capillarity_total_energy(const orc_model* m, const double* n, int i, int j,
                                              double dx, double dy, const sd_ops* S) {
  prim_fn r = orc_mass_density;
  const double ux = S->g(m, n, i, j, orc_velocity_x);
  const double uy = S->g(m, n, i, j, orc_velocity_y);
  const double duxdx = S->dgdx(m, n, i, j, orc_velocity_x, dx);
  const double duydy = S->dgdy(m, n, i, j, orc_velocity_y, dy);
  const double drhodx = S->dgdx(m, n, i, j, r, dx);
  const double drhody = S->dgdy(m, n, i, j, r, dy);
  return 1.5 / m->cv_r * S->g(m, n, i, j, capillarity_pressure_ywork) *
             (sq(S->dgdx(m, n, i, j, r, dx)) + sq(S->dgdy(m, n, i, j, r, dy))) +
         S->g(m, n, i, j, momentum_y) * (S->d2gdx2(m, n, i, j, r, dx) + S->d2gdy2(m, n, i, j, r, dy)) +
         0.5 * uy * (sq(drhodx) - sq(drhody)) - drhodx * drhody * ux -
         S->g(m, n, i, j, r) * drhody * (duxdx + duydy) -
         1.0 / (m->re * m->pr) * S->dgdy_nl(m, n, i, j, capillarity_temperature_eff, dx, dy);
}

/* total fluxes: inviscid - 1/Re viscid - 1/We capillarity */
static double flux_x_momentum_x(const orc_model* m, const double* n, int i, int j, double dx,
                                double dy, const sd_ops* S) {
  return flux_x_inviscid_momentum_x(m, n, i, j, S) -
         1.0 / m->re * flux_x_viscid_momentum_x(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_x_capillarity_momentum_x(m, n, i, j, dx, dy, S);
}
static double flux_y_momentum_x(const orc_model* m, const double* n, int i, int j, double dx,
                                double dy, const sd_ops* S) {
  return flux_y_inviscid_momentum_x(m, n, i, j, S) -
         1.0 / m->re * flux_y_viscid_momentum_x(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_y_capillarity_momentum_x(m, n, i, j, dx, dy, S);
}
static double flux_x_momentum_y(const orc_model* m, const double* n, int i, int j, double dx,
                                double dy, const sd_ops* S) {
  return flux_x_inviscid_momentum_y(m, n, i, j, S) -
         1.0 / m->re * flux_x_viscid_momentum_y(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_x_capillarity_momentum_y(m, n, i, j, dx, dy, S);
}
static double flux_y_momentum_y(const orc_model* m, const double* n, int i, int j, double dx,
                                double dy, const sd_ops* S) {
  return flux_y_inviscid_momentum_y(m, n, i, j, S) -
         1.0 / m->re * flux_y_viscid_momentum_y(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_y_capillarity_momentum_y(m, n, i, j, dx, dy, S);
}
static double flux_x_total_energy(const orc_model* m, const double* n, int i, int j, double dx,
                                  double dy, const sd_ops* S) {
  return flux_x_inviscid_total_energy(m, n, i, j, S) -
         1.0 / m->re * flux_x_viscid_total_energy(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_x_capillarity_total_energy(m, n, i, j, dx, dy, S);
}
static double flux_y_total_energy(const orc_model* m, const double* n, int i, int j, double dx,
                                  double dy, const sd_ops* S) {
  return flux_y_inviscid_total_energy(m, n, i, j, S) -
         1.0 / m->re * flux_y_viscid_total_energy(m, n, i, j, dx, dy, S) -
         1.0 / m->we * flux_y_capillarity_total_energy(m, n, i, j, dx, dy, S);
}

double orc_flux_term(const orc_model* m, const double* n, int which, int i, int j, double dx,
                     double dy) {
  switch (which) {
    case 0: return flux_x_mass_density(m, n, i, j, &MATTSSON);
    case 1: return flux_y_mass_density(m, n, i, j, &MATTSSON);
    case 2: return flux_x_inviscid_momentum_x(m, n, i, j, &MATTSSON);
    case 3: return flux_x_viscid_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 4: return flux_x_capillarity_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 5: return flux_x_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 6: return flux_y_inviscid_momentum_x(m, n, i, j, &MATTSSON);
    case 7: return flux_y_viscid_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 8: return flux_y_capillarity_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 9: return flux_y_momentum_x(m, n, i, j, dx, dy, &MATTSSON);
    case 10: return flux_x_inviscid_momentum_y(m, n, i, j, &MATTSSON);
    case 11: return flux_x_viscid_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 12: return flux_x_capillarity_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 13: return flux_x_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 14: return flux_y_inviscid_momentum_y(m, n, i, j, &MATTSSON);
    case 15: return flux_y_viscid_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 16: return flux_y_capillarity_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 17: return flux_y_momentum_y(m, n, i, j, dx, dy, &MATTSSON);
    case 18: return flux_x_inviscid_total_energy(m, n, i, j, &MATTSSON);
    case 19: return flux_x_viscid_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    case 20: return flux_x_capillarity_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    case 21: return flux_x_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    case 22: return flux_y_inviscid_total_energy(m, n, i, j, &MATTSSON);
    case 23: return flux_y_viscid_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    case 24: return flux_y_capillarity_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    case 25: return flux_y_total_energy(m, n, i, j, dx, dy, &MATTSSON);
    default: return NAN;
  }
}

/* ------------------------------------------------------------------ */
/* state equation fits: dim2d_state_eq_module.f (dim2d_lowTemperature=F) */
/* ------------------------------------------------------------------ */
double orc_surface_tension(const orc_model* m, double T) {
  return 2.0 * sqrt(2.0) / (3.0 * sqrt(m->we)) * (9.691686257306483 * pow(sqrt(1.0 - T), 3));
}
double orc_mass_density_liquid(double T) { return 1.0 + 2.06 * sqrt(1.0 - T); }
double orc_mass_density_vapor(double T) { return 1.0 - 1.91 * sqrt(1.0 - T); }

/* ------------------------------------------------------------------ */
/* wall boundary: wall_xy_equilibrium_module.f, ridders_method_module.f */
/* ------------------------------------------------------------------ */

/* sign_normal, sign_trans: wall_xy_equilibrium_module.f:136-140 */
#define WALL_SIGN_NORMAL 1.0
#define WALL_SIGN_TRANS (-1.0)

/* get_wall_heat_flux with a constant heat source (no source when 0) */
static double wall_heat_flux(const orc_model* m) { return m->wall_heat_flux; }

/* dwallInternalEnergy_dmd, :2238-2293.  The reference tests an uninitialised
 * local against 90 degrees; that test never holds, so the cosine branch is the
 * one that runs. */
double orc_dwall_internal_energy_dmd(const orc_model* m, double md, double temperature,
                                     double micro_angle) {
  const double md_vap = orc_mass_density_vapor(temperature);
  const double md_liq = orc_mass_density_liquid(temperature);
  const double surface_tension = orc_surface_tension(m, temperature);
  const double angle = micro_angle * acos(-1.0) / 180.0;
  return 6.0 * (md - md_liq) * (md - md_vap) / pow(md_liq - md_vap, 3) * surface_tension *
         cos(angle);
}

/* temperature_average, :2295-2336 */
static double temperature_average(const orc_model* m, double T0, double delta_x, double q) {
  return T0 + 0.5 * delta_x * m->re * m->pr * 3.0 / (8.0 * m->cv_r) * q;
}
/* md_average, :2337-2380 */
static double md_average(double md0, double md1) { return 0.5 * (md0 + md1); }
/* temperature, :2382-2432 */
double orc_wall_temperature(const orc_model* m, double md, double md_grad_squared, double ux,
                            double uy, double Ed) {
  return 3.0 / (8.0 * m->cv_r) *
         (Ed / md - 0.5 * (sq(ux) + sq(uy)) - 0.5 / m->we * md_grad_squared / md + 3.0 * md);
}
/* dmddx, :2433-2475 (interior gradient type only: 0 interior, 1 L0, 2 R0) */
static double dmddx(double delta_x, const double t[3], int type) {
  if (type == 1) return (t[2] - t[1]) / delta_x;
  if (type == 2) return (t[1] - t[0]) / delta_x;
  return (t[2] - t[0]) / (2.0 * delta_x);
}

typedef struct wall_root_ctx {
  const orc_model* m;
  double md_x[2], md_y[2], ux, uy, Ed, delta_x, delta_y, angle, q;
  int grad_x, grad_y;
} wall_root_ctx;

/* wall_x_equilibrium_root_fct, :2151-2236 */
static double wall_equilibrium(const wall_root_ctx* c, double md, const double md_x[2]) {
  const orc_model* m = c->m;
  const double dmddx_half = (-md + md_x[1]) / c->delta_x;
  const double md_av = md_average(md, md_x[1]);
  const double tx[3] = {md_x[0], md, md_x[1]};
  const double ty[3] = {c->md_y[0], md, c->md_y[1]};
  const double md_grad_squared =
      sq(dmddx(c->delta_x, tx, c->grad_x)) + sq(dmddx(c->delta_y, ty, c->grad_y));
  const double temperature1 = orc_wall_temperature(m, md, md_grad_squared, c->ux, c->uy, c->Ed);
  const double T_average = temperature_average(m, temperature1, c->delta_x, c->q);
  return 1.0 / m->we * dmddx_half + orc_dwall_internal_energy_dmd(m, md_av, T_average, c->angle);
}

/* wall_x_root_f, :2081-2102: f(x) = equilibrium(md_x(2), [md_x(1), x]) */
static double wall_root_f(void* vc, double x) {
  const wall_root_ctx* c = (const wall_root_ctx*)vc;
  const double mx[2] = {c->md_x[0], x};
  return wall_equilibrium(c, c->md_x[1], mx);
}

double orc_wall_root_f(const orc_model* m, double md, const double md_x[2], const double md_y[2],
                       double velocity_x, double velocity_y, double Ed, double delta_x,
                       double delta_y) {
  wall_root_ctx c = {m, {md_x[0], md_x[1]}, {md_y[0], md_y[1]}, velocity_x, velocity_y, Ed,
                     delta_x, delta_y, m->wall_contact_angle, wall_heat_flux(m), 0, 0};
  return wall_equilibrium(&c, md, md_x);
}

typedef double (*root_fn)(void*, double);

static int have_same_sign(double x1, double x2) {
  return (copysign(1.0, x1) * copysign(1.0, x2)) > 0;
}

/* is_real_validated(var, cst): |var - cst| < 1e-10 (check_data_module.f:28-46) */
static int validated(double v, double c) { return fabs(v - c) < 1e-10; }

/* get_wall_x_root_brackets, :1846-1941; returns 0 or -1 when no bracket found */
int orc_root_brackets(root_fn f, void* ctx, double T_guess, double md_guess, double out[2]) {
  const int MAXIT = 99;
  double xl = orc_mass_density_vapor(T_guess);
  double xh = orc_mass_density_liquid(T_guess);
  double fl = f(ctx, xl);
  double fh = f(ctx, xh);
  if (have_same_sign(fl, fh)) {
    const double dmd_l = md_guess / (double)(float)(MAXIT + 1);
    const double dmd_r = (3.0 - md_guess) / (double)(float)(MAXIT + 1);
    const double fmd = f(ctx, md_guess);
    for (int j = 1; j <= MAXIT; ++j) {
      xl = md_guess - j * dmd_l;
      xh = md_guess + j * dmd_r;
      fl = f(ctx, xl);
      fh = f(ctx, xh);
      if (!have_same_sign(fl, fmd)) {
        out[0] = xl;
        out[1] = md_guess;
        return 0;
      } else if (!have_same_sign(fh, fmd)) {
        out[0] = md_guess;
        out[1] = xh;
        return 0;
      }
    }
    return -1;
  }
  out[0] = xl;
  out[1] = xh;
  return 0;
}

/* get_root_ridder_method, ridders_method_module.f:112-216; status -1 on failure */
double orc_ridders(root_fn f, void* ctx, double x1, double x2, double xacc, int* status) {
  const int MAXIT = 60;
  const double UNUSED = (double)(-1.11e30f);
  double fl = f(ctx, x1);
  double fh = f(ctx, x2);
  *status = 0;
  if ((fl > 0.0 && fh < 0.0) || (fl < 0.0 && fh > 0.0)) {
    double xl = x1, xh = x2, x = UNUSED;
    for (int j = 1; j <= MAXIT; ++j) {
      const double xm = 0.5 * (xl + xh);
      const double fm = f(ctx, xm);
      const double s = sqrt(sq(fm) - fl * fh);
      if (validated(s, 0.0)) return x;
      const double xnew = xm + (xm - xl) * (copysign(1.0, fl - fh) * fm / s);
      if (fabs(xnew - x) <= xacc) return x;
      x = xnew;
      const double fnew = f(ctx, x);
      if (validated(fnew, 0.0)) return x;
      if (!validated(copysign(fm, fnew), fm)) {
        xl = xm;
        fl = fm;
        xh = x;
        fh = fnew;
      } else if (!validated(copysign(fl, fnew), fl)) {
        xh = x;
        fh = fnew;
      } else if (!validated(copysign(fh, fnew), fh)) {
        xl = x;
        fl = fnew;
      } else {
        *status = -1;
        return x;
      }
      if (fabs(xh - xl) <= xacc) return x;
    }
    *status = -1;
    return x;
  } else if (validated(fl, 0.0)) {
    return x1;
  } else if (validated(fh, 0.0)) {
    return x2;
  }
  *status = -1;
  return NAN;
}

static double wall_md_ghost_ctx(wall_root_ctx* c, double T_guess, double md_guess) {
  double br[2];
  int st = 0;
  if (orc_root_brackets(wall_root_f, c, T_guess, md_guess, br) != 0) return NAN;
  /* xacc = 1.0d0*1e-12 with a default-kind literal */
  const double r = orc_ridders(wall_root_f, c, br[0], br[1], 1.0 * (double)1e-12f, &st);
  return st == 0 ? r : NAN;
}

/* get_wall_x_md_ghost_cell, :1747-1821 */
double orc_wall_md_ghost(const orc_model* m, double T_guess, double md_guess, const double md_x[2],
                         const double md_y[2], double velocity_x, double velocity_y, double Ed,
                         double delta_x, double delta_y) {
  wall_root_ctx c = {m, {md_x[0], md_x[1]}, {md_y[0], md_y[1]}, velocity_x, velocity_y, Ed,
                     delta_x, delta_y, m->wall_contact_angle, wall_heat_flux(m), 0, 0};
  return wall_md_ghost_ctx(&c, T_guess, md_guess);
}

/* get_wall_md_extra_ghost_cell with wall_dmddx_high_order, :1439-1465 */
static double wall_md_extra(double md_far, double md_first, double md_ghost) {
  return md_far + 3.0 * (-md_first + md_ghost);
}

/* compute_wall_xy_guess, bc_operators_wall_xy_class.f:898-930 */
static void wall_guess(const orc_model* m, const double* nodes_tmp, int ti, int tj, int mi, int mj,
                       double dx, double dy, double* T_guess, double* md_guess) {
  *T_guess = 3.0 / (8.0 * m->cv_r) *
             temperature_eff(m, nodes_tmp, ti, tj, dx, dy, gradient_x_interior, gradient_y_interior);
  *md_guess = ND(nodes_tmp, mi, mj, 1);
}

/* compute_wall_{S,N}_ghost_cell, wall_xy_equilibrium_module.f:948-1178.
 * S: j is the inner ghost row, the interior neighbour is j+1, the far one j+2.
 * N: j is the inner ghost row, the interior neighbour is j-1, the far one j-2. */
static void wall_y_ghost(const orc_model* m, double* nodes, int i, int j, int north, double dx,
                         double dy, double T_guess, double md_guess) {
  const int j1 = north ? j - 1 : j + 1;
  const int j2 = north ? j - 2 : j + 2;
  const int jo = north ? j + 1 : j - 1;
  wall_root_ctx c;
  c.m = m;
  c.md_x[0] = ND(nodes, i, j2, 1);
  c.md_x[1] = ND(nodes, i, j1, 1);
  c.md_y[0] = ND(nodes, i - 1, j1, 1);
  c.md_y[1] = ND(nodes, i + 1, j1, 1);
  /* S: velocity_x = -qy/rho, velocity_y = qx/rho; N: velocity_x = qy/rho, velocity_y = qx/rho */
  c.ux = north ? ND(nodes, i, j1, 3) / ND(nodes, i, j1, 1) : -ND(nodes, i, j1, 3) / ND(nodes, i, j1, 1);
  c.uy = ND(nodes, i, j1, 2) / ND(nodes, i, j1, 1);
  c.Ed = ND(nodes, i, j1, 4);
  c.delta_x = dy;
  c.delta_y = dx;
  c.angle = m->wall_contact_angle;
  c.q = wall_heat_flux(m);
  c.grad_x = 0;
  c.grad_y = 0;
  ND(nodes, i, j, 1) = wall_md_ghost_ctx(&c, T_guess, md_guess);
  ND(nodes, i, jo, 1) = wall_md_extra(ND(nodes, i, j2, 1), ND(nodes, i, j1, 1), ND(nodes, i, j, 1));
  ND(nodes, i, j, 2) = WALL_SIGN_TRANS * ND(nodes, i, j1, 2) / ND(nodes, i, j1, 1) * ND(nodes, i, j, 1);
  ND(nodes, i, j, 3) = WALL_SIGN_NORMAL * ND(nodes, i, j1, 3) / ND(nodes, i, j1, 1) * ND(nodes, i, j, 1);
}

/* compute_wall_{W,E}_ghost_cell, :1211-1437 (transposed roles) */
static void wall_x_ghost(const orc_model* m, double* nodes, int i, int j, int east, double dx,
                         double dy, double T_guess, double md_guess) {
  const int i1 = east ? i - 1 : i + 1;
  const int i2 = east ? i - 2 : i + 2;
  const int io = east ? i + 1 : i - 1;
  wall_root_ctx c;
  c.m = m;
  c.md_x[0] = ND(nodes, i2, j, 1);
  c.md_x[1] = ND(nodes, i1, j, 1);
  c.md_y[0] = ND(nodes, i1, j - 1, 1);
  c.md_y[1] = ND(nodes, i1, j + 1, 1);
  /* W: velocity_x = -qx/rho, velocity_y = qy/rho; E: velocity_x = qx/rho, velocity_y = qy/rho */
  c.ux = east ? ND(nodes, i1, j, 2) / ND(nodes, i1, j, 1) : -ND(nodes, i1, j, 2) / ND(nodes, i1, j, 1);
  c.uy = ND(nodes, i1, j, 3) / ND(nodes, i1, j, 1);
  c.Ed = ND(nodes, i1, j, 4);
  c.delta_x = dx;
  c.delta_y = dy;
  c.angle = m->wall_contact_angle;
  c.q = wall_heat_flux(m);
  c.grad_x = 0;
  c.grad_y = 0;
  ND(nodes, i, j, 1) = wall_md_ghost_ctx(&c, T_guess, md_guess);
  ND(nodes, io, j, 1) = wall_md_extra(ND(nodes, i2, j, 1), ND(nodes, i1, j, 1), ND(nodes, i, j, 1));
  ND(nodes, i, j, 2) = WALL_SIGN_NORMAL * ND(nodes, i1, j, 2) / ND(nodes, i1, j, 1) * ND(nodes, i, j, 1);
  ND(nodes, i, j, 3) = WALL_SIGN_TRANS * ND(nodes, i1, j, 3) / ND(nodes, i1, j, 1) * ND(nodes, i, j, 1);
}

/* wall face fluxes, wall_xy_equilibrium_module.f:249-445, 445-946 (constant
 * heat flux, no extra heat source). side_right=1 for N/E.  The stencils are
 * the interior operator set's (the module takes `s`, the compiled
 * sd_operators, :263,356: Mattsson or Cockburn-Gau); the temperature's
 * gradients are the central ones of sd_operators_fd_module (:97-99). */
static void wall_flux_y(const orc_model* m, const double* n, int i, int j, double dx, double dy,
                        int side_right, const sd_ops* S, double out[4]) {
  prim_fn r = orc_mass_density;
  const double q = wall_heat_flux(m);
  double Tl;
  if (!side_right)
    Tl = 3.0 / (8.0 * m->cv_r) *
         temperature_eff(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior);
  else
    Tl = 3.0 / (8.0 * m->cv_r) *
         temperature_eff(m, n, i, j - 1, dx, dy, gradient_x_interior, gradient_y_interior);
  const double Th = Tl + dy * m->re * m->pr * (3.0 / (8.0 * m->cv_r)) * q;
  const double T_half = 0.5 * (Tl + Th);
  const double md_half = 0.5 * (ND(n, i, j - 1, 1) + ND(n, i, j, 1));
  const double inv_mom_y = pressure_T_md(T_half, md_half);
  const double visc_mom_x = S->dgdy(m, n, i, j, orc_velocity_x, dy);
  const double cap_mom_x = -S->dgdx(m, n, i, j, r, dx) * S->dgdy(m, n, i, j, r, dy);
  const double cap_mom_y =
      S->g(m, n, i, j, r) * (S->d2gdx2(m, n, i, j, r, dx) + S->d2gdy2(m, n, i, j, r, dy)) +
      0.5 * (sq(S->dgdx(m, n, i, j, r, dx)) - sq(S->dgdy(m, n, i, j, r, dy)));
  out[0] = 0.0;
  out[1] = -1.0 / m->re * visc_mom_x - 1.0 / m->we * cap_mom_x;
  out[2] = inv_mom_y - 1.0 / m->we * cap_mom_y;
  out[3] = side_right ? -q : q;
}

static void wall_flux_x(const orc_model* m, const double* n, int i, int j, double dx, double dy,
                        int side_right, const sd_ops* S, double out[4]) {
  prim_fn r = orc_mass_density;
  const double q = wall_heat_flux(m);
  double Tl;
  if (!side_right)
    Tl = 3.0 / (8.0 * m->cv_r) *
         temperature_eff(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_interior);
  else
    Tl = 3.0 / (8.0 * m->cv_r) *
         temperature_eff(m, n, i - 1, j, dx, dy, gradient_x_interior, gradient_y_interior);
  const double Th = Tl + dx * 3.0 / (8.0 * m->cv_r) * m->re * m->pr * q;
  const double T_half = 0.5 * (Tl + Th);
  const double md_half = 0.5 * (ND(n, i - 1, j, 1) + ND(n, i, j, 1));
  const double inv_mom_x = pressure_T_md(T_half, md_half);
  const double cap_mom_x =
      S->f(m, n, i, j, r) * (S->d2fdx2(m, n, i, j, r, dx) + S->d2fdy2(m, n, i, j, r, dy)) +
      0.5 * (-sq(S->dfdx(m, n, i, j, r, dx)) + sq(S->dfdy(m, n, i, j, r, dy)));
  const double visc_mom_y = S->dfdx(m, n, i, j, orc_velocity_y, dx);
  const double cap_mom_y = -S->dfdx(m, n, i, j, r, dx) * S->dfdy(m, n, i, j, r, dy);
  out[0] = 0.0;
  out[1] = inv_mom_x - 1.0 / m->we * cap_mom_x;
  out[2] = -1.0 / m->re * visc_mom_y - 1.0 / m->we * cap_mom_y;
  out[3] = side_right ? -q : q;
}

void orc_wall_flux(const orc_model* m, const double* n, int dir_y, int i, int j, double dx,
                   double dy, int side_right, double out[4]) {
  if (dir_y)
    wall_flux_y(m, n, i, j, dx, dy, side_right, interior_ops(m), out);
  else
    wall_flux_x(m, n, i, j, dx, dy, side_right, interior_ops(m), out);
}

/* ------------------------------------------------------------------ */
/* compute_time_dev: td_operators_class.f:119-235                      */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* open boundaries (hedstrom_xy): one-sided stencils and the           */
/* characteristic time derivative                                       */
/* ------------------------------------------------------------------ */

/* one-sided gradients, mattsson_operators/sd_operators_fd_module.f:182-450 */
static double gradient_x_x_oneside_L0(const orc_model* m, const double* n, int i, int j,
                                      prim_fn p, double dn) {
  return 1.0 / dn * (-p(m, n, i, j) + p(m, n, i + 1, j));
}
static double gradient_x_x_oneside_R0(const orc_model* m, const double* n, int i, int j,
                                      prim_fn p, double dn) {
  return 1.0 / dn * (-p(m, n, i - 1, j) + p(m, n, i, j));
}
static double gradient_y_y_oneside_L0(const orc_model* m, const double* n, int i, int j,
                                      prim_fn p, double dn) {
  return 1.0 / dn * (-p(m, n, i, j) + p(m, n, i, j + 1));
}
static double gradient_y_y_oneside_R0(const orc_model* m, const double* n, int i, int j,
                                      prim_fn p, double dn) {
  return 1.0 / dn * (-p(m, n, i, j - 1) + p(m, n, i, j));
}

/* x-face stencils of the outer S ghost row, sd_operators_y_oneside_L0_class.f */
static double dfdy_y_L0(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / dy *
         (-0.5 * p(m, n, i - 1, j) - 0.5 * p(m, n, i, j) + 0.5 * p(m, n, i - 1, j + 1) +
          0.5 * p(m, n, i, j + 1));
}
static double d2fdy2_y_L0(const orc_model* m, const double* n, int i, int j, prim_fn p,
                          double dy) {
  return 1.0 / (dy * dy) *
         (0.5 * p(m, n, i - 1, j) + 0.5 * p(m, n, i, j) -
          2.0 * (+0.5 * p(m, n, i - 1, j + 1) + 0.5 * p(m, n, i, j + 1)) +
          0.5 * p(m, n, i - 1, j + 2) + 0.5 * p(m, n, i, j + 2));
}
static double dfdx_y_L0_nl(const orc_model* m, const double* n, int i, int j, sec_fn p,
                           double dx, double dy) {
  return 1.0 / dx *
         (-p(m, n, i - 1, j, dx, dy, gradient_x_interior, gradient_y_y_oneside_L0) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_y_oneside_L0));
}
/* x-face stencils of the outer N ghost row, sd_operators_y_oneside_R0_class.f */
static double dfdy_y_R0(const orc_model* m, const double* n, int i, int j, prim_fn p, double dy) {
  return 1.0 / dy *
         (-0.5 * p(m, n, i - 1, j - 1) - 0.5 * p(m, n, i, j - 1) + 0.5 * p(m, n, i - 1, j) +
          0.5 * p(m, n, i, j));
}
static double d2fdy2_y_R0(const orc_model* m, const double* n, int i, int j, prim_fn p,
                          double dy) {
  return 1.0 / (dy * dy) *
         (0.5 * p(m, n, i - 1, j - 2) + 0.5 * p(m, n, i, j - 2) -
          2.0 * (+0.5 * p(m, n, i - 1, j - 1) + 0.5 * p(m, n, i, j - 1)) +
          0.5 * p(m, n, i - 1, j) + 0.5 * p(m, n, i, j));
}
static double dfdx_y_R0_nl(const orc_model* m, const double* n, int i, int j, sec_fn p,
                           double dx, double dy) {
  return 1.0 / dx *
         (-p(m, n, i - 1, j, dx, dy, gradient_x_interior, gradient_y_y_oneside_R0) +
          p(m, n, i, j, dx, dy, gradient_x_interior, gradient_y_y_oneside_R0));
}
/* y-face stencils of the outer W ghost column, sd_operators_x_oneside_L0_class.f */
static double dgdx_x_L0(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / dx *
         (-0.5 * p(m, n, i, j - 1) - 0.5 * p(m, n, i, j) + 0.5 * p(m, n, i + 1, j - 1) +
          0.5 * p(m, n, i + 1, j));
}
static double d2gdx2_x_L0(const orc_model* m, const double* n, int i, int j, prim_fn p,
                          double dx) {
  return 1.0 / (dx * dx) *
         (0.5 * p(m, n, i, j - 1) + 0.5 * p(m, n, i, j) -
          2.0 * (+0.5 * p(m, n, i + 1, j - 1) + 0.5 * p(m, n, i + 1, j)) +
          0.5 * p(m, n, i + 2, j - 1) + 0.5 * p(m, n, i + 2, j));
}
static double dgdy_x_L0_nl(const orc_model* m, const double* n, int i, int j, sec_fn p,
                           double dx, double dy) {
  return 1.0 / dy *
         (-p(m, n, i, j - 1, dx, dy, gradient_x_x_oneside_L0, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_x_oneside_L0, gradient_y_interior));
}
/* y-face stencils of the outer E ghost column, sd_operators_x_oneside_R0_class.f */
static double dgdx_x_R0(const orc_model* m, const double* n, int i, int j, prim_fn p, double dx) {
  return 1.0 / dx *
         (-0.5 * p(m, n, i - 1, j - 1) - 0.5 * p(m, n, i - 1, j) + 0.5 * p(m, n, i, j - 1) +
          0.5 * p(m, n, i, j));
}
static double d2gdx2_x_R0(const orc_model* m, const double* n, int i, int j, prim_fn p,
                          double dx) {
  return 1.0 / (dx * dx) *
         (0.5 * p(m, n, i - 2, j - 1) + 0.5 * p(m, n, i - 2, j) -
          2.0 * (+0.5 * p(m, n, i - 1, j - 1) + 0.5 * p(m, n, i - 1, j)) +
          0.5 * p(m, n, i, j - 1) + 0.5 * p(m, n, i, j));
}
static double dgdy_x_R0_nl(const orc_model* m, const double* n, int i, int j, sec_fn p,
                           double dx, double dy) {
  return 1.0 / dy *
         (-p(m, n, i, j - 1, dx, dy, gradient_x_x_oneside_R0, gradient_y_interior) +
          p(m, n, i, j, dx, dy, gradient_x_x_oneside_R0, gradient_y_interior));
}
/* members a face family does not use keep the interior stencils */
static const sd_ops Y_L0 = {s_f,    s_dfdx, dfdx_y_L0_nl, dfdy_y_L0, s_d2fdx2, d2fdy2_y_L0,
                            s_g,    s_dgdx, s_dgdy,       s_dgdy_nl, s_d2gdx2, s_d2gdy2};
static const sd_ops Y_R0 = {s_f,    s_dfdx, dfdx_y_R0_nl, dfdy_y_R0, s_d2fdx2, d2fdy2_y_R0,
                            s_g,    s_dgdx, s_dgdy,       s_dgdy_nl, s_d2gdx2, s_d2gdy2};
static const sd_ops X_L0 = {s_f,    s_dfdx,    s_dfdx_nl, s_dfdy,       s_d2fdx2,    s_d2fdy2,
                            s_g,    dgdx_x_L0, s_dgdy,    dgdy_x_L0_nl, d2gdx2_x_L0, s_d2gdy2};
static const sd_ops X_R0 = {s_f,    s_dfdx,    s_dfdx_nl, s_dfdy,       s_d2fdx2,    s_d2fdy2,
                            s_g,    dgdx_x_R0, s_dgdy,    dgdy_x_R0_nl, d2gdx2_x_R0, s_d2gdy2};

/* the one-sided stencils of the density, for the pins: 16 dfdy, 17 d2fdy2,
 * 18 dfdx_nl of sd_operators_y_oneside_L0; 19-21 the same of _y_oneside_R0;
 * 22 dgdx, 23 d2gdx2, 24 dgdy_nl of sd_operators_x_oneside_L0; 25-27 of
 * _x_oneside_R0 (nl: of gradient_x + gradient_y of the density) */
double orc_oneside_density_stencil(const orc_model* m, const double* n, int which, int i, int j,
                                   double dx, double dy) {
  prim_fn p = orc_mass_density;
  switch (which) {
    case 16: return dfdy_y_L0(m, n, i, j, p, dy);
    case 17: return d2fdy2_y_L0(m, n, i, j, p, dy);
    case 18: return dfdx_y_L0_nl(m, n, i, j, gradient_av, dx, dy);
    case 19: return dfdy_y_R0(m, n, i, j, p, dy);
    case 20: return d2fdy2_y_R0(m, n, i, j, p, dy);
    case 21: return dfdx_y_R0_nl(m, n, i, j, gradient_av, dx, dy);
    case 22: return dgdx_x_L0(m, n, i, j, p, dx);
    case 23: return d2gdx2_x_L0(m, n, i, j, p, dx);
    case 24: return dgdy_x_L0_nl(m, n, i, j, gradient_av, dx, dy);
    case 25: return dgdx_x_R0(m, n, i, j, p, dx);
    case 26: return d2gdx2_x_R0(m, n, i, j, p, dx);
    case 27: return dgdy_x_R0_nl(m, n, i, j, gradient_av, dx, dy);
    default: return NAN;
  }
}

/* compute_flux_x_oneside / compute_flux_y_oneside with obc_edge_flux_capillarity
 * (pmodel_eq_class.f:1389-1595): the full fluxes with the given operators */
static void flux_x_ops(const orc_model* m, const double* n, int i, int j, double dx, double dy,
                       const sd_ops* S, double f[4]) {
  f[0] = flux_x_mass_density(m, n, i, j, S);
  f[1] = flux_x_momentum_x(m, n, i, j, dx, dy, S);
  f[2] = flux_x_momentum_y(m, n, i, j, dx, dy, S);
  f[3] = flux_x_total_energy(m, n, i, j, dx, dy, S);
}
static void flux_y_ops(const orc_model* m, const double* n, int i, int j, double dx, double dy,
                       const sd_ops* S, double f[4]) {
  f[0] = flux_y_mass_density(m, n, i, j, S);
  f[1] = flux_y_momentum_x(m, n, i, j, dx, dy, S);
  f[2] = flux_y_momentum_y(m, n, i, j, dx, dy, S);
  f[3] = flux_y_total_energy(m, n, i, j, dx, dy, S);
}

/* get_prim_obc_eigenqties with obc_eigenqties_bc (pmodel_eq_class.f:2142-2195):
 * [rho, ux, uy, P, c] of the boundary point itself */
static void obc_eigenqties_bc(const orc_model* m, const double v[4], double pe[5]) {
  pe[0] = v[0];
  pe[1] = v[1] / v[0];
  pe[2] = v[2] / v[0];
  pe[3] = classical_pressure_local(m, v);
  pe[4] = orc_speed_of_sound_local(m, v);
}

/* eigenvalues and left/right eigenvectors of the primitive variables
 * [rho, u, v, P] in direction x (ydir 0) or y (ydir 1), ns_vdw2d_prim_module.f:
 * compute_x_eigenvalues/left/righteigenvector_ns_vdw2d :504-722,
 * compute_y_... :539-891; un is the velocity along the direction.  Fortran
 * matrices A(row,col) are stored as A[row-1][col-1]. */
static void ns_vdw2d_eigen(int ydir, double md, double un, double c, double lam[4],
                           double L[4][4], double R[4][4]) {
  const double q_c = 0.5 * md * c;
  const double inv_c = 1.0 / (c * c);
  const double inv_q_c = 1.0 / (md * c);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) L[a][b] = R[a][b] = 0.0;
  lam[0] = un; lam[1] = un; lam[2] = un - c; lam[3] = un + c;
  if (!ydir) {
    L[2][0] = 1.0;
    L[0][1] = 1.0; L[3][1] = -1.0 / (c * c);
    L[1][2] = -q_c; L[3][2] = 0.5;
    L[1][3] = q_c; L[3][3] = 0.5;
    R[1][0] = 1.0; R[2][0] = inv_c; R[3][0] = inv_c;
    R[2][1] = -inv_q_c; R[3][1] = inv_q_c;
    R[0][2] = 1.0;
    R[2][3] = 1.0; R[3][3] = 1.0;
  } else {
    L[1][0] = 1.0;
    L[0][1] = 1.0; L[3][1] = -1.0 / (c * c);
    L[2][2] = -q_c; L[3][2] = 0.5;
    L[2][3] = q_c; L[3][3] = 0.5;
    R[1][0] = 1.0; R[2][0] = inv_c; R[3][0] = inv_c;
    R[0][1] = 1.0;
    R[2][2] = -inv_q_c; R[3][2] = inv_q_c;
    R[2][3] = 1.0; R[3][3] = 1.0;
  }
}

/* compute_timedev_with_openbc (hedstrom_xy_module.f:445-558) with the
 * eigen-quantities of ns_vdw2d_prim_module.f:249-891 (Fortran matrices
 * A(row,col) stored as A[row-1][col-1]) */
void orc_openbc_timedev(const orc_model* m, const double* n, int i, int j, int ydir, int right,
                        int grad_type, double dn, const double pe_in[5], double out[4]) {
  double v[4] = {ND(n, i, j, 1), ND(n, i, j, 2), ND(n, i, j, 3), ND(n, i, j, 4)};
  double pe[5];
  if (pe_in)
    for (int q = 0; q < 5; ++q) pe[q] = pe_in[q];
  else
    obc_eigenqties_bc(m, v, pe);
  const double md = pe[0], ux = pe[1], uy = pe[2], P = pe[3], c = pe[4];
  double lam[4], L[4][4], R[4][4], J[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) J[a][b] = 0.0;
  ns_vdw2d_eigen(ydir, md, ydir ? uy : ux, c, lam, L, R);
  /* compute_jacobian_cons_to_prim_ns_vdw2d :249-320 */
  const double dEdrho = 0.5 * (ux * ux + uy * uy) - 6.0 * md -
                        1.0 / 3.0 * m->cv_r * (P + 9.0 * (md - 2.0) * md);
  J[0][0] = 1.0;
  J[0][1] = ux; J[1][1] = md;
  J[0][2] = uy; J[2][2] = md;
  J[0][3] = dEdrho; J[1][3] = md * ux; J[2][3] = md * uy; J[3][3] = m->cv_r * (1.0 - md / 3.0);
  /* compute_gradient_prim (pmodel_eq_class.f:3113-3153) */
  grad_fn gr = grad_type == 0 ? (ydir ? gradient_y_y_oneside_L0 : gradient_x_x_oneside_L0)
                              : (ydir ? gradient_y_y_oneside_R0 : gradient_x_x_oneside_R0);
  const double g[4] = {gr(m, n, i, j, orc_mass_density, dn), gr(m, n, i, j, orc_velocity_x, dn),
                       gr(m, n, i, j, orc_velocity_y, dn),
                       gr(m, n, i, j, orc_classical_pressure, dn)};
  /* incoming_left: lambda >= 0, incoming_right: lambda <= 0
   * (openbc_operators_module.f:97-136): no amplitude for incoming waves */
  double amp[4];
  for (int k = 0; k < 4; ++k) {
    const int incoming = right ? (lam[k] <= 0.0) : (lam[k] >= 0.0);
    if (incoming) {
      amp[k] = 0.0;
    } else {
      double dot = 0.0;
      for (int r = 0; r < 4; ++r) dot += L[r][k] * g[r];
      amp[k] = -lam[k] * dot;
    }
  }
  double tp[4];
  for (int col = 0; col < 4; ++col) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += amp[k] * R[k][col];
    tp[col] = acc;
  }
  for (int col = 0; col < 4; ++col) {
    double acc = 0.0;
    for (int l = 0; l < 4; ++l) acc += tp[l] * J[l][col];
    out[col] = acc;
  }
}

/* compute_body_forces, pmodel_eq_class.f:1840-1918 (conservative form) */
static double body_force(const orc_model* m, const double* v, int k) {
  if (!m->gravity_ac) return 0.0;
  if (k == 3) return -v[0] * m->gravity;
  if (k == 4) return -v[2] * m->gravity;
  return 0.0;
}

void orc_compute_time_dev(const orc_model* m, double t, const double* nodes, const double* x_map,
                          const double* y_map, double* time_dev) {
  (void)t;
  const int nx = m->nx, ny = m->ny;
  const size_t plane = (size_t)nx * ny;
  for (size_t q = 0; q < plane * 4; ++q) time_dev[q] = DEBUG_REAL;
  const double dx = x_map[1] - x_map[0];
  const double dy = y_map[1] - y_map[0];

  const sd_ops* ops = interior_ops(m);
  double* flux_x = (double*)calloc((size_t)(nx + 1) * ny * 4, sizeof(double));
  double* flux_y = (double*)calloc((size_t)nx * (ny + 1) * 4, sizeof(double));
#define FX(i, j, k) flux_x[((size_t)((k) - 1) * ny + ((j) - 1)) * (nx + 1) + ((i) - 1)]
#define FY(i, j, k) flux_y[((size_t)((k) - 1) * (ny + 1) + ((j) - 1)) * nx + ((i) - 1)]

  /* compute_flux_x, pmodel_eq_class.f:1073-1139 */
  for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j)
    for (int i = 1 + BC_SIZE; i <= nx + 1 - BC_SIZE; ++i) {
      FX(i, j, 1) = flux_x_mass_density(m, nodes, i, j, ops);
      FX(i, j, 2) = flux_x_momentum_x(m, nodes, i, j, dx, dy, ops);
      FX(i, j, 3) = flux_x_momentum_y(m, nodes, i, j, dx, dy, ops);
      FX(i, j, 4) = flux_x_total_energy(m, nodes, i, j, dx, dy, ops);
    }
  /* compute_flux_y, :1140-1214 */
  for (int j = 1 + BC_SIZE; j <= ny + 1 - BC_SIZE; ++j)
    for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
      FY(i, j, 1) = flux_y_mass_density(m, nodes, i, j, ops);
      FY(i, j, 2) = flux_y_momentum_x(m, nodes, i, j, dx, dy, ops);
      FY(i, j, 3) = flux_y_momentum_y(m, nodes, i, j, dx, dy, ops);
      FY(i, j, 4) = flux_y_total_energy(m, nodes, i, j, dx, dy, ops);
    }

  /* apply_bc_on_fluxes for wall edges (bc_operators_wall_xy_class.f:702-854),
   * interior columns/rows of each edge section */
  double w[4];
  if (m->bc_s == ORC_BC_WALL)
    for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
      wall_flux_y(m, nodes, i, 1 + BC_SIZE, dx, dy, 0, ops, w);
      for (int k = 1; k <= 4; ++k) FY(i, 1 + BC_SIZE, k) = w[k - 1];
    }
  if (m->bc_n == ORC_BC_WALL)
    for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
      wall_flux_y(m, nodes, i, ny - BC_SIZE + 1, dx, dy, 1, ops, w);
      for (int k = 1; k <= 4; ++k) FY(i, ny - BC_SIZE + 1, k) = w[k - 1];
    }
  if (m->bc_w == ORC_BC_WALL)
    for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j) {
      wall_flux_x(m, nodes, 1 + BC_SIZE, j, dx, dy, 0, ops, w);
      for (int k = 1; k <= 4; ++k) FX(1 + BC_SIZE, j, k) = w[k - 1];
    }
  if (m->bc_e == ORC_BC_WALL)
    for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j) {
      wall_flux_x(m, nodes, nx - BC_SIZE + 1, j, dx, dy, 1, ops, w);
      for (int k = 1; k <= 4; ++k) FX(nx - BC_SIZE + 1, j, k) = w[k - 1];
    }

  /* divergence, td_operators_class.f:181-194 */
  for (int k = 1; k <= 4; ++k)
    for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j)
      for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
        double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
        ND(time_dev, i, j, k) = (FX(i, j, k) / dx - FX(i + 1, j, k) / dx) +
                                (FY(i, j, k) / dy - FY(i, j + 1, k) / dy) + body_force(m, v, k);
      }

  /* apply_bc_on_timedev of hedstrom_xy on every edge
   * (bc_operators_hedstrom_xy_class.f:182-496): sections of
   * bc_operators_gen_class.f:95-102, edges with one-sided fluxes
   * (bc_operators_openbc_normal_class.f:563-787) and the characteristic term
   * (hedstrom_xy_module.f:175-400), corners with both characteristic terms */
  if (m->bc_n == ORC_BC_OPEN) {
    const int B = BC_SIZE;
    double f[4], o[4];
    /* N and S edges: x-fluxes of the ghost rows, i in [B+1, nx-B+1] */
    for (int e = 0; e < 2; ++e) {
      const int j0 = e == 0 ? ny - B + 1 : 1;
      for (int jj = 0; jj < 2; ++jj) {
        const int j = j0 + jj;
        const sd_ops* S = e == 0 ? (jj == 0 ? &MATTSSON : &Y_R0) : (jj == 0 ? &Y_L0 : &MATTSSON);
        for (int i = B + 1; i <= nx - B + 1; ++i) {
          flux_x_ops(m, nodes, i, j, dx, dy, S, f);
          for (int k = 1; k <= 4; ++k) FX(i, j, k) = f[k - 1];
        }
        for (int i = B + 1; i <= nx - B; ++i) {
          double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
          orc_openbc_timedev(m, nodes, i, j, 1, e == 0, e == 0, dy, NULL, o);
          for (int k = 1; k <= 4; ++k)
            ND(time_dev, i, j, k) = 1.0 / dx * (FX(i, j, k) - FX(i + 1, j, k)) + o[k - 1] +
                                    body_force(m, v, k);
        }
      }
    }
    /* E and W edges: y-fluxes of the ghost columns, j in [B+1, ny-B+1] */
    for (int e = 0; e < 2; ++e) {
      const int i0 = e == 0 ? nx - B + 1 : 1;
      for (int ii = 0; ii < 2; ++ii) {
        const int i = i0 + ii;
        const sd_ops* S = e == 0 ? (ii == 0 ? &MATTSSON : &X_R0) : (ii == 0 ? &X_L0 : &MATTSSON);
        for (int j = B + 1; j <= ny - B + 1; ++j) {
          flux_y_ops(m, nodes, i, j, dx, dy, S, f);
          for (int k = 1; k <= 4; ++k) FY(i, j, k) = f[k - 1];
        }
        for (int j = B + 1; j <= ny - B; ++j) {
          double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
          orc_openbc_timedev(m, nodes, i, j, 0, e == 0, e == 0, dx, NULL, o);
          for (int k = 1; k <= 4; ++k)
            ND(time_dev, i, j, k) = o[k - 1] + 1.0 / dy * (FY(i, j, k) - FY(i, j + 1, k)) +
                                    body_force(m, v, k);
        }
      }
    }
    /* corners: x- and y-characteristic terms */
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx)
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            const int i = (cx ? nx - B + 1 : 1) + ii, j = (cy ? ny - B + 1 : 1) + jj;
            double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
            double ox[4], oy[4];
            orc_openbc_timedev(m, nodes, i, j, 0, cx, cx, dx, NULL, ox);
            orc_openbc_timedev(m, nodes, i, j, 1, cy, cy, dy, NULL, oy);
            for (int k = 1; k <= 4; ++k)
              ND(time_dev, i, j, k) = ox[k - 1] + oy[k - 1] + body_force(m, v, k);
          }
  }
#undef FX
#undef FY
  free(flux_x);
  free(flux_y);
}

/* ------------------------------------------------------------------ */
/* ghost cells                                                          */
/* ------------------------------------------------------------------ */

/* reflection_x (bc_operators_reflection_x_class.f:178-296) over a column range
 * [i0, i0+1] and a row range [j0, j1]; west mirrors about the W face. */
static void reflect_x(const orc_model* m, double* n, int west, int i0, int j0, int j1) {
  const int refl = west ? 2 * i0 + 2 * BC_SIZE - 1 : 2 * i0 - 1;
  for (int k = 1; k <= 4; ++k) {
    const double pf = (k == 2) ? -1.0 : 1.0;
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i0 + BC_SIZE - 1; ++i) ND(n, i, j, k) = pf * ND(n, refl - i, j, k);
  }
}
/* reflection_y (bc_operators_reflection_y_class.f:178-291) */
static void reflect_y(const orc_model* m, double* n, int south, int j0, int i0, int i1) {
  const int refl = south ? 2 * j0 + 2 * BC_SIZE - 1 : 2 * j0 - 1;
  for (int k = 1; k <= 4; ++k) {
    const double pf = (k == 3) ? -1.0 : 1.0;
    for (int j = j0; j <= j0 + BC_SIZE - 1; ++j)
      for (int i = i0; i <= i1; ++i) ND(n, i, j, k) = pf * ND(n, i, refl - j, k);
  }
}

/* periodic x over rows [j0, j1] (periodic_xy/bc_operators_class.f:156-175) */
static void periodic_x(const orc_model* m, double* n, int j0, int j1) {
  const int px = m->nx - 2 * BC_SIZE;
  for (int k = 1; k <= 4; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = 1; i <= BC_SIZE; ++i) {
        ND(n, i, j, k) = ND(n, i + px, j, k);
        ND(n, i + px + BC_SIZE, j, k) = ND(n, i + BC_SIZE, j, k);
      }
}
/* periodic y over columns [i0, i1] (:176-185) */
static void periodic_y(const orc_model* m, double* n, int i0, int i1) {
  const int py = m->ny - 2 * BC_SIZE;
  for (int k = 1; k <= 4; ++k)
    for (int j = 1; j <= BC_SIZE; ++j)
      for (int i = i0; i <= i1; ++i) {
        ND(n, i, j, k) = ND(n, i, j + py, k);
        ND(n, i, j + py + BC_SIZE, k) = ND(n, i, j + BC_SIZE, k);
      }
}

/* ini_debug_nodes, bc_operators_wall_xy_class.f:951-975 */
static void debug_fill(const orc_model* m, double* n, int i0, int i1, int j0, int j1) {
  for (int k = 1; k <= 4; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) ND(n, i, j, k) = DEBUG_REAL;
}

/* wall edges over the interior columns (S/N) or rows (W/E), apply_bc_on_nodes
 * of bc_operators_wall_xy_class.f:232-330 */
static void wall_edge(const orc_model* m, double* n, const double* ntmp, int edge, double dx,
                      double dy) {
  const int nx = m->nx, ny = m->ny;
  double Tg, mg;
  if (edge == 0) { /* S: inner ghost row j = bc_size */
    debug_fill(m, n, 1 + BC_SIZE, nx - BC_SIZE, 1, BC_SIZE);
    const int j = BC_SIZE;
    for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
      wall_guess(m, ntmp, i, j + 1, i, j, dx, dy, &Tg, &mg);
      wall_y_ghost(m, n, i, j, 0, dx, dy, Tg, mg);
    }
  } else if (edge == 1) { /* N: inner ghost row j = ny-bc_size+1 */
    debug_fill(m, n, 1 + BC_SIZE, nx - BC_SIZE, ny - BC_SIZE + 1, ny);
    const int j = ny - BC_SIZE + 1;
    for (int i = 1 + BC_SIZE; i <= nx - BC_SIZE; ++i) {
      wall_guess(m, ntmp, i, j - 1, i, j, dx, dy, &Tg, &mg);
      wall_y_ghost(m, n, i, j, 1, dx, dy, Tg, mg);
    }
  } else if (edge == 2) { /* W: inner ghost column i = bc_size */
    debug_fill(m, n, 1, BC_SIZE, 1 + BC_SIZE, ny - BC_SIZE);
    const int i = BC_SIZE;
    for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j) {
      wall_guess(m, ntmp, i + 1, j, i, j, dx, dy, &Tg, &mg);
      wall_x_ghost(m, n, i, j, 0, dx, dy, Tg, mg);
    }
  } else { /* E: inner ghost column i = nx-bc_size+1 */
    debug_fill(m, n, nx - BC_SIZE + 1, nx, 1 + BC_SIZE, ny - BC_SIZE);
    const int i = nx - BC_SIZE + 1;
    for (int j = 1 + BC_SIZE; j <= ny - BC_SIZE; ++j) {
      wall_guess(m, ntmp, i - 1, j, i, j, dx, dy, &Tg, &mg);
      wall_x_ghost(m, n, i, j, 1, dx, dy, Tg, mg);
    }
  }
}

/* apply_bc_on_nodes, bc_operators_gen_class.f:238-381, for the supported
 * configurations:
 *  - reflection on all edges: library_config.py reflection_xy_choice, corners
 *    reflection_y, order W,E,SW,S,SE,NW,N,NE;
 *  - periodic on all edges: periodic_xy/bc_operators_class.f:156-210, x then y;
 *  - wall on S (and/or N) with reflection elsewhere: wall_S_reflection_choice,
 *    corners reflection_x, order W,E,S,N,SW,SE,NW,NE;
 *  - periodic in x with walls in y (no counterpart in the reference's active
 *    dispatcher): x-periodic fill of the interior rows, the wall rows, then an
 *    x-periodic fill of the ghost rows, so corners are periodic images of the
 *    wall ghost cells;
 *  - walls on W and/or E with reflection or periodic edges in y: the
 *    transposes of the two cases above. */
void orc_apply_bc_on_nodes(const orc_model* m, double t, const double* x_map, const double* y_map,
                           const double* nodes_tmp, double* nodes) {
  (void)t;
  const int nx = m->nx, ny = m->ny;
  const double dx = x_map[1] - x_map[0];
  const double dy = y_map[1] - y_map[0];
  const int B = BC_SIZE;
  if (m->bc_n == ORC_BC_OPEN) return; /* hedstrom_xy: the ghost cells are integrated */
  const int allrefl = m->bc_w == ORC_BC_REFLECTION && m->bc_e == ORC_BC_REFLECTION &&
                      m->bc_s == ORC_BC_REFLECTION && m->bc_n == ORC_BC_REFLECTION;
  if (allrefl) {
    if (ORC_PLANT == 6) { /* planted: corners from the W/E ghosts of the last fill */
      reflect_y(m, nodes, 1, 1, 1, B);
      reflect_y(m, nodes, 1, 1, nx - B + 1, nx);
      reflect_y(m, nodes, 0, ny - B + 1, 1, B);
      reflect_y(m, nodes, 0, ny - B + 1, nx - B + 1, nx);
    }
    reflect_x(m, nodes, 1, 1, B + 1, ny - B);              /* W */
    reflect_x(m, nodes, 0, nx - B + 1, B + 1, ny - B);     /* E */
    if (ORC_PLANT == 6) {
      reflect_y(m, nodes, 1, 1, B + 1, nx - B);
      reflect_y(m, nodes, 0, ny - B + 1, B + 1, nx - B);
      return;
    }
    reflect_y(m, nodes, 1, 1, 1, B);                       /* SW (reflection_y) */
    reflect_y(m, nodes, 1, 1, B + 1, nx - B);              /* S */
    reflect_y(m, nodes, 1, 1, nx - B + 1, nx);             /* SE */
    reflect_y(m, nodes, 0, ny - B + 1, 1, B);              /* NW */
    reflect_y(m, nodes, 0, ny - B + 1, B + 1, nx - B);     /* N */
    reflect_y(m, nodes, 0, ny - B + 1, nx - B + 1, nx);    /* NE */
    return;
  }
  if (m->bc_w == ORC_BC_PERIODIC && m->bc_s == ORC_BC_PERIODIC) {
    periodic_x(m, nodes, 1 + B, ny - B);
    if (ORC_PLANT == 9) periodic_y(m, nodes, 1 + B, nx - B); /* planted */
    else periodic_y(m, nodes, 1, nx);
    return;
  }
  if (m->bc_w == ORC_BC_PERIODIC) { /* periodic x, walls (or reflection) in y */
    if (ORC_PLANT != 4) periodic_x(m, nodes, 1 + B, ny - B);
    if (m->bc_s == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 0, dx, dy);
    else reflect_y(m, nodes, 1, 1, B + 1, nx - B);
    if (m->bc_n == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 1, dx, dy);
    else reflect_y(m, nodes, 0, ny - B + 1, B + 1, nx - B);
    if (ORC_PLANT == 4) periodic_x(m, nodes, 1 + B, ny - B); /* planted: after the walls */
    if (ORC_PLANT == 5) return;                              /* planted: no ghost-row fill */
    periodic_x(m, nodes, 1, B);
    periodic_x(m, nodes, ny - B + 1, ny);
    return;
  }
  if (m->bc_w == ORC_BC_WALL || m->bc_e == ORC_BC_WALL) {
    /* walls on W and/or E: the W/E sections of the wall class
     * (bc_operators_wall_xy_class.f:232-330, compute_wall_{W,E}_ghost_cell of
     * wall_xy_equilibrium_module.f:1211-1437), which no choice of
     * library_config.py selects; composed as the transpose of the S/N cases */
    if (m->bc_s == ORC_BC_PERIODIC) {
      periodic_y(m, nodes, 1 + B, nx - B);
      if (m->bc_w == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 2, dx, dy);
      else reflect_x(m, nodes, 1, 1, B + 1, ny - B);
      if (m->bc_e == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 3, dx, dy);
      else reflect_x(m, nodes, 0, nx - B + 1, B + 1, ny - B);
      if (ORC_PLANT == 8) return; /* planted: no ghost-column fill */
      periodic_y(m, nodes, 1, B);
      periodic_y(m, nodes, nx - B + 1, nx);
      return;
    }
    reflect_y(m, nodes, 1, 1, B + 1, nx - B);          /* S */
    reflect_y(m, nodes, 0, ny - B + 1, B + 1, nx - B); /* N */
    if (m->bc_w == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 2, dx, dy);
    else reflect_x(m, nodes, 1, 1, B + 1, ny - B);
    if (m->bc_e == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 3, dx, dy);
    else reflect_x(m, nodes, 0, nx - B + 1, B + 1, ny - B);
    reflect_y(m, nodes, 1, 1, 1, B);                   /* SW */
    reflect_y(m, nodes, 1, 1, nx - B + 1, nx);         /* SE */
    reflect_y(m, nodes, 0, ny - B + 1, 1, B);          /* NW */
    reflect_y(m, nodes, 0, ny - B + 1, nx - B + 1, nx); /* NE */
    return;
  }
  /* reflection in x; S and N are wall or reflection; corners reflection_x */
  reflect_x(m, nodes, 1, 1, B + 1, ny - B);          /* W */
  reflect_x(m, nodes, 0, nx - B + 1, B + 1, ny - B); /* E */
  if (m->bc_s == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 0, dx, dy);
  else reflect_y(m, nodes, 1, 1, B + 1, nx - B);
  if (m->bc_n == ORC_BC_WALL) wall_edge(m, nodes, nodes_tmp, 1, dx, dy);
  else reflect_y(m, nodes, 0, ny - B + 1, B + 1, nx - B);
  reflect_x(m, nodes, 1, 1, 1, B);                   /* SW */
  reflect_x(m, nodes, 0, nx - B + 1, 1, B);          /* SE */
  reflect_x(m, nodes, 1, 1, ny - B + 1, ny);         /* NW */
  reflect_x(m, nodes, 0, nx - B + 1, ny - B + 1, ny); /* NE */
}

/* ------------------------------------------------------------------ */
/* RK3-TVD: td_integrator_class.f:103-161, rk3tvd_steps_module.f        */
/* ------------------------------------------------------------------ */
void orc_rk3_step(const orc_model* m, double t, double dt, const double* x_map,
                  const double* y_map, double* nodes, double* work) {
  const int nx = m->nx, ny = m->ny;
  const size_t plane = (size_t)nx * ny;
  double* nodes_tmp = work;
  double* time_dev = work + plane * 4;
  const int B = BC_SIZE;
  const double b2 = 0.75;
  const double b3 = ORC_PLANT == 3 ? 0.25 : 1.0 / 3.0;
  /* the state stages 2 and 3 blend with (u^n) and differentiate */
  const double* blend = ORC_PLANT == 2 ? nodes : nodes_tmp;
  const double* dsrc = ORC_PLANT == 7 ? nodes_tmp : nodes;

  /* stage 1 (compute_1st_step :52-100) over the whole array */
  orc_compute_time_dev(m, t, nodes, x_map, y_map, time_dev);
  for (size_t q = 0; q < plane * 4; ++q) {
    nodes_tmp[q] = nodes[q];
    nodes[q] = nodes[q] + dt * time_dev[q];
  }
  if (ORC_PLANT != 1) orc_apply_bc_on_nodes(m, t, x_map, y_map, nodes_tmp, nodes);

  /* x/y_borders (field_abstract_class.f:244-270): widened to the ghost cells
   * on edges whose time derivative is computed (open boundaries) */
  const int op = m->bc_n == ORC_BC_OPEN;
  const int i0 = op ? 1 : B + 1, i1 = op ? nx : nx - B;
  const int j0 = op ? 1 : B + 1, j1 = op ? ny : ny - B;

  /* stage 2 (compute_2nd_step :205-269) over x/y_borders */
  orc_compute_time_dev(m, t, dsrc, x_map, y_map, time_dev);
  for (int k = 1; k <= 4; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i)
        ND(nodes, i, j, k) =
            b2 * ND(blend, i, j, k) + (1.0 - b2) * (ND(nodes, i, j, k) + dt * ND(time_dev, i, j, k));
  orc_apply_bc_on_nodes(m, t, x_map, y_map, nodes_tmp, nodes);

  /* stage 3 (compute_3rd_step :424-490) */
  orc_compute_time_dev(m, t, dsrc, x_map, y_map, time_dev);
  for (int k = 1; k <= 4; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i)
        ND(nodes, i, j, k) =
            b3 * ND(blend, i, j, k) + (1.0 - b3) * (ND(nodes, i, j, k) + dt * ND(time_dev, i, j, k));
  orc_apply_bc_on_nodes(m, t, x_map, y_map, nodes_tmp, nodes);
}

/* Low-storage RK3, Williamson (1980) 2N-storage scheme (coefficients from
 * that paper's third-order case).  Not in the reference; the stage recurrence
 * is shared by the field step and the scalar amplification factor below. */
static const double LS_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double LS_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

static inline void lsrk_point(int k, double dtL, double* q, double* u) {
  *q = LS_A[k] * *q + dtL;
  *u = *u + LS_B[k] * *q;
}

double orc_lsrk3_amplification(double z) {
  double u = 1.0, q = 0.0;
  for (int k = 0; k < 3; ++k) lsrk_point(k, z * u, &q, &u);
  return u;
}

void orc_lsrk3_step(const orc_model* m, double t, double dt, const double* x_map,
                    const double* y_map, double* nodes, double* work) {
  const int nx = m->nx, ny = m->ny;
  const size_t plane = (size_t)nx * ny;
  double* q = work;
  double* time_dev = work + plane * 4;
  double* guess = work + plane * 8;
  const int B = BC_SIZE;
  const int op = m->bc_n == ORC_BC_OPEN; /* borders as in orc_rk3_step */
  const int i0 = op ? 1 : B + 1, i1 = op ? nx : nx - B;
  const int j0 = op ? 1 : B + 1, j1 = op ? ny : ny - B;
  for (size_t i = 0; i < plane * 4; ++i) q[i] = 0.0;
  for (int s = 0; s < 3; ++s) {
    memcpy(guess, nodes, plane * 4 * sizeof(double));
    orc_compute_time_dev(m, t, nodes, x_map, y_map, time_dev);
    for (int k = 1; k <= 4; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i)
          lsrk_point(s, dt * ND(time_dev, i, j, k), &ND(q, i, j, k), &ND(nodes, i, j, k));
    orc_apply_bc_on_nodes(m, t, x_map, y_map, guess, nodes);
  }
}

int orc_run(const orc_model* m, double t0, int nsteps, const double* dts, const double* x_map,
            const double* y_map, double* nodes) {
  const size_t plane = (size_t)m->nx * m->ny;
  double* work = (double*)malloc(plane * 12 * sizeof(double));
  double t = t0;
  int bad = 0;
  for (int s = 0; s < nsteps && !bad; ++s) {
    if (m->integrator == ORC_INT_LSRK3)
      orc_lsrk3_step(m, t, dts[s], x_map, y_map, nodes, work);
    else
      orc_rk3_step(m, t, dts[s], x_map, y_map, nodes, work);
    t += dts[s];
    for (int k = 1; k <= 4 && !bad; ++k)
      for (int j = 1 + BC_SIZE; j <= m->ny - BC_SIZE && !bad; ++j)
        for (int i = 1 + BC_SIZE; i <= m->nx - BC_SIZE; ++i)
          if (!isfinite(ND(nodes, i, j, k))) {
            bad = s + 1;
            break;
          }
  }
  free(work);
  return bad;
}

/* ini_coordinates, field_abstract_class.f:286-315 */
void orc_coordinates(int nx, int ny, double x_min, double x_max, double y_min, double y_max,
                     double* x_map, double* y_map) {
  const double dx = (x_max - x_min) / (nx - 1 - 2 * BC_SIZE);
  const double dy = (y_max - y_min) / (ny - 1 - 2 * BC_SIZE);
  for (int i = 1; i <= nx; ++i) x_map[i - 1] = x_min + (i - 1 - BC_SIZE) * dx;
  for (int j = 1; j <= ny; ++j) y_map[j - 1] = y_min + (j - 1 - BC_SIZE) * dy;
}

/* sound speed of the time step rule: the reference's speed_of_sound
 * (dim2d_prim_module.f:1179-1215) reduces to
 * c^2 = 3 (e + (P + 3 rho^2)/cv) / (rho (3 - rho)), e = P + rho^2 (2 rho - 3);
 * where c^2 <= 0 (inside the spinodal region of the interface) c is taken as 0. */
static double cfl_sound_speed(const orc_model* m, const double v[4]) {
  const double P = classical_pressure_local(m, v);
  const double e = P + sq(v[0]) * (-3.0 + 2.0 * v[0]);
  const double c2 = 3.0 * (e + (P + 3.0 * sq(v[0])) / m->cv_r) / (v[0] * (3.0 - v[0]));
  return c2 > 0.0 ? sqrt(c2) : 0.0;
}

void orc_cfl_dt(const orc_model* m, const double* nodes, double dx, double dy, double cfl,
                double out[4]) {
  double amax = 0.0, rmin = INFINITY, rmax = -INFINITY;
  for (int j = 1 + BC_SIZE; j <= m->ny - BC_SIZE; ++j)
    for (int i = 1 + BC_SIZE; i <= m->nx - BC_SIZE; ++i) {
      const double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
      const double u = v[1] / v[0], w = v[2] / v[0];
      const double a = sqrt(u * u + w * w) + cfl_sound_speed(m, v);
      if (a > amax) amax = a;
      if (v[0] < rmin) rmin = v[0];
      if (v[0] > rmax) rmax = v[0];
    }
  const double h = dx < dy ? dx : dy;
  double dt = h / amax;
  const double dv = m->re * rmin * h * h;
  const double dc = sqrt(m->we / rmax) * h * h;
  if (dv < dt) dt = dv;
  if (dc < dt) dt = dc;
  out[0] = cfl * dt;
  out[1] = amax;
  out[2] = rmin;
  out[3] = rmax;
}

/* are_openbc_undermined (pmodel_eq_class.f:1991-2070) with its default
 * criterion (bf_openbc_md_threshold_ac = .false., parameters_input.f:833):
 * the speed of sound of the point is imaginary, P + rho^2 (-3 + 2 rho) < 0,
 * with P = classical_pressure_local. */
int orc_openbc_undermined_point(const orc_model* m, const double v[4]) {
  const double P = classical_pressure_local(m, v);
  const double coeff = P + sq(v[0]) * (-3.0 + 2.0 * v[0]);
  return coeff < 0.0;
}

/* are_openbc_undermined at the interior point (i, j) (1-based) of a field:
 * the default criterion, then, when md_thr >= 0 (bf_openbc_md_threshold_ac
 * with bf_openbc_md_threshold = md_thr, parameters_input.f:833-846), a density
 * inside [md_vap_thr, md_liq_thr] at T = 3/(8 cv) temperature_eff
 * (dim2d_prim_module.f:652-691) with central density gradients
 * (gradient_x/y_interior) and the saturation fits of dim2d_state_eq_module.f:145-227 */
int orc_openbc_undermined_at(const orc_model* m, const double* nodes, int i, int j, double dx,
                             double dy, double md_thr) {
  const double v[4] = {ND(nodes, i, j, 1), ND(nodes, i, j, 2), ND(nodes, i, j, 3), ND(nodes, i, j, 4)};
  if (orc_openbc_undermined_point(m, v)) return 1;
  if (md_thr < 0.0) return 0;
  const double gx = (ND(nodes, i + 1, j, 1) - ND(nodes, i - 1, j, 1)) / (2.0 * dx);
  const double gy = (ND(nodes, i, j + 1, 1) - ND(nodes, i, j - 1, 1)) / (2.0 * dy);
  const double teff = 1.0 / v[0] *
                      (v[3] - 0.5 * v[0] * (sq(v[1] / v[0]) + sq(v[2] / v[0])) -
                       0.5 / m->we * (sq(gx) + sq(gy)) + 3.0 * sq(v[0]));
  const double T = 3.0 / (8.0 * m->cv_r) * teff;
  const double md_liq = orc_mass_density_liquid(T);
  const double md_vap = orc_mass_density_vapor(T);
  const double md_mid = 0.5 * (md_vap + md_liq);
  const double md_vap_thr = md_vap + md_thr * (md_mid - md_vap);
  const double md_liq_thr = md_liq - md_thr * (md_liq - md_mid);
  return v[0] >= md_vap_thr && v[0] <= md_liq_thr;
}

/* undermined points among the `dist` interior columns nearest the W and E
 * edges and the `dist` interior rows nearest the S and N edges
 * (out[0..3] = W, E, S, N; a corner point counts for both of its edges) */
void orc_openbc_undermined(const orc_model* m, const double* nodes, int dist, double dx, double dy,
                           double md_thr, int out[4]) {
  const int i0 = 1 + BC_SIZE, i1 = m->nx - BC_SIZE, j0 = 1 + BC_SIZE, j1 = m->ny - BC_SIZE;
  for (int k = 0; k < 4; ++k) out[k] = 0;
  for (int j = j0; j <= j1; ++j)
    for (int i = i0; i <= i1; ++i) {
      const int near = (i - i0 < dist) || (i1 - i < dist) || (j - j0 < dist) || (j1 - j < dist);
      if (!near || !orc_openbc_undermined_at(m, nodes, i, j, dx, dy, md_thr)) continue;
      if (i - i0 < dist) out[0] += 1;
      if (i1 - i < dist) out[1] += 1;
      if (j - j0 < dist) out[2] += 1;
      if (j1 - j < dist) out[3] += 1;
    }
}

void orc_sums(const orc_model* m, const double* nodes, double out[4]) {
  for (int k = 1; k <= 4; ++k) {
    double s = 0.0;
    for (int j = 1 + BC_SIZE; j <= m->ny - BC_SIZE; ++j)
      for (int i = 1 + BC_SIZE; i <= m->nx - BC_SIZE; ++i) s += ND(nodes, i, j, k);
    out[k - 1] = s;
  }
}

/* ------------------------------------------------------------------ */
/* new grid points of the buffer layers: the characteristic            */
/* extrapolation of bf_layer_newgrdpt/bf_newgrdpt_class.f              */
/* ------------------------------------------------------------------ */

#define GN(g, i, j, k) \
  (g)->nodes[((size_t)((k) - 1) * (size_t)(g)->ny + (size_t)((j) - 1)) * (size_t)(g)->nx + \
             (size_t)((i) - 1)]

/* compute_prim_var_ns_vdw2d, ns_vdw2d_prim_module.f:51-64 */
static void prim_var(const orc_model* m, const double v[4], double p[4]) {
  p[0] = v[0];
  p[1] = v[1] / v[0];
  p[2] = v[2] / v[0];
  p[3] = classical_pressure_local(m, v);
}

/* compute_cons_var_ns_vdw2d, ns_vdw2d_prim_module.f:83-108 */
static void cons_var(const orc_model* m, const double p[4], double v[4]) {
  const double T = (p[3] + 3.0 * sq(p[0])) * (3.0 - p[0]) / (8.0 * p[0]);
  v[0] = p[0];
  v[1] = p[0] * p[1];
  v[2] = p[0] * p[2];
  v[3] = 0.5 * p[0] * (sq(p[1]) + sq(p[2])) + p[0] * (8.0 / 3.0 * m->cv_r * T - 3.0 * p[0]);
}

/* compute_body_forces(..., prim=.true.), pmodel_eq_class.f:1840-1918 */
static void body_forces_prim(const orc_model* m, double b[4]) {
  b[0] = 0.0;
  b[1] = 0.0;
  b[2] = m->gravity_ac ? -m->gravity : 0.0;
  b[3] = 0.0;
}

/* get_prim_obc_eigenqties, pmodel_eq_class.f:2142-2195: [rho, ux, uy, P, c]
 * of the point (obc_eigenqties_bc) or of the far field (obc_eigenqties_lin) */
static void obc_eigenqties(const orc_model* m, const double ff[4], int eigen_lin,
                           const double v[4], double pe[5]) {
  obc_eigenqties_bc(m, eigen_lin ? ff : v, pe);
}

/* n_coords_module.f: the diagonal coordinates.  get_x_coord and get_y_coord
 * are restated as written there (sqrt(2)/4 (n1 +- n2)); they only place the
 * far-field evaluation, which is uniform here. */
static double n1_coord(double x, double y) { return sqrt(2.0) * 0.5 * (x - y); }
static double n2_coord(double x, double y) { return sqrt(2.0) * 0.5 * (x + y); }

/* compute_xy_to_n_var / compute_n_to_xy_var, pmodel_eq_class.f:3173-3231 */
static void xy_to_n(const double a[4], double b[4]) {
  const double a1 = a[1], a2 = a[2];
  b[0] = a[0];
  b[1] = 0.5 * sqrt(2.0) * (a1 - a2);
  b[2] = 0.5 * sqrt(2.0) * (a1 + a2);
  b[3] = a[3];
}
static void n_to_xy(const double a[4], double b[4]) {
  const double a1 = a[1], a2 = a[2];
  b[0] = a[0];
  b[1] = 0.5 * sqrt(2.0) * (a1 + a2);
  b[2] = 0.5 * sqrt(2.0) * (a2 - a1);
  b[3] = a[3];
}

/* compute_x_transM_ns_vdw2d(md, uy, c) and compute_y_transM_ns_vdw2d(md, ux,
 * c), ns_vdw2d_prim_module.f:345-482, applied as MATMUL(gradient, transM) */
static void trans_term(int ydir, const double pe[5], const double g[4], double out[4]) {
  const double md = pe[0], c = pe[4];
  if (!ydir) {
    const double uy = pe[2];
    out[0] = g[0] * uy + g[2] * md;
    out[1] = g[1] * uy;
    out[2] = g[2] * uy + g[3] * (1.0 / md);
    out[3] = g[2] * (c * c * md) + g[3] * uy;
  } else {
    const double ux = pe[1];
    out[0] = g[0] * ux + g[1] * md;
    out[1] = g[1] * ux + g[3] * (1.0 / md);
    out[2] = g[2] * ux;
    out[3] = g[1] * (c * c * md) + g[3] * ux;
  }
}

/* the gradient procedures the new-point procedures pass
 * (sd_operators_fd_module.f: gradient_{x,y}_interior and the one-sided
 * gradient_x_x_oneside_{L0,R0}, gradient_y_y_oneside_{L0,R0}) */
static grad_fn gradient_proc(int ydir, int type) {
  if (!ydir)
    return type == ORC_GRAD_I ? gradient_x_interior
                              : (type == ORC_GRAD_L0 ? gradient_x_x_oneside_L0 : gradient_x_x_oneside_R0);
  return type == ORC_GRAD_I ? gradient_y_interior
                            : (type == ORC_GRAD_L0 ? gradient_y_y_oneside_L0 : gradient_y_y_oneside_R0);
}

/* compute_gradient_prim, pmodel_eq_class.f:3113-3153 */
static void gradient_prim(const orc_model* m, const orc_bf_grid* g, int i, int j, grad_fn gr,
                          double dn, double out[4]) {
  orc_model mg = *m;
  mg.nx = g->nx;
  mg.ny = g->ny;
  out[0] = gr(&mg, g->nodes, i, j, orc_mass_density, dn);
  out[1] = gr(&mg, g->nodes, i, j, orc_velocity_x, dn);
  out[2] = gr(&mg, g->nodes, i, j, orc_velocity_y, dn);
  out[3] = gr(&mg, g->nodes, i, j, orc_classical_pressure, dn);
}

static void grid_point(const orc_bf_grid* g, int i, int j, double v[4]) {
  for (int k = 1; k <= 4; ++k) v[k - 1] = GN(g, i, j, k);
}

/* get_interpolation_coeff_1D, bf_newgrdpt_class.f:2560-2590: per variable,
 * the slope and intercept of the line through (x_map(1), nodes(1,k)) and
 * (x_map(2), nodes(2,k)) */
void orc_interp_coeff_1d(const double x_map[2], const double nodes[8], double coeff[8]) {
  for (int k = 0; k < 4; ++k) {
    coeff[k] = (nodes[4 + k] - nodes[k]) / (x_map[1] - x_map[0]);
    coeff[4 + k] = nodes[k] - coeff[k] * x_map[0];
  }
}

/* interpolate_1D, :2592-2612 */
void orc_interpolate_1d(double x, const double coeff[8], double out[4]) {
  for (int k = 0; k < 4; ++k) out[k] = x * coeff[k] + coeff[4 + k];
}

/* get_interpolation_coeff_2D with get_plane_normal_vector and cross_product,
 * :2665-2858: the plane through the three points (x(l), y(l), nodes(l,k)) as
 * z = c1 x + c2 y + c3 from the normal n = AB x AC */
void orc_interp_coeff_2d(const double x_map[3], const double y_map[3], const double nodes[12],
                         double coeff[12]) {
  for (int k = 0; k < 4; ++k) {
    const double A[3] = {x_map[0], y_map[0], nodes[k]};
    const double B[3] = {x_map[1], y_map[1], nodes[4 + k]};
    const double C[3] = {x_map[2], y_map[2], nodes[8 + k]};
    double AB[3], AC[3];
    for (int q = 0; q < 3; ++q) {
      AB[q] = B[q] - A[q];
      AC[q] = C[q] - A[q];
    }
    const double n0 = AB[1] * AC[2] - AB[2] * AC[1];
    const double n1 = -AB[0] * AC[2] + AB[2] * AC[0];
    const double n2 = AB[0] * AC[1] - AB[1] * AC[0];
    coeff[k] = -n0 / n2;
    coeff[4 + k] = -n1 / n2;
    coeff[8 + k] = -A[0] * coeff[k] - A[1] * coeff[4 + k] + A[2];
  }
}

/* interpolate_2D, :2767-2790 */
void orc_interpolate_2d(double x, double y, const double coeff[12], double out[4]) {
  for (int k = 0; k < 4; ++k) out[k] = x * coeff[k] + y * coeff[4 + k] + coeff[8 + k];
}

/* compute_NewtonCotes_integration, :2858-2875: trapezoidal rule over dt */
void orc_newton_cotes(const double d0[4], const double d1[4], double dt, double out[4]) {
  for (int k = 0; k < 4; ++k) out[k] = (d0[k] + 0.5 * (d1[k] - d0[k])) * dt;
}

/* the far-field state in the primitive variables and its body forces, for
 * an incoming characteristic (the else branches of :1300-1320) */
static void far_field_amp(const orc_model* m, const double ff[4], int rotate, double n_amp[4],
                          double t_amp[4], double b_amp[4]) {
  double p[4], b[4];
  prim_var(m, ff, p);
  body_forces_prim(m, b);
  if (rotate) {
    xy_to_n(p, n_amp);
    xy_to_n(b, b_amp);
  } else {
    for (int k = 0; k < 4; ++k) {
      n_amp[k] = p[k];
      b_amp[k] = b[k];
    }
  }
  for (int k = 0; k < 4; ++k) t_amp[k] = 0.0;
}

/* compute_newgrdpt_x (ydir 0), :1029-1361, and compute_newgrdpt_y (ydir 1),
 * :1417-1746, in one body: the direction swaps the roles of the indices, the
 * coordinate maps, the transverse gradient and the eigen-decomposition. */
static void newgrdpt_dir(const orc_model* m, const double ff[4], int eigen_lin, double dt,
                         const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int ydir,
                         int side_right, int grad_type, double out[4]) {
  const int a_n = ydir ? g1->align_y - g0->align_y : g1->align_x - g0->align_x;
  const int a_t = ydir ? g1->align_x - g0->align_x : g1->align_y - g0->align_y;
  const int p1 = ydir ? j1 : i1;  /* index along the direction */
  const int q1 = ydir ? i1 : j1;  /* transverse index */
  int p_eigen, p0a, p0b, p1a, p1b;
  if (side_right) {
    p_eigen = p1 - 1;
    p0a = a_n + p_eigen - 1;
    p0b = a_n + p_eigen;
    p1a = p_eigen - 1;
    p1b = p_eigen;
  } else {
    p_eigen = p1 + 1;
    p0a = a_n + p_eigen;
    p0b = a_n + p_eigen + 1;
    p1a = p_eigen;
    p1b = p_eigen + 1;
  }
  const int q0 = a_t + q1;
  const double* map0 = ydir ? g0->y_map : g0->x_map;
  const double* map1 = ydir ? g1->y_map : g1->x_map;
  const double* tmap1 = ydir ? g1->x_map : g1->y_map;
  const double dt_n = tmap1[1] - tmap1[0]; /* spacing of the transverse gradient */
  const grad_fn gr = gradient_proc(!ydir, grad_type);
  const double s1 = map1[p1 - 1];
#define PT(g, pp, qq, v) (ydir ? grid_point(g, qq, pp, v) : grid_point(g, pp, qq, v))
#define GRAD(g, pp, qq, o) \
  (ydir ? gradient_prim(m, g, qq, pp, gr, dt_n, o) : gradient_prim(m, g, pp, qq, gr, dt_n, o))

  double xm[2], pts[8], v[4], pe[5], gv[4];
  double inter_nodes0[8], inter_trans0[8], inter_bf0[8], inter_trans1[8], inter_bf1[8];
  /* time t - dt */
  xm[0] = map0[p0a - 1];
  xm[1] = map0[p0b - 1];
  PT(g0, p0a, q0, v);
  prim_var(m, v, pts);
  PT(g0, p0b, q0, v);
  prim_var(m, v, pts + 4);
  orc_interp_coeff_1d(xm, pts, inter_nodes0);
  for (int l = 0; l < 2; ++l) {
    const int pp = l ? p0b : p0a;
    PT(g0, pp, q0, v);
    obc_eigenqties(m, ff, eigen_lin, v, pe);
    GRAD(g0, pp, q0, gv);
    trans_term(ydir, pe, gv, pts + 4 * l);
  }
  orc_interp_coeff_1d(xm, pts, inter_trans0);
  body_forces_prim(m, pts);
  body_forces_prim(m, pts + 4);
  orc_interp_coeff_1d(xm, pts, inter_bf0);
  /* time t */
  xm[0] = map1[p1a - 1];
  xm[1] = map1[p1b - 1];
  for (int l = 0; l < 2; ++l) {
    const int pp = l ? p1b : p1a;
    PT(g1, pp, q1, v);
    obc_eigenqties(m, ff, eigen_lin, v, pe);
    GRAD(g1, pp, q1, gv);
    trans_term(ydir, pe, gv, pts + 4 * l);
  }
  orc_interp_coeff_1d(xm, pts, inter_trans1);
  body_forces_prim(m, pts);
  body_forces_prim(m, pts + 4);
  orc_interp_coeff_1d(xm, pts, inter_bf1);
  double t_amp1[4], b_amp1[4];
  orc_interpolate_1d(s1, inter_trans1, t_amp1);
  orc_interpolate_1d(s1, inter_bf1, b_amp1);
  /* eigen-decomposition at the neighbour (p_eigen, q1) at time t */
  PT(g1, p_eigen, q1, v);
  obc_eigenqties(m, ff, eigen_lin, v, pe);
  double lam[4], L[4][4], R[4][4];
  ns_vdw2d_eigen(ydir, pe[0], ydir ? pe[2] : pe[1], pe[4], lam, L, R);
  double char_amp[4];
  for (int k = 0; k < 4; ++k) {
    const double s0 = s1 - lam[k] * dt;
    const int inside = side_right ? (lam[k] >= 0) : !(lam[k] > 0);
    double n_amp0[4], t_amp0[4], b_amp0[4];
    if (inside) {
      orc_interpolate_1d(s0, inter_nodes0, n_amp0);
      orc_interpolate_1d(s0, inter_trans0, t_amp0);
      orc_interpolate_1d(s0, inter_bf0, b_amp0);
    } else {
      far_field_amp(m, ff, 0, n_amp0, t_amp0, b_amp0);
    }
    double d0[4], d1[4], integ[4];
    for (int l = 0; l < 4; ++l) {
      d0[l] = t_amp0[l] - b_amp0[l];
      d1[l] = t_amp1[l] - b_amp1[l];
    }
    orc_newton_cotes(d0, d1, dt, integ);
    double acc = 0.0;
    for (int l = 0; l < 4; ++l) acc += (n_amp0[l] - integ[l]) * L[l][k];
    char_amp[k] = acc;
  }
  double prim[4];
  for (int col = 0; col < 4; ++col) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += char_amp[k] * R[k][col];
    prim[col] = acc;
  }
  cons_var(m, prim, out);
#undef PT
#undef GRAD
}

void orc_newgrdpt_x(const orc_model* m, const double ff[4], int eigen_lin, double dt,
                    const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int side_right,
                    int grad_y, double out[4]) {
  newgrdpt_dir(m, ff, eigen_lin, dt, g0, g1, i1, j1, 0, side_right, grad_y, out);
}

void orc_newgrdpt_y(const orc_model* m, const double ff[4], int eigen_lin, double dt,
                    const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int side_right,
                    int grad_x, double out[4]) {
  newgrdpt_dir(m, ff, eigen_lin, dt, g0, g1, i1, j1, 1, side_right, grad_x, out);
}

/* get_n_transverse_data_for_interpolation, :2402-2556: at each of the three
 * points, the gradient across the direction n (the n2-derivative for n1,
 * the n1-derivative for n2) of the primitive variables, its velocity turned
 * into (n1, n2) components, times the transverse matrix of the point's
 * eigen-quantities in (n1, n2) components */
static void n_transverse(const orc_model* m, const double ff[4], int eigen_lin,
                         const orc_bf_grid* g, const int idx[6], int n_dir, const int grads[6],
                         double dx, double dy, double nodes_inter[12]) {
  for (int k = 0; k < 3; ++k) {
    const int ix = idx[2 * k], iy = idx[2 * k + 1];
    double xg[4], yg[4], ngp[4], ng[4], v[4], pe[5], pn[4];
    gradient_prim(m, g, ix, iy, gradient_proc(0, grads[2 * k]), dx, xg);
    gradient_prim(m, g, ix, iy, gradient_proc(1, grads[2 * k + 1]), dy, yg);
    for (int l = 0; l < 4; ++l)
      ngp[l] = n_dir == ORC_N1 ? n2_coord(xg[l], yg[l]) : n1_coord(xg[l], yg[l]);
    xy_to_n(ngp, ng);
    grid_point(g, ix, iy, v);
    obc_eigenqties(m, ff, eigen_lin, v, pe);
    xy_to_n(pe, pn);
    for (int l = 0; l < 4; ++l) pe[l] = pn[l];
    trans_term(n_dir == ORC_N1 ? 0 : 1, pe, ng, nodes_inter + 4 * k);
  }
}

/* compute_newgrdpt_xy, :1825-2400: characteristic extrapolation along the
 * diagonal direction n_dir (ORC_N1: n1 = (x - y)/sqrt2, ORC_N2: n2 = (x +
 * y)/sqrt2) from a plane through three points.  grads: the x and y gradient
 * procedures (ORC_GRAD_*) of points 1, 2, 3; eigen_idx: (i, j) of the
 * eigen-decomposition point in grid 1; inter_idx1: (i, j) of the three
 * points in grid 1. */
void orc_newgrdpt_xy(const orc_model* m, const double ff[4], int eigen_lin, double dt,
                     const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1, int n_dir,
                     int side_right, const int grads[6], const int eigen_idx[2],
                     const int inter_idx1[6], double out[4]) {
  const double x1 = g1->x_map[i1 - 1], y1 = g1->y_map[j1 - 1];
  const double dx = g1->x_map[1] - g1->x_map[0];
  const double dy = g1->y_map[1] - g1->y_map[0];
  const double n1_1 = n1_coord(x1, y1), n2_1 = n2_coord(x1, y1);
  int inter_idx0[6];
  for (int k = 0; k < 3; ++k) {
    inter_idx0[2 * k] = g1->align_x - g0->align_x + inter_idx1[2 * k];
    inter_idx0[2 * k + 1] = g1->align_y - g0->align_y + inter_idx1[2 * k + 1];
  }
  double n1i[3], n2i[3], nodes_inter[12], v[4], p[4], b[4];
  double inter_nodes0[12], inter_trans0[12], inter_bf0[12], inter_trans1[12], inter_bf1[12];
  for (int k = 0; k < 3; ++k) {
    const int ix = inter_idx0[2 * k], iy = inter_idx0[2 * k + 1];
    const double x = g0->x_map[ix - 1], y = g0->y_map[iy - 1];
    n1i[k] = n1_coord(x, y);
    n2i[k] = n2_coord(x, y);
    grid_point(g0, ix, iy, v);
    prim_var(m, v, p);
    xy_to_n(p, nodes_inter + 4 * k);
  }
  orc_interp_coeff_2d(n1i, n2i, nodes_inter, inter_nodes0);
  n_transverse(m, ff, eigen_lin, g0, inter_idx0, n_dir, grads, dx, dy, nodes_inter);
  orc_interp_coeff_2d(n1i, n2i, nodes_inter, inter_trans0);
  for (int k = 0; k < 3; ++k) {
    body_forces_prim(m, b);
    xy_to_n(b, nodes_inter + 4 * k);
  }
  orc_interp_coeff_2d(n1i, n2i, nodes_inter, inter_bf0);
  for (int k = 0; k < 3; ++k) {
    const int ix = inter_idx1[2 * k], iy = inter_idx1[2 * k + 1];
    const double x = g1->x_map[ix - 1], y = g1->y_map[iy - 1];
    n1i[k] = n1_coord(x, y);
    n2i[k] = n2_coord(x, y);
    body_forces_prim(m, b);
    xy_to_n(b, nodes_inter + 4 * k);
  }
  orc_interp_coeff_2d(n1i, n2i, nodes_inter, inter_bf1);
  n_transverse(m, ff, eigen_lin, g1, inter_idx1, n_dir, grads, dx, dy, nodes_inter);
  orc_interp_coeff_2d(n1i, n2i, nodes_inter, inter_trans1);
  double t_amp1[4], b_amp1[4];
  orc_interpolate_2d(n1_1, n2_1, inter_trans1, t_amp1);
  orc_interpolate_2d(n1_1, n2_1, inter_bf1, b_amp1);
  double pe[5], pn[4];
  grid_point(g1, eigen_idx[0], eigen_idx[1], v);
  obc_eigenqties(m, ff, eigen_lin, v, pe);
  xy_to_n(pe, pn);
  for (int l = 0; l < 4; ++l) pe[l] = pn[l];
  double lam[4], L[4][4], R[4][4];
  const int ydir = n_dir == ORC_N2;
  ns_vdw2d_eigen(ydir, pe[0], ydir ? pe[2] : pe[1], pe[4], lam, L, R);
  double char_amp[4];
  for (int k = 0; k < 4; ++k) {
    const double n1_0 = ydir ? n1_1 : n1_1 - lam[k] * dt;
    const double n2_0 = ydir ? n2_1 - lam[k] * dt : n2_1;
    const int inside = side_right ? (lam[k] >= 0) : !(lam[k] > 0);
    double n_amp0[4], t_amp0[4], b_amp0[4];
    if (inside) {
      orc_interpolate_2d(n1_0, n2_0, inter_nodes0, n_amp0);
      orc_interpolate_2d(n1_0, n2_0, inter_trans0, t_amp0);
      orc_interpolate_2d(n1_0, n2_0, inter_bf0, b_amp0);
    } else {
      far_field_amp(m, ff, 1, n_amp0, t_amp0, b_amp0);
    }
    double d0[4], d1[4], integ[4];
    for (int l = 0; l < 4; ++l) {
      d0[l] = t_amp0[l] - b_amp0[l];
      d1[l] = t_amp1[l] - b_amp1[l];
    }
    orc_newton_cotes(d0, d1, dt, integ);
    double acc = 0.0;
    for (int l = 0; l < 4; ++l) acc += (n_amp0[l] - integ[l]) * L[l][k];
    char_amp[k] = acc;
  }
  double prim_n[4], prim[4];
  for (int col = 0; col < 4; ++col) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += char_amp[k] * R[k][col];
    prim_n[col] = acc;
  }
  n_to_xy(prim_n, prim);
  cons_var(m, prim, out);
}

/* compute_newgrdpt_local, bf_newgrdpt_class.f:245-1000, for the corner
 * procedures and the edge procedures (the diagonal *_edge_type procedures of
 * irregular buffer layers are not restated) */
void orc_newgrdpt_local(const orc_model* m, const double ff[4], int eigen_lin, double dt,
                        const orc_bf_grid* g0, const orc_bf_grid* g1, int i1, int j1,
                        int procedure, int grad_type, double out[4]) {
  enum { I = ORC_GRAD_I, L0 = ORC_GRAD_L0, R0 = ORC_GRAD_R0 };
  switch (procedure) {
    case ORC_NG_SW_CORNER: {
      const int gr[6] = {L0, L0, I, L0, L0, I};
      const int e[2] = {i1 + 1, j1 + 1};
      const int idx[6] = {i1 + 1, j1 + 1, i1 + 2, j1 + 1, i1 + 1, j1 + 2};
      orc_newgrdpt_xy(m, ff, eigen_lin, dt, g0, g1, i1, j1, ORC_N2, 0, gr, e, idx, out);
      return;
    }
    case ORC_NG_SE_CORNER: {
      const int gr[6] = {I, L0, R0, L0, R0, I};
      const int e[2] = {i1 - 1, j1 + 1};
      const int idx[6] = {i1 - 2, j1 + 1, i1 - 1, j1 + 1, i1 - 1, j1 + 2};
      orc_newgrdpt_xy(m, ff, eigen_lin, dt, g0, g1, i1, j1, ORC_N1, 1, gr, e, idx, out);
      return;
    }
    case ORC_NG_NW_CORNER: {
      const int gr[6] = {L0, I, L0, R0, I, R0};
      const int e[2] = {i1 + 1, j1 - 1};
      const int idx[6] = {i1 + 1, j1 - 2, i1 + 1, j1 - 1, i1 + 2, j1 - 1};
      orc_newgrdpt_xy(m, ff, eigen_lin, dt, g0, g1, i1, j1, ORC_N1, 0, gr, e, idx, out);
      return;
    }
    case ORC_NG_NE_CORNER: {
      const int gr[6] = {R0, I, I, R0, R0, R0};
      const int e[2] = {i1 - 1, j1 - 1};
      const int idx[6] = {i1 - 1, j1 - 2, i1 - 2, j1 - 1, i1 - 1, j1 - 1};
      orc_newgrdpt_xy(m, ff, eigen_lin, dt, g0, g1, i1, j1, ORC_N2, 1, gr, e, idx, out);
      return;
    }
    case ORC_NG_S_EDGE: orc_newgrdpt_y(m, ff, eigen_lin, dt, g0, g1, i1, j1, 0, grad_type, out); return;
    case ORC_NG_N_EDGE: orc_newgrdpt_y(m, ff, eigen_lin, dt, g0, g1, i1, j1, 1, grad_type, out); return;
    case ORC_NG_W_EDGE: orc_newgrdpt_x(m, ff, eigen_lin, dt, g0, g1, i1, j1, 0, grad_type, out); return;
    case ORC_NG_E_EDGE: orc_newgrdpt_x(m, ff, eigen_lin, dt, g0, g1, i1, j1, 1, grad_type, out); return;
    default:
      for (int k = 0; k < 4; ++k) out[k] = NAN;
  }
}

/* get_config_id and get_config_procedures, bf_newgrdpt_procedure_module.f:
 * the 3x3 neighbourhood of a new point, bit k-1 set when neighbour k is a
 * boundary point of the grid (k = 1..8: SW, S, SE, W, E, NW, N, NE).  The
 * configurations a rectangular grid meets when it grows by one layer, with
 * the reference's procedure for each; -1 for the others. */
int orc_newgrdpt_config_procedure(int config_id, int* procedure, int* grad_type) {
  static const int table[][3] = {
      {1, ORC_NG_NE_CORNER, ORC_GRAD_I},  {3, ORC_NG_N_EDGE, ORC_GRAD_R0},
      {4, ORC_NG_NW_CORNER, ORC_GRAD_I},  {6, ORC_NG_N_EDGE, ORC_GRAD_L0},
      {7, ORC_NG_N_EDGE, ORC_GRAD_I},     {9, ORC_NG_E_EDGE, ORC_GRAD_R0},
      {20, ORC_NG_W_EDGE, ORC_GRAD_R0},   {32, ORC_NG_SE_CORNER, ORC_GRAD_I},
      {40, ORC_NG_E_EDGE, ORC_GRAD_L0},   {41, ORC_NG_E_EDGE, ORC_GRAD_I},
      {96, ORC_NG_S_EDGE, ORC_GRAD_R0},   {128, ORC_NG_SW_CORNER, ORC_GRAD_I},
      {144, ORC_NG_W_EDGE, ORC_GRAD_L0},  {148, ORC_NG_W_EDGE, ORC_GRAD_I},
      {192, ORC_NG_S_EDGE, ORC_GRAD_L0},  {224, ORC_NG_S_EDGE, ORC_GRAD_I}};
  for (size_t q = 0; q < sizeof table / sizeof table[0]; ++q)
    if (table[q][0] == config_id) {
      *procedure = table[q][1];
      *grad_type = table[q][2];
      return 0;
    }
  return -1;
}

/* A rectangular grid grown by one layer at the edges with add[e] = 1 (W, E,
 * S, N): the old state at time t (nodes1, old extents m->nx x m->ny) moves to
 * its place in the new grid, and every new point comes from
 * orc_newgrdpt_local with the procedure of its configuration of old points,
 * the state at t - dt (nodes0, old extents) and the far field ff.  Every new
 * point reads old points only, so the order of the points does not matter.
 * out: (m->nx + add W + add E) x (m->ny + add S + add N) x 4.
 * Returns 0, or -1 for an unexpected configuration. */
int orc_extend_newgrdpt(const orc_model* m, const double ff[4], double dt, const double* nodes0,
                        const double* nodes1, const double* x_map, const double* y_map,
                        const int add[4], double* out) {
  const int nxo = m->nx, nyo = m->ny;
  const int aw = add[0], as = add[2];
  const int nxn = nxo + add[0] + add[1], nyn = nyo + add[2] + add[3];
  const double dx = x_map[1] - x_map[0], dy = y_map[1] - y_map[0];
  double* xn = (double*)malloc(sizeof(double) * (size_t)nxn);
  double* yn = (double*)malloc(sizeof(double) * (size_t)nyn);
  for (int i = 1; i <= nxn; ++i) xn[i - 1] = x_map[0] + (i - 1 - aw) * dx;
  for (int j = 1; j <= nyn; ++j) yn[j - 1] = y_map[0] + (j - 1 - as) * dy;
  for (int i = 1; i <= nxo; ++i) xn[i - 1 + aw] = x_map[i - 1];
  for (int j = 1; j <= nyo; ++j) yn[j - 1 + as] = y_map[j - 1];
  for (size_t q = 0; q < (size_t)nxn * nyn * 4; ++q) out[q] = NAN;
  for (int k = 1; k <= 4; ++k)
    for (int j = 1; j <= nyo; ++j)
      for (int i = 1; i <= nxo; ++i)
        out[((size_t)(k - 1) * nyn + (j - 1 + as)) * nxn + (i - 1 + aw)] =
            nodes1[((size_t)(k - 1) * nyo + (j - 1)) * nxo + (i - 1)];
  const orc_bf_grid g0 = {nxo, nyo, aw, as, x_map, y_map, nodes0};
  const orc_bf_grid g1 = {nxn, nyn, 0, 0, xn, yn, out};
  int status = 0;
  for (int j = 1; j <= nyn; ++j)
    for (int i = 1; i <= nxn; ++i) {
      const int io = i - aw, jo = j - as;
      if (io >= 1 && io <= nxo && jo >= 1 && jo <= nyo) continue;
      int config = 0, bit = 0;
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
          if (!di && !dj) continue;
          const int a = io + di, b = jo + dj;
          if (a >= 1 && a <= nxo && b >= 1 && b <= nyo) config |= 1 << bit;
          ++bit;
        }
      int proc, gt;
      double v[4];
      if (orc_newgrdpt_config_procedure(config, &proc, &gt)) {
        status = -1;
        continue;
      }
      orc_newgrdpt_local(m, ff, 0, dt, &g0, &g1, i, j, proc, gt, v);
      for (int k = 1; k <= 4; ++k) out[((size_t)(k - 1) * nyn + (j - 1)) * nxn + (i - 1)] = v[k - 1];
    }
  free(xn);
  free(yn);
  return status;
}
