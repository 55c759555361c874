// dim2d_kernels.hip -- gfx950 kernels of the DIM RK3-TVD hot path.
//
// One fused kernel per RK stage.  A workgroup is one wave of 64 lanes that owns
// a strip of 64 columns (60 output columns plus a 2-column halo on each side,
// one lane per column) and `rows` output rows, and marches up the strip one row
// per iteration.  Per row:
//   * issue the loads of the cells of row r+4 (two 16-byte loads per lane from
//     the interleaved cell layout: 2 KiB contiguous per wave; ghost cells of
//     reflection/periodic edges are resolved by index maps, so no ghost-fill
//     launch exists for them) and of the RK inputs of row r+1;
//   * compute the primitives of row r+1 once per cell (1/rho, u, v, P, T,
//     1/(3-rho) and the transport products the fluxes need) and the capillary
//     temperature of that row;
//   * x-face flux at (i-1/2, r) and y-face flux at (i, r+1/2); the values of
//     the neighbouring columns move between lanes by DPP wavefront shifts (no
//     LDS, no barrier);
//   * flux-difference divergence with the right x-face from the next lane and
//     the lower y-face from the previous iteration, then the RK stage update.
// Stage 3 also reduces max(|u|+c) and min/max rho of the new state for the
// next CFL time step.
//
// Reference semantics followed here (jdesmarais/dim2D_finite_volume):
//   fluxes  physical_models/dim2d/dim2d_fluxes_module.f:103-1649
//   stencils sd_operators/mattsson_operators/sd_operators_class.f:185-1039
//   divergence td_operators/finitevolume/td_operators_class.f:181-194
//   RK stages td_integrator/rungekutta3rdtvd/rk3tvd_steps_module.f:52-490
//   wall  boundary_conditions/wall_xy/wall_xy_equilibrium_module.f, ridders_method_module.f
#include "dim2d_internal.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

namespace dim2d {

// ------------------------------------------------------------------------
// virtual ghost cells: index maps of reflection and periodic edges
// ------------------------------------------------------------------------
// reflection_x: ghost i mirrors about the boundary face (bc_operators_reflection_x_class.f:206-234),
// periodic: i +/- (nx - 2*bc_size) (periodic_xy/bc_operators_class.f:127-128,156-175);
// wall ghosts are stored cells.
__device__ __forceinline__ int map_col(const DevParams& p, int c, bool& flip) {
  flip = false;
  if (c < kBcSize) {
    if (p.bc_w == kBcReflection) { flip = true; return 2 * kBcSize - 1 - c; }
    if (p.bc_w == kBcPeriodic) return c + p.nx - 2 * kBcSize;
  } else if (c >= p.nx - kBcSize) {
    if (p.bc_e == kBcReflection) { flip = true; return 2 * p.nx - 2 * kBcSize - 1 - c; }
    if (p.bc_e == kBcPeriodic) return c - (p.nx - 2 * kBcSize);
  }
  return c;
}
__device__ __forceinline__ int map_row(const DevParams& p, int r, bool& flip) {
  flip = false;
  if (r < kBcSize) {
    if (p.bc_s == kBcReflection) { flip = true; return 2 * kBcSize - 1 - r; }
    if (p.bc_s == kBcPeriodic) return r + p.ny - 2 * kBcSize;
  } else if (r >= p.ny - kBcSize) {
    if (p.bc_n == kBcReflection) { flip = true; return 2 * p.ny - 2 * kBcSize - 1 - r; }
    if (p.bc_n == kBcPeriodic) return r - (p.ny - 2 * kBcSize);
  }
  return r;
}

// ------------------------------------------------------------------------
// cell access: a cell is 4 consecutive FP64 (two 16-byte loads) or FP32 (one
// 16-byte load) values; the row offset is uniform across the wave, the column
// a 32-bit lane offset.
// ------------------------------------------------------------------------
__device__ __forceinline__ void ld_cell(const double* __restrict__ row, unsigned col, double v[4]) {
  const double2* q = reinterpret_cast<const double2*>(row) + 2 * col;
  const double2 a = q[0], b = q[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void ld_cell(const float* __restrict__ row, unsigned col, double v[4]) {
  const float4 a = reinterpret_cast<const float4*>(row)[col];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}
__device__ __forceinline__ void st_cell(double* __restrict__ row, unsigned col, const double v[4]) {
  double2* q = reinterpret_cast<double2*>(row) + 2 * col;
  q[0] = make_double2(v[0], v[1]);
  q[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void st_cell(float* __restrict__ row, unsigned col, const float v[4]) {
  reinterpret_cast<float4*>(row)[col] = make_float4(v[0], v[1], v[2], v[3]);
}

// state value of cell (cm, rm) (already mapped) with the reflection signs
// (momentum normal to a reflecting edge changes sign, reflection_xy_module.f:50-110).
// SRC 0: FP64 field S64; 1: U alone; 2: U + S32, the increment kept in `inc`
template <int SRC>
__device__ __forceinline__ void load_state(const DevParams& p, const double* __restrict__ U,
                                           const double* __restrict__ S64,
                                           const float* __restrict__ S32, int rm, unsigned cm,
                                           bool fx, bool fy, double v[4], float inc[4]) {
  const long long row = (long long)rm * p.pitch * 4;
  if (SRC == 0) {
    ld_cell(S64 + row, cm, v);
  } else if (SRC == 1) {
    ld_cell(U + row, cm, v);
  } else {
    const float4 d = reinterpret_cast<const float4*>(S32 + row)[cm];
    ld_cell(U + row, cm, v);
    inc[0] = d.x;
    inc[1] = d.y;
    inc[2] = d.z;
    inc[3] = d.w;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += (double)inc[k];
  }
  if (fx) v[1] = -v[1];
  if (fy) v[2] = -v[2];
}

// ------------------------------------------------------------------------
// per-cell quantities (dim2d_prim_module.f)
// ------------------------------------------------------------------------
// 1/x by the hardware reciprocal and two Newton steps (error ~1 ulp)
__device__ __forceinline__ double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// constant p.<name> in the flux precision R: the double itself, or its FP32
// copy in p.f (the same value as (float)p.<name>)
#define PC(R, name) (std::is_same<R, float>::value ? (R)p.f.name : (R)p.name)

// Own-column values a lane keeps in registers (what the y-faces need)
template <typename RD, typename R>
struct Cell {
  RD rho;
  double irho;
  R u, v, w, Tcl, Tcap;
  R qy, qxv, Pyy, Hv, vw;  // qy, qx*v, qy*v+P, (E+P)v, v w
};
// x-face values of a cell, exchanged through LDS
template <typename R>
struct XVals {
  R qx, Pxx, Hu, uw;  // qx, qx*u+P, (E+P)u, u w
};

// velocities, classical pressure (:425-453), classical temperature (:515-544),
// capillarity pressure 1/(3-rho) (:1030-1062) and the transport products
// (:830-1150), computed in FP64 from the FP64 state.  F32P (FP32 fluxes
// only): the transport products from the FP32 operands instead, 8
// conversions per cell instead of 13, each product ~1.5 FP32 ulp instead of
// 0.5.  Every mixed stage but the last of a step takes it (stages 1 and 2
// alone or as the stage pair, which is bound by instruction issue); the last
// stage does not, since its fewer registers would let it run 3 waves per
// SIMD, which is slower for that HBM-bound kernel.
template <typename RD, typename R, bool F32P = false>
__device__ __forceinline__ void make_cell(const DevParams& p, const double v[4], Cell<RD, R>& c,
                                          XVals<R>& x) {
  const double rho = v[0];
  const double ir = fast_rcp(rho);
  const double u = v[1] * ir;
  const double vv = v[2] * ir;
  const double rho2 = rho * rho;
  const double Tn = v[3] - 0.5 * rho * (u * u + vv * vv) + 3.0 * rho2;
  const double w = fast_rcp(3.0 - rho);
  const double P = p.c3cv * w * Tn - 3.0 * rho2;
  const double H = v[3] + P;
  c.rho = (RD)rho;
  c.irho = ir;
  c.u = (R)u;
  c.v = (R)vv;
  c.w = (R)w;
  c.Tcl = (R)(Tn * ir);
  c.Tcap = (R)0;
  c.qy = (R)v[2];
  x.qx = (R)v[1];
  if (F32P && std::is_same<R, float>::value) {
    const R Pr = (R)P, Hr = (R)H;
    c.qxv = x.qx * c.v;
    c.Pyy = c.qy * c.v + Pr;
    c.Hv = Hr * c.v;
    c.vw = c.v * c.w;
    x.Pxx = x.qx * c.u + Pr;
    x.Hu = Hr * c.u;
    x.uw = c.u * c.w;
  } else {
    c.qxv = (R)(v[1] * vv);
    c.Pyy = (R)(v[2] * vv + P);
    c.Hv = (R)(H * vv);
    c.vw = (R)(vv * w);
    x.Pxx = (R)(v[1] * u + P);
    x.Hu = (R)(H * u);
    x.uw = (R)(u * w);
  }
}

// ------------------------------------------------------------------------
// cross-lane exchange: one wave is one strip of 64 columns, so the values of
// the neighbouring columns live in the neighbouring lanes' registers and move
// with DPP wavefront shifts (VALU register moves, no LDS, no barrier).
// shr: lane i receives lane i-1; shl: lane i receives lane i+1; a lane
// without a source (lane 0 for shr, 63 for shl) receives 0.  Only lanes 0, 1,
// 62 and 63 ever depend on those zeros, and they are halo lanes whose results
// are never stored nor reduced.
// ------------------------------------------------------------------------
__device__ __forceinline__ int dpp_shr_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true);  // wave_shr:1, bound_ctrl
}
__device__ __forceinline__ int dpp_shl_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true);  // wave_shl:1, bound_ctrl
}
__device__ __forceinline__ float shr(float x) { return __int_as_float(dpp_shr_i(__float_as_int(x))); }
__device__ __forceinline__ float shl(float x) { return __int_as_float(dpp_shl_i(__float_as_int(x))); }
__device__ __forceinline__ double shr(double x) {
  return __hiloint2double(dpp_shr_i(__double2hiint(x)), dpp_shr_i(__double2loint(x)));
}
__device__ __forceinline__ double shl(double x) {
  return __hiloint2double(dpp_shl_i(__double2hiint(x)), dpp_shl_i(__double2loint(x)));
}

// density and velocities of a neighbouring cell
template <typename RD, typename R>
struct Nb {
  RD rho;
  R u, v;
};
template <typename RD, typename R>
__device__ __forceinline__ Nb<RD, R> nb_shr(const Cell<RD, R>& c) {
  return Nb<RD, R>{shr(c.rho), shr(c.u), shr(c.v)};
}
template <typename RD, typename R>
__device__ __forceinline__ Nb<RD, R> nb_shl(const Cell<RD, R>& c) {
  return Nb<RD, R>{shl(c.rho), shl(c.u), shl(c.v)};
}

// capillarity_temperature_eff (:582-613) with gradient_x/y_interior
template <typename RD, typename R>
__device__ __forceinline__ R cap_temperature(const DevParams& p, double irho, RD rho_w, RD rho_e,
                                             RD rho_s, RD rho_n) {
  const RD gx = (rho_e - rho_w) * (RD)p.half_idx;
  const RD gy = (rho_n - rho_s) * (RD)p.half_idy;
  return (R)(0.5 * irho * (double)(gx * gx + gy * gy));
}

// pressure(T, rho) of the wall (dim2d_prim_module.f:476-488)
__device__ __forceinline__ double wall_pressure(double T, double md) {
  return 8.0 * md * T / (3.0 - md) - 3.0 * md * md;
}

// x-face flux at (i-1/2, j): L = (i-1, j), R = (i, j)  (dim2d_fluxes_module.f,
// Mattsson f/dfdx/dfdy/d2fdx2/d2fdy2/dfdx_nl stencils), returned doubled
// (F2 = 2 F) so the face averages are plain sums; the divergence applies the
// 1/2.  L: left cell (its x-values in Lx); LL/RR: densities of (i-2, j) and
// (i+1, j); Ldn/Lup, Rdn/Rup: rows j-1 and j+1 of the two columns.
template <typename RD, typename R>
__device__ __forceinline__ void x_face(const DevParams& p, const Nb<RD, R>& L, const Cell<RD, R>& Lc,
                                       const XVals<R>& Lx, RD rhoLL, RD rhoRR,
                                       const Nb<RD, R>& Ldn, const Nb<RD, R>& Lup,
                                       const Cell<RD, R>& Rc, const XVals<R>& Rx,
                                       const Nb<RD, R>& Rdn, const Nb<RD, R>& Rup, bool wall_w,
                                       bool wall_e, R F2[4]) {
  const RD srho = L.rho + Rc.rho;
  const R rx = (R)((Rc.rho - L.rho) * (RD)p.idx);
  const R ry = (R)(((Lup.rho + Rup.rho) - (Ldn.rho + Rdn.rho)) * (RD)p.q_idy);
  const R rxx = (R)(((rhoLL + rhoRR) - srho) * (RD)p.h2x);
  const R ryy = (R)(((Ldn.rho + Rdn.rho) + (Lup.rho + Rup.rho) - (RD)2 * srho) * (RD)p.h2y);
  const R su = L.u + Rc.u;
  const R sv = L.v + Rc.v;
  const R dudx = (Rc.u - L.u) * PC(R, idx);
  const R dvdx = (Rc.v - L.v) * PC(R, idx);
  const R dudy = ((Lup.u + Rup.u) - (Ldn.u + Rdn.u)) * PC(R, q_idy);
  const R dvdy = ((Lup.v + Rup.v) - (Ldn.v + Rdn.v)) * PC(R, q_idy);
  const R rxy = rx * ry;
  const R grad2 = rx * rx + ry * ry;
  const R lap = rxx + ryy;
  const R rx2d = ry * ry - rx * rx;
  const R srhoR = (R)srho;

  if (wall_w || wall_e) {
    // compute_wall_flux_x (wall_xy_equilibrium_module.f:249-341)
    const double TeffL = (double)Lc.Tcl - p.inv_we * (double)Lc.Tcap;
    const double TeffR = (double)Rc.Tcl - p.inv_we * (double)Rc.Tcap;
    const double Tl = p.c38cv * (wall_w ? TeffR : TeffL);
    const double Th = Tl + p.dx * p.c38cv * p.re * p.pr * p.wall_q;
    const double Pw = wall_pressure(0.5 * (Tl + Th), 0.5 * (double)srho);
    F2[0] = (R)0;
    F2[1] = (R)(2.0 * Pw - p.inv_we * (double)(srhoR * lap + rx2d));
    F2[2] = (R)(2.0 * (-p.inv_re * (double)dvdx + p.inv_we * (double)rxy));
    F2[3] = (R)(2.0 * (wall_w ? p.wall_q : -p.wall_q));
    return;
  }
  const R visc = PC(R, two_vr) * dudx + PC(R, vr) * dvdy;
  const R shear = dudy + dvdx;
  F2[0] = Lx.qx + Rx.qx;
  F2[1] = (Lx.Pxx + Rx.Pxx) - PC(R, two_inv_re) * visc -
          PC(R, inv_we) * (PC(R, c15cv) * (Lc.w + Rc.w) * grad2 + srhoR * lap + rx2d);
  F2[2] = (Lc.qxv + Rc.qxv) - PC(R, two_inv_re) * shear + PC(R, two_inv_we) * rxy;
  const R visc_e2 = su * visc + sv * shear + PC(R, k_heat_x) * (Rc.Tcl - Lc.Tcl);
  const R cap_e2 = PC(R, c15cv) * (Lx.uw + Rx.uw) * grad2 + F2[0] * lap + (R)0.5 * su * rx2d -
                   rxy * sv - srhoR * rx * (dudx + dvdy) - PC(R, k_cap_x) * (Rc.Tcap - Lc.Tcap);
  F2[3] = (Lx.Hu + Rx.Hu) - PC(R, inv_re) * visc_e2 - PC(R, inv_we) * cap_e2;
}

// y-face flux at (i, j-1/2): D = (i, j-1), U = (i, j)  (g/dgdx/dgdy/d2gdx2/d2gdy2/dgdy_nl),
// returned doubled like x_face.  WD/ED, WU/EU: columns i-1 and i+1 of both rows.
template <typename RD, typename R>
__device__ __forceinline__ void y_face(const DevParams& p, const Cell<RD, R>& D,
                                       const Cell<RD, R>& Uc, const Nb<RD, R>& WD,
                                       const Nb<RD, R>& ED, const Nb<RD, R>& WU,
                                       const Nb<RD, R>& EU, RD rho_dd, RD rho_uu, bool wall_s,
                                       bool wall_n, R G2[4]) {
  const RD srho = D.rho + Uc.rho;
  const R ry = (R)((Uc.rho - D.rho) * (RD)p.idy);
  const R rx = (R)(((ED.rho + EU.rho) - (WD.rho + WU.rho)) * (RD)p.q_idx);
  const R ryy = (R)(((rho_dd + rho_uu) - srho) * (RD)p.h2y);
  const R rxx = (R)(((WD.rho + WU.rho) + (ED.rho + EU.rho) - (RD)2 * srho) * (RD)p.h2x);
  const R su = D.u + Uc.u;
  const R sv = D.v + Uc.v;
  const R dudy = (Uc.u - D.u) * PC(R, idy);
  const R dvdy = (Uc.v - D.v) * PC(R, idy);
  const R dudx = ((ED.u + EU.u) - (WD.u + WU.u)) * PC(R, q_idx);
  const R dvdx = ((ED.v + EU.v) - (WD.v + WU.v)) * PC(R, q_idx);
  const R rxy = rx * ry;
  const R grad2 = rx * rx + ry * ry;
  const R lap = rxx + ryy;
  const R ry2d = rx * rx - ry * ry;
  const R srhoR = (R)srho;

  if (wall_s || wall_n) {
    // compute_wall_flux_y (wall_xy_equilibrium_module.f:342-444)
    const double TeffD = (double)D.Tcl - p.inv_we * (double)D.Tcap;
    const double TeffU = (double)Uc.Tcl - p.inv_we * (double)Uc.Tcap;
    const double Tl = p.c38cv * (wall_s ? TeffU : TeffD);
    const double Th = Tl + p.dy * p.re * p.pr * p.c38cv * p.wall_q;
    const double Pw = wall_pressure(0.5 * (Tl + Th), 0.5 * (double)srho);
    G2[0] = (R)0;
    G2[1] = (R)(2.0 * (-p.inv_re * (double)dudy + p.inv_we * (double)rxy));
    G2[2] = (R)(2.0 * Pw - p.inv_we * (double)(srhoR * lap + ry2d));
    G2[3] = (R)(2.0 * (wall_s ? p.wall_q : -p.wall_q));
    return;
  }
  const R shear = dudy + dvdx;
  const R visc = PC(R, vr) * dudx + PC(R, two_vr) * dvdy;
  G2[0] = D.qy + Uc.qy;
  G2[1] = (D.qxv + Uc.qxv) - PC(R, two_inv_re) * shear + PC(R, two_inv_we) * rxy;
  G2[2] = (D.Pyy + Uc.Pyy) - PC(R, two_inv_re) * visc -
          PC(R, inv_we) * (PC(R, c15cv) * (D.w + Uc.w) * grad2 + srhoR * lap + ry2d);
  const R visc_e2 = su * shear + sv * visc + PC(R, k_heat_y) * (Uc.Tcl - D.Tcl);
  const R cap_e2 = PC(R, c15cv) * (D.vw + Uc.vw) * grad2 + G2[0] * lap + (R)0.5 * sv * ry2d -
                   rxy * su - srhoR * ry * (dudx + dvdy) - PC(R, k_cap_y) * (Uc.Tcap - D.Tcap);
  G2[3] = (D.Hv + Uc.Hv) - PC(R, inv_re) * visc_e2 - PC(R, inv_we) * cap_e2;
}

// speed_of_sound (dim2d_prim_module.f:1179-1215), algebraically reduced to
// c^2 = 3 (e + (P + 3 rho^2)/cv) / (rho (3 - rho)),  e = P + rho^2 (2 rho - 3);
// c = 0 where c^2 <= 0 (the spinodal part of an interface has no real sound speed)
__device__ __forceinline__ double sound_speed(const DevParams& p, const double v[4],
                                              double* speed) {
  const double rho = v[0];
  const double ir = fast_rcp(rho);
  const double u = v[1] * ir, vv = v[2] * ir;
  const double rho2 = rho * rho;
  const double Tn = v[3] - 0.5 * rho * (u * u + vv * vv) + 3.0 * rho2;
  const double w = fast_rcp(3.0 - rho);
  const double P = p.c3cv * w * Tn - 3.0 * rho2;
  const double e = P + rho2 * (2.0 * rho - 3.0);
  const double c2 = 3.0 * (e + (P + 3.0 * rho2) * p.c3cv_3) * w * ir;
  *speed = sqrt(u * u + vv * vv);
  return c2 > 0.0 ? sqrt(c2) : 0.0;
}
// The same sound speed for the time step of mixed-precision runs: 1/rho, the
// velocities and the temperature recovered from the total energy (a sum of
// terms that nearly cancel) in FP64; the EOS terms, the reciprocal of 3 - rho
// and both square roots in FP32.  The time step then differs from the FP64
// rule by a few FP32 roundings (~1e-7 relative), far below what the FP32
// stage increments already carry.
__device__ __forceinline__ float sound_speed_f32(const DevParams& p, const double v[4],
                                                 float* speed) {
  const double rho = v[0];
  const double ir = fast_rcp(rho);
  const double u = v[1] * ir, vv = v[2] * ir;
  const double rho2 = rho * rho;
  const double Tn = v[3] - 0.5 * rho * (u * u + vv * vv) + 3.0 * rho2;
  const float rf = (float)rho, r2 = (float)rho2, uf = (float)u, vf = (float)vv;
  const float w = __builtin_amdgcn_rcpf(3.0f - rf);
  const float c3cv = (float)p.c3cv;
  const float P = c3cv * w * (float)Tn - 3.0f * r2;
  const float e = P + r2 * (2.0f * rf - 3.0f);
  const float c2 = 3.0f * (e + (P + 3.0f * r2) * p.f.c3cv_3) * w * (float)ir;
  *speed = __builtin_amdgcn_sqrtf(uf * uf + vf * vf);
  return c2 > 0.0f ? __builtin_amdgcn_sqrtf(c2) : 0.0f;
}

// Reductions over the 64 lanes of a wave (all lanes active), result in every
// lane.  Within each row of 16 lanes by DPP moves (VALU, no LDS round trip):
// quad_perm [1,0,3,2] and [2,3,0,1] combine the 4 lanes of a quad, row_ror:4
// and row_ror:8 the 4 quads of a row; then the four row results are read
// from lanes 0, 16, 32 and 48.
template <int CTRL>
__device__ __forceinline__ int dpp_move(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_move<CTRL>((int)b), hi = dpp_move<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int lane_value(int x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}
__device__ __forceinline__ double lane_value(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T x, Op op) {
  x = op(x, dpp_move<0xB1>(x));   // quad_perm [1,0,3,2]
  x = op(x, dpp_move<0x4E>(x));   // quad_perm [2,3,0,1]
  x = op(x, dpp_move<0x124>(x));  // row_ror:4
  x = op(x, dpp_move<0x128>(x));  // row_ror:8
  return op(op(lane_value(x, 0), lane_value(x, 16)), op(lane_value(x, 32), lane_value(x, 48)));
}
__device__ __forceinline__ double wave_sum(double x) {
  return wave_reduce(x, [](double a, double b) { return a + b; });
}
__device__ __forceinline__ double wave_max(double x) {
  return wave_reduce(x, [](double a, double b) { return fmax(a, b); });
}
__device__ __forceinline__ double wave_min(double x) {
  return wave_reduce(x, [](double a, double b) { return fmin(a, b); });
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned x) {
  return (unsigned)wave_reduce((int)x, [](int a, int b) { return (int)((unsigned)a + (unsigned)b); });
}

// relaxed device-scope read of a reduction word
__device__ __forceinline__ unsigned long long peek(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// reductions of a state: the CFL quantities (max |u|+c, min/max rho, count of
// non-finite cells) and, when SUMS, the conserved sums
template <bool SUMS>
struct RedAcc {
  double s[4];
  double amax, rmin, rmax;
  unsigned nonfinite;
  __device__ void init() {
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] = 0.0;
    amax = 0.0;
    rmin = INFINITY;
    rmax = 0.0;
    nonfinite = 0;
  }
  // F32: the sound speed of sound_speed_f32 (mixed-precision stages)
  template <bool F32 = false>
  __device__ void add(const DevParams& p, const double v[4]) {
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (SUMS) s[k] += v[k];
      bad |= !isfinite(v[k]);
    }
    double a;
    if (F32) {
      float sp;
      const float c = sound_speed_f32(p, v, &sp);
      a = (double)(sp + c);
    } else {
      double sp;
      const double c = sound_speed(p, v, &sp);
      a = sp + c;
    }
    bad |= !isfinite(a);
    nonfinite += bad ? 1u : 0u;
    if (!bad) {
      amax = fmax(amax, a);
      rmin = fmin(rmin, v[0]);
      rmax = fmax(rmax, v[0]);
    }
  }
  // one wave's contribution into slot `slot` of ctl
  __device__ void flush(StepCtl* ctl, int slot) {
    double t[4];
    if (SUMS) {
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = wave_sum(s[k]);
    }
    const double am = wave_max(amax);
    const double rn = wave_min(rmin);
    const double rx = wave_max(rmax);
    const unsigned nf = wave_sum_u(nonfinite);
    if ((threadIdx.x & 63) == 0) {
      if (SUMS) {
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&ctl->sum[slot][k], t[k]);
      }
      const unsigned long long pk[3] = {peek(&ctl->amax[slot]), peek(&ctl->rmin[slot]),
                                        peek(&ctl->rmax[slot])};
      commit(ctl, slot, am, rn, rx, nf, pk);
    }
  }
  // The contribution of a whole workgroup of up to 4 waves (all lanes
  // active): wave reductions, then the waves' results combined through LDS
  // and committed by one lane, so a launch makes one set of atomics per
  // workgroup instead of per wave (the conserved sums cannot skip theirs).
  __device__ void flush_block(StepCtl* ctl, int slot) {
    __shared__ double sh[4][8];
    __shared__ unsigned shn[4];
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    if (SUMS) {
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = wave_sum(s[k]);
    }
    const double am = wave_max(amax);
    const double rn = wave_min(rmin);
    const double rx = wave_max(rmax);
    const unsigned nf = wave_sum_u(nonfinite);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[w][k] = t[k];
      sh[w][4] = am;
      sh[w][5] = rn;
      sh[w][6] = rx;
      shn[w] = nf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = (int)(blockDim.x >> 6);
      double tt[4] = {sh[0][0], sh[0][1], sh[0][2], sh[0][3]};
      double bam = sh[0][4], brn = sh[0][5], brx = sh[0][6];
      unsigned bnf = shn[0];
      for (int q = 1; q < nw; ++q) {
        for (int k = 0; k < 4; ++k) tt[k] += sh[q][k];
        bam = fmax(bam, sh[q][4]);
        brn = fmin(brn, sh[q][5]);
        brx = fmax(brx, sh[q][6]);
        bnf += shn[q];
      }
      if (SUMS) {
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&ctl->sum[slot][k], tt[k]);
      }
      const unsigned long long pk[3] = {peek(&ctl->amax[slot]), peek(&ctl->rmin[slot]),
                                        peek(&ctl->rmax[slot])};
      commit(ctl, slot, bam, brn, brx, bnf, pk);
    }
  }
  // Combined max |u|+c, min/max rho and non-finite count into slot `slot`.
  // The stored max (min) only grows (shrinks) while the kernel runs, so a
  // value that cannot change it skips the device-scope atomic: most do, which
  // keeps thousands of waves off one cache line.  pk: the stored amax, rmin,
  // rmax read at any time since the slot was cleared (earlier reads are only
  // weaker bounds).
  static __device__ void commit(StepCtl* ctl, int slot, double am, double rn, double rx,
                                unsigned nf, const unsigned long long pk[3]) {
    const unsigned long long ab = (unsigned long long)__double_as_longlong(am);
    if (ab > pk[0]) atomicMax(&ctl->amax[slot], ab);
    if (rn >= 0.0 && rn < INFINITY) {
      const unsigned long long nb = (unsigned long long)__double_as_longlong(rn);
      if (nb < pk[1]) atomicMin(&ctl->rmin[slot], nb);
    }
    const unsigned long long xb = (unsigned long long)__double_as_longlong(fmax(rx, 0.0));
    if (xb > pk[2]) atomicMax(&ctl->rmax[slot], xb);
    if (nf) atomicAdd(&ctl->nonfinite[slot], nf);
  }
};

__device__ __forceinline__ void clear_slot(StepCtl* ctl, int slot) {
  for (int k = 0; k < 4; ++k) ctl->sum[slot][k] = 0.0;
  ctl->amax[slot] = 0ull;
  ctl->rmin[slot] = (unsigned long long)__double_as_longlong(INFINITY);
  ctl->rmax[slot] = 0ull;
  ctl->nonfinite[slot] = 0u;
}

// time step rule of dim2d_cfl_dt (include/dim2d_hip.h)
__device__ __forceinline__ double cfl_dt_from(const StepCtl* ctl, int slot) {
  const double amax = __longlong_as_double((long long)ctl->amax[slot]);
  const double rmin = __longlong_as_double((long long)ctl->rmin[slot]);
  const double rmax = __longlong_as_double((long long)ctl->rmax[slot]);
  const double h = ctl->h;
  double dt = h / amax;
  dt = fmin(dt, ctl->re * rmin * h * h);
  dt = fmin(dt, sqrt(ctl->we / rmax) * h * h);
  return ctl->cfl * dt;
}

// ------------------------------------------------------------------------
// the fused stage tile
// ------------------------------------------------------------------------
// One wave (64 lanes) owns a strip of 64 columns (60 output columns plus a
// 2-column halo on each side) and `rows` output rows, and marches up it.
// STAGE 0: time derivative into O64; 1, 2, 3: RK3-TVD stages; 4, 5, 6: the
// stages of the low-storage RK3 (FP64 only).
// FP64 mode (MIXED=0): stage input S64; out = b*U + (1-b)*(S64 + dt L) with
// b = 0 (stage 1, S64 = U), 3/4, 1/3 (rk3tvd_steps_module.f:225,444).
// Mixed mode: state = U + S32 and out increments
//   stage 1: S32' = dt L;  stage 2: S32' = 1/4 (S32 + dt L);
//   stage 3: O64 = U + (1 - 1/3) (S32 + dt L)   (FP64 accumulation, formed
//   as (U + S32) + (2/3 dt L - 1/3 S32) from the loaded stencil row).
// Low-storage RK3 (Williamson's 2N-storage scheme, FP64): stage s reads the
// stencil input S64 and q, writes q' = A_s q + dt L to Q (stages 1, 2) and
// S64 + B_s q' to O64; A = (0, -5/9, -153/128), B = (1/3, 15/16, 8/15).
// Row pipeline: the loads of a row are issued four rows before it is needed
// (its density two rows before), and the RK inputs of (i, r) one row before.
// Per output row r the wave builds the cell of row r+1, fetches the
// neighbouring columns' values by DPP, computes the x-face of row r and the
// y-face r+1/2, and writes the updated cell (i, r).  The loop-carried values
// live in rings of kRing slots indexed by compile-time constants; the loop body
// is unrolled kRing times so that the ring rotation is register renaming (no
// moves, and no wait for a load before its value is used).  The loop starts two
// rows early to fill the pipeline (no outputs there), so every output row runs
// the same instruction stream whatever the tiling.
// FP64 stages take S64(i, r) from the raw row already in registers (an output
// cell is interior, where no ghost map applies).
constexpr int kLanes = 64;
constexpr int kOutCols = kLanes - 2 * kBcSize;
constexpr int kRing = 5;

template <int K>
struct Slot {
  static constexpr int k = K;
};

// The pipeline rings of one strip: slot (q - q0) % kRing holds row q, q0 the
// first row the wave processes.
template <typename RD, typename R>
struct StripRings {
  double raw[kRing][4];    // stencil rows (FP64 state)
  Cell<RD, R> cell[kRing];
  XVals<R> xv[kRing];
  Nb<RD, R> wst[kRing], est[kRing];
  R fy[kRing][4];          // y-face fluxes (doubled) of the rows' upper faces
  float inc[kRing][4];     // stage increments of the stencil rows (mixed stages 2, 3)
  double rkU[kRing][4];    // RK inputs loaded apart from the stencil rows
};

// the first cell of a strip's pipeline (row q0, slot 0), with no capillary
// temperature: the warm-up rows that use it produce no output
template <typename RD, typename R, bool F32P = false>
__device__ __forceinline__ void strip_start(const DevParams& p, StripRings<RD, R>& s) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    s.fy[kRing - 1][k] = (R)0;
    s.rkU[0][k] = 0.0;
  }
  make_cell<RD, R, F32P>(p, s.raw[0], s.cell[0], s.xv[0]);
  s.cell[0].Tcap = (R)0;
  s.cell[kRing - 1] = s.cell[0];
  s.wst[0] = nb_shr(s.cell[0]);
  s.est[0] = nb_shl(s.cell[0]);
  s.wst[kRing - 1] = s.wst[0];
}

// Time derivative L of row r (slot K) of a strip: builds the cell of row r+1
// from raw[K1] (its capillary temperature takes the density of row r+2 from
// raw[K2]), exchanges the neighbouring columns by DPP, and forms the x-face
// flux at (i-1/2, r), the y-face flux at (i, r+1/2) and the flux-difference
// divergence plus body force (td_operators_class.f:181-194).
template <typename RD, typename R, int K, bool F32P = false>
__device__ __forceinline__ void strip_time_dev(const DevParams& p, StripRings<RD, R>& s, int r,
                                               bool wall_w, bool wall_e, R L[4]) {
  constexpr int KM = (K + kRing - 1) % kRing;  // row r-1
  constexpr int K1 = (K + 1) % kRing;          // row r+1
  constexpr int K2 = (K + 2) % kRing;          // row r+2
  // fluxes are doubled (face sums), hence the 1/2
  const R hidx = PC(R, half_idx), hidy = PC(R, half_idy);
  Cell<RD, R>& cm1 = s.cell[KM];
  Cell<RD, R>& c0 = s.cell[K];
  Cell<RD, R>& c1 = s.cell[K1];
  const XVals<R>& x0 = s.xv[K];
  // cell of row r+1 and its neighbours
  make_cell<RD, R, F32P>(p, s.raw[K1], c1, s.xv[K1]);
  s.wst[K1] = nb_shr(c1);
  s.est[K1] = nb_shl(c1);
  const Nb<RD, R>& c1W = s.wst[K1];
  const Nb<RD, R>& c1E = s.est[K1];
  const Nb<RD, R>& c0W = s.wst[K];
  const Nb<RD, R>& c0E = s.est[K];
  const Nb<RD, R>& cm1W = s.wst[KM];
  c1.Tcap = cap_temperature<RD, R>(p, c1.irho, c1W.rho, c1E.rho, c0.rho, (RD)s.raw[K2][0]);

  // the west cell of row r for the x-face
  Cell<RD, R> Lc;
  XVals<R> Lx;
  Lc.w = shr(c0.w);
  Lc.Tcl = shr(c0.Tcl);
  Lc.Tcap = shr(c0.Tcap);
  Lc.qxv = shr(c0.qxv);
  Lx.qx = shr(x0.qx);
  Lx.Pxx = shr(x0.Pxx);
  Lx.Hu = shr(x0.Hu);
  Lx.uw = shr(x0.uw);
  const RD rhoLL = shr(c0W.rho);

  R F[4];
  R* G = s.fy[K];
  const R* fy_prev = s.fy[KM];
  x_face<RD, R>(p, c0W, Lc, Lx, rhoLL, c0E.rho, cm1W, c1W, c0, x0,
                Nb<RD, R>{cm1.rho, cm1.u, cm1.v}, Nb<RD, R>{c1.rho, c1.u, c1.v}, wall_w, wall_e,
                F);
  y_face<RD, R>(p, c0, c1, c0W, c0E, c1W, c1E, cm1.rho, (RD)s.raw[K2][0],
                p.bc_s == kBcWall && r == kBcSize - 1,
                p.bc_n == kBcWall && r == p.ny - kBcSize - 1, G);
#pragma unroll
  for (int k = 0; k < 4; ++k) L[k] = (F[k] - shl(F[k])) * hidx + (fy_prev[k] - G[k]) * hidy;
  if (p.gravity_ac) {
    // compute_body_forces (pmodel_eq_class.f:1840-1918)
    L[2] -= (R)((double)c0.rho * p.gravity);
    L[3] -= (R)((double)c0.qy * p.gravity);
  }
}

template <typename RD, typename R, int MIXED, int STAGE, bool REDUCE>
__device__ __forceinline__ void stage_tile(const DevParams& p, const StageArgs& a, double dt,
                                           int bx, int by, int rows, RedAcc<false>& red) {
  // mixed RK3-TVD stages 2 and 3 read u^n + increment, mixed stages 0/1 u^n
  // alone; the low-storage stages read their FP64 state (S64) in both modes
  constexpr int SRC = (MIXED && STAGE < 4) ? ((STAGE == 2 || STAGE == 3) ? 2 : 1) : 0;
  // FP32 transport products in every mixed stage but a step's last
  constexpr bool F32P = MIXED && STAGE != 3 && STAGE != 6;
  // u^n (RK3-TVD FP64 stages 2, 3) or q (low-storage stages 2, 3; FP32 in
  // mixed precision) is loaded apart from the stencil rows; every other RK
  // input is the stencil row itself (S64, or U + S32 with S32 kept)
  constexpr bool LOAD_U = (!MIXED && (STAGE == 2 || STAGE == 3)) || STAGE == 5 || STAGE == 6;
  const double* __restrict__ rk_src = STAGE >= 5 ? a.Q : a.U;
  const int lane = threadIdx.x & (kLanes - 1);
  const int c = bx * kOutCols + lane;  // column of this lane (0-based)
  const int r0 = kBcSize + by * rows;
  const int r1 = min(r0 + rows, p.ny - kBcSize);
  const bool out_col = lane >= kBcSize && lane < kLanes - kBcSize && c <= p.nx - kBcSize - 1;
  const int ccl = min(max(c, 0), p.nx - 1);
  const unsigned uc = (unsigned)ccl;

  bool flx;
  const int cm = map_col(p, ccl, flx);
  const bool wall_w = (p.bc_w == kBcWall) && c == kBcSize;
  const bool wall_e = (p.bc_e == kBcWall) && c == p.nx - kBcSize;

  auto load_raw = [&](int r, double v[4], float inc[4]) {
    bool fly = false;
    int rm = r;
    if (r < kBcSize || r >= p.ny - kBcSize) rm = map_row(p, min(max(r, 0), p.ny - 1), fly);
    load_state<SRC>(p, a.U, a.S64, a.S32, rm, (unsigned)cm, flx, fly, v, inc);
  };
  // u^n at (i, r); r stays within [r0 - 1, r1 - 1] (inside the array).  Halo
  // lanes read the nearest output column of the strip instead of their own
  // (their value is never used): no lines beyond the output columns' are
  // fetched, and no lane is masked.
  const unsigned ucu = (unsigned)min(max(c, bx * kOutCols + kBcSize),
                                     min(bx * kOutCols + kLanes - kBcSize - 1,
                                         p.nx - kBcSize - 1));
  auto load_u = [&](int r, double u[4]) {
    if (MIXED && STAGE >= 5) ld_cell(a.Q32 + (long long)r * p.pitch * 4, ucu, u);
    else ld_cell(rk_src + (long long)r * p.pitch * 4, ucu, u);
  };

  // rings: slot (q - (r0 - 2)) % kRing holds row q
  StripRings<RD, R> sr;
  auto& raw = sr.raw;
  auto& inc = sr.inc;
  auto& rkU = sr.rkU;

  load_raw(r0 - 2, raw[0], inc[0]);
  load_raw(r0 - 1, raw[1], inc[1]);
  load_raw(r0, raw[2], inc[2]);
  load_raw(r0 + 1, raw[3], inc[3]);
  strip_start<RD, R, F32P>(p, sr);

  auto row = [&](auto slot, int r) __attribute__((always_inline)) {
    constexpr int K = decltype(slot)::k;
    constexpr int K1 = (K + 1) % kRing;          // row r+1
    constexpr int K4 = (K + 4) % kRing;          // row r+4
    // prefetch: stencil row r+4 and u^n at (i, r+1).  Rows past the last one
    // the tile reads (r1 + 1, resp. r1 - 1) are clamped to it: the repeated
    // load hits in cache and its value is never used.
    load_raw(min(r + 4, r1 + 1), raw[K4], inc[K4]);
    if (LOAD_U) load_u(min(r + 1, r1 - 1), rkU[K1]);
    R L[4];
    strip_time_dev<RD, R, K, F32P>(p, sr, r, wall_w, wall_e, L);

    const long long row = (long long)r * p.pitch * 4;  // uniform row offset
    if (out_col && r >= r0) {
      const double* S = raw[K];  // the stencil row at (i, r): S64, or u^n + S32
      double vout[4];
      if (STAGE == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) vout[k] = (double)L[k];
        st_cell(a.O64 + row, uc, vout);
      } else if (STAGE >= 4) {
        // low-storage RK3 stage s: q = A_s q + dt L, u = u + B_s q (the state
        // FP64 in both modes; mixed precision keeps q in FP32, the state
        // update taking q before its rounding)
        constexpr int s = STAGE - 4;
        constexpr double LA = s == 0 ? 0.0 : (s == 1 ? -5.0 / 9.0 : -153.0 / 128.0);
        constexpr double LB = s == 0 ? 1.0 / 3.0 : (s == 1 ? 15.0 / 16.0 : 8.0 / 15.0);
        double q[4];
        float q32[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          q[k] = (s == 0 ? 0.0 : LA * rkU[K][k]) + dt * (double)L[k];
          vout[k] = S[k] + LB * q[k];
          q32[k] = (float)q[k];
        }
        st_cell(a.O64 + row, uc, vout);
        if (s < 2) {
          if (MIXED) st_cell(a.Q32 + row, uc, q32);
          else st_cell(a.Q + row, uc, q);
        }
      } else if (!MIXED) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double inc = S[k] + dt * (double)L[k];
          if (STAGE == 1) vout[k] = inc;
          else if (STAGE == 2) vout[k] = 0.75 * rkU[K][k] + (1.0 - 0.75) * inc;
          else vout[k] = (1.0 / 3.0) * rkU[K][k] + (1.0 - 1.0 / 3.0) * inc;
        }
        st_cell(a.O64 + row, uc, vout);
      } else if (STAGE == 1 || STAGE == 2) {
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double dl = dt * (double)L[k];
          o[k] = STAGE == 1 ? (float)dl : (float)(0.25 * ((double)inc[K][k] + dl));
        }
        st_cell(a.O32 + row, uc, o);
      } else {
        // u^n + 2/3 (S32 + dt L) = (u^n + S32) + (2/3 dt L - 1/3 S32), from the
        // stencil row (u^n + S32 rounded once) and its kept increment
#pragma unroll
        for (int k = 0; k < 4; ++k)
          vout[k] = S[k] + ((2.0 / 3.0) * (dt * (double)L[k]) - (1.0 / 3.0) * (double)inc[K][k]);
        st_cell(a.O64 + row, uc, vout);
      }
      if (REDUCE) red.template add<MIXED != 0>(p, vout);
    }
  };

  int r = r0 - 2;
  for (;;) {
    row(Slot<0>{}, r);
    if (++r >= r1) break;
    row(Slot<1>{}, r);
    if (++r >= r1) break;
    row(Slot<2>{}, r);
    if (++r >= r1) break;
    row(Slot<3>{}, r);
    if (++r >= r1) break;
    row(Slot<4>{}, r);
    if (++r >= r1) break;
  }
}

// minimum waves per SIMD the register allocation must allow: 2, i.e. up to 256
// VGPRs (stage 3 with the CFL reduction one fewer).  Tighter bounds make the
// allocator spill to scratch; the 5-row pipeline needs about 200-230 VGPRs.
#ifndef DIM2D_WAVES_FP64
#define DIM2D_WAVES_FP64 2
#endif
#ifndef DIM2D_WAVES_MIXED
#define DIM2D_WAVES_MIXED 2
#endif
// A workgroup holds one strip (64 lanes) or two neighbouring strips (128
// lanes): the two waves of a pair run on one CU, so the 128-byte line that
// holds both strips' shared halo columns is fetched from L2 once.  Waves have
// no barrier between them; a wave past the last strip exits at once.
template <int MIXED, int STAGE>
__global__ void __launch_bounds__(2 * kLanes)
__attribute__((amdgpu_waves_per_eu((MIXED ? DIM2D_WAVES_MIXED : DIM2D_WAVES_FP64) -
                                       (STAGE == 3 || STAGE == 6 ? 1 : 0),
                                   8)))
stage_kernel(const DevParams p, const StageArgs a, int rows) {
  const int bx = blockIdx.x * (int)(blockDim.x / kLanes) + (int)(threadIdx.x / kLanes);
  if (bx * kOutCols >= p.nx - 2 * kBcSize) return;
  // density and its differences stay FP64 in both modes: the capillary
  // second differences of O(1) densities are scaled by 1/dx^2, so FP32
  // densities (absolute rounding ~1e-7) would leave errors of order 1e-7/dx^2
  using RD = double;
  using R = typename std::conditional<MIXED, float, double>::type;
  constexpr bool FIRST = STAGE == 1 || STAGE == 4;  // takes dt, clears the next slot
  constexpr bool LAST = STAGE == 3 || STAGE == 6;    // reduces the new state, t += dt
  double dt = a.dt_fixed;
  if (FIRST) {
    if (a.cfl_mode) dt = cfl_dt_from(a.ctl, a.parity);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      a.ctl->dt = dt;
      clear_slot(a.ctl, a.parity ^ 1);
    }
  } else if (STAGE != 0) {
    dt = a.ctl->dt;
  }
  RedAcc<false> red;
  red.init();
  stage_tile<RD, R, MIXED, STAGE, LAST>(p, a, dt, bx, blockIdx.y, rows, red);
  if (LAST) {
    red.flush(a.ctl, a.parity ^ 1);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.ctl->t += dt;
  }
}

// persistent batch kernel: one grid walks the tile list of all simulations;
// each simulation has a fixed dt (a.S64/U/O pointers are batch bases, offset
// by DevParams::offset); no reductions.
template <int MIXED, int STAGE>
__global__ void __launch_bounds__(kLanes)
__attribute__((amdgpu_waves_per_eu(MIXED ? DIM2D_WAVES_MIXED : DIM2D_WAVES_FP64, 8)))
stage_batch_kernel(const DevParams* __restrict__ P, const WorkItem* __restrict__ items, int nitems,
                   const StageArgs a, const double* __restrict__ dts) {
  // density and its differences stay FP64 in both modes: the capillary
  // second differences of O(1) densities are scaled by 1/dx^2, so FP32
  // densities (absolute rounding ~1e-7) would leave errors of order 1e-7/dx^2
  using RD = double;
  using R = typename std::conditional<MIXED, float, double>::type;
  RedAcc<false> red;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = items[it];
    const DevParams p = P[w.sim];
    StageArgs b = a;
    b.U = a.U ? a.U + p.offset : nullptr;
    b.S64 = a.S64 ? a.S64 + p.offset : nullptr;
    b.S32 = a.S32 ? a.S32 + p.offset : nullptr;
    b.O64 = a.O64 ? a.O64 + p.offset : nullptr;
    b.O32 = a.O32 ? a.O32 + p.offset : nullptr;
    stage_tile<RD, R, MIXED, STAGE, false>(p, b, dts[w.sim], w.bx, w.by, w.rows, red);
  }
}

// ------------------------------------------------------------------------
// wall ghost cells (bc_operators_wall_xy_class.f:232-330 and
// wall_xy_equilibrium_module.f:948-1437, 1747-2290, ridders_method_module.f:112-216)
// ------------------------------------------------------------------------
struct WallRoot {
  double far, first;   // md_x = [far, first] along the wall normal
  double ys0, ys1;     // md_y: transverse neighbours of the first interior cell
  double ux, uy, Ed;
  double dn, dt;       // normal and transverse spacings
  // the parts of the root function that do not depend on the unknown, set
  // by wall_root_prepare: 1/dn, the transverse gradient squared, the
  // temperature terms without the normal gradient and the wall heat flux term
  double idn, gt2, Ta, Tb, Tq;
};

__device__ void wall_root_prepare(const DevParams& p, WallRoot& w) {
  const double md = w.first;
  const double gt = (w.ys1 - w.ys0) / (2.0 * w.dt);
  w.idn = 1.0 / w.dn;
  w.gt2 = gt * gt;
  w.Ta = w.Ed / md - 0.5 * (w.ux * w.ux + w.uy * w.uy) + 3.0 * md;
  w.Tb = 0.5 / p.we / md;
  w.Tq = 0.5 * w.dn * p.re * p.pr * p.c38cv * p.wall_q;
}

__device__ double wall_root_f(const DevParams& p, const WallRoot& w, double x) {
  // wall_x_equilibrium_root_fct(md = first, md_x = [far, x]) (:2151-2236),
  // its terms regrouped around the unknown x (wall_root_prepare)
  const double md = w.first;
  const double dmddx_half = (-md + x) * w.idn;
  const double md_av = 0.5 * (md + x);
  const double gn = (x - w.far) * (0.5 * w.idn);
  const double g2 = gn * gn + w.gt2;
  const double T1 = p.c38cv * (w.Ta - w.Tb * g2);
  const double Tav = T1 + w.Tq;
  // dwallInternalEnergy_dmd (:2238-2293)
  const double sq = sqrt(1.0 - Tav);
  const double ml = 1.0 + 2.06 * sq;
  const double mv = 1.0 - 1.91 * sq;
  const double sigma = p.sigma_pref * (9.691686257306483 * sq * sq * sq);
  const double d = ml - mv;
  const double dw = 6.0 * (md_av - ml) * (md_av - mv) / (d * d * d) * sigma * p.wall_cos_angle;
  return p.inv_we * dmddx_half + dw;
}

__device__ __forceinline__ bool same_sign(double a, double b) {
  return copysign(1.0, a) * copysign(1.0, b) > 0.0;
}
__device__ __forceinline__ bool validated(double v, double c) { return fabs(v - c) < 1e-10; }

__device__ double wall_ghost_density(const DevParams& p, const WallRoot& w, double T_guess,
                                     double md_guess) {
  // get_wall_x_root_brackets (:1846-1941)
  double xl = 1.0 - 1.91 * sqrt(1.0 - T_guess);
  double xh = 1.0 + 2.06 * sqrt(1.0 - T_guess);
  double fl = wall_root_f(p, w, xl);
  double fh = wall_root_f(p, w, xh);
  if (same_sign(fl, fh)) {
    const double dl = md_guess / 100.0, dr = (3.0 - md_guess) / 100.0;
    const double fmd = wall_root_f(p, w, md_guess);
    bool found = false;
    for (int j = 1; j <= 99 && !found; ++j) {
      const double a = md_guess - j * dl, b = md_guess + j * dr;
      const double fa = wall_root_f(p, w, a), fb = wall_root_f(p, w, b);
      if (!same_sign(fa, fmd)) {
        xl = a; fl = fa; xh = md_guess; fh = fmd; found = true;
      } else if (!same_sign(fb, fmd)) {
        xl = md_guess; fl = fmd; xh = b; fh = fb; found = true;
      }
    }
    if (!found) return NAN;
  }
  // get_root_ridder_method with xacc = 1e-12 (default-kind literal)
  const double xacc = (double)1e-12f;
  if ((fl > 0.0 && fh < 0.0) || (fl < 0.0 && fh > 0.0)) {
    double x = (double)(-1.11e30f);
    for (int j = 1; j <= 60; ++j) {
      const double xm = 0.5 * (xl + xh);
      const double fm = wall_root_f(p, w, xm);
      const double s = sqrt(fm * fm - fl * fh);
      if (validated(s, 0.0)) return x;
      const double xnew = xm + (xm - xl) * (copysign(1.0, fl - fh) * fm / s);
      if (fabs(xnew - x) <= xacc) return x;
      x = xnew;
      const double fnew = wall_root_f(p, w, x);
      if (validated(fnew, 0.0)) return x;
      if (!validated(copysign(fm, fnew), fm)) {
        xl = xm; fl = fm; xh = x; fh = fnew;
      } else if (!validated(copysign(fl, fnew), fl)) {
        xh = x; fh = fnew;
      } else if (!validated(copysign(fh, fnew), fh)) {
        xl = x; fl = fnew;
      } else {
        return NAN;
      }
      if (fabs(xh - xl) <= xacc) return x;
    }
    return NAN;
  }
  if (validated(fl, 0.0)) return xl;
  if (validated(fh, 0.0)) return xh;
  return NAN;
}

// state of cell (c, r) (mapped) from an FP64 field or U + increment
__device__ __forceinline__ void wall_load(const DevParams& p, const double* U, const double* S64,
                                          const float* S32, int mixed, int c, int r, double v[4]) {
  bool fx, fy;
  const int cm = map_col(p, c, fx), rm = map_row(p, r, fy);
  const long long idx = ((long long)rm * p.pitch + cm) * 4;
  for (int k = 0; k < 4; ++k)
    v[k] = mixed ? U[idx + k] + (double)S32[idx + k] : S64[idx + k];
  if (fx) v[1] = -v[1];
  if (fy) v[2] = -v[2];
}
__device__ __forceinline__ double rho_at(const DevParams& p, const double* F, int c, int r) {
  bool fx, fy;
  const int cm = map_col(p, c, fx), rm = map_row(p, r, fy);
  return F[((long long)rm * p.pitch + cm) * 4];
}

// temperature_eff (:652-691) at (c, r) of the FP64 field F (for the bracket guess)
__device__ double temperature_eff_at(const DevParams& p, const double* F, int c, int r) {
  double v[4];
  wall_load(p, nullptr, F, nullptr, 0, c, r, v);
  const double gx = 0.5 / p.dx * (-rho_at(p, F, c - 1, r) + rho_at(p, F, c + 1, r));
  const double gy = 0.5 / p.dy * (-rho_at(p, F, c, r - 1) + rho_at(p, F, c, r + 1));
  const double rho = v[0];
  return 1.0 / rho *
         (v[3] - 0.5 * rho * ((v[1] / rho) * (v[1] / rho) + (v[2] / rho) * (v[2] / rho)) -
          0.5 / p.we * (gx * gx + gy * gy) + 3.0 * rho * rho);
}

// The ghost and outer ghost cells of one wall cell (edge e: 0 S, 1 N, 2 W,
// 3 E) from the stage state of the first (v1) and far (v2) interior cells
// along the wall normal, the densities ys0, ys1 of the first interior cell's
// two transverse neighbours, and the bracket guess from u^n:
// compute_wall_{S,N,W,E}_ghost_cell (wall_xy_equilibrium_module.f:948-1437)
// with the outer density of get_wall_md_extra_ghost_cell (:1439-1465).
__device__ __forceinline__ void wall_ghost_values(const DevParams& p, int e, const double v1[4],
                                                  const double v2[4], double ys0, double ys1,
                                                  double T_guess, double md_guess, double g[4],
                                                  double go[4]) {
  WallRoot w;
  w.far = v2[0];
  w.first = v1[0];
  w.Ed = v1[3];
  if (e <= 1) {
    w.ux = (e == 0 ? -v1[2] : v1[2]) / v1[0];  // normal velocity (sign of the S/N cell)
    w.uy = v1[1] / v1[0];
    w.dn = p.dy;
    w.dt = p.dx;
  } else {
    w.ux = (e == 2 ? -v1[1] : v1[1]) / v1[0];
    w.uy = v1[2] / v1[0];
    w.dn = p.dx;
    w.dt = p.dy;
  }
  w.ys0 = ys0;
  w.ys1 = ys1;
  wall_root_prepare(p, w);
  const double rg = wall_ghost_density(p, w, T_guess, md_guess);
  // sign_normal = +1, sign_trans = -1 (wall_xy_equilibrium_module.f:136-140)
  g[0] = rg;
  go[0] = v2[0] + 3.0 * (-v1[0] + rg);  // wall_dmddx_high_order
  if (e <= 1) {
    g[1] = -1.0 * v1[1] / v1[0] * rg;
    g[2] = 1.0 * v1[2] / v1[0] * rg;
  } else {
    g[1] = 1.0 * v1[1] / v1[0] * rg;
    g[2] = -1.0 * v1[2] / v1[0] * rg;
  }
  // the reference leaves the ghost energy and the outer momenta unset; they are
  // never read by an interior stencil.  Keep finite values there.
  g[3] = v1[3];
  go[1] = g[1];
  go[2] = g[2];
  go[3] = v1[3];
}

// one thread per wall cell; edge e: 0 S, 1 N, 2 W, 3 E
__device__ void wall_cell(const DevParams& p, int e, int t, int mixed_in, const double* U,
                          const double* S64, const float* S32, double* O64, float* O32,
                          const double* tmp) {
  int c, r, c1, r1, c2, r2, co, ro;  // ghost, first interior, far interior, outer ghost
  if (e <= 1) {
    c = c1 = c2 = co = kBcSize + t;
    if (e == 0) { r = kBcSize - 1; r1 = r + 1; r2 = r + 2; ro = r - 1; }
    else { r = p.ny - kBcSize; r1 = r - 1; r2 = r - 2; ro = r + 1; }
  } else {
    r = r1 = r2 = ro = kBcSize + t;
    if (e == 2) { c = kBcSize - 1; c1 = c + 1; c2 = c + 2; co = c - 1; }
    else { c = p.nx - kBcSize; c1 = c - 1; c2 = c - 2; co = c + 1; }
  }
  const long long ig = ((long long)r * p.pitch + c) * 4, io = ((long long)ro * p.pitch + co) * 4;
  // u^n at the two ghost cells, for the increments of a mixed-precision
  // output: loaded with the stencil values, not after the root
  double ug[4] = {0.0, 0.0, 0.0, 0.0}, uo[4] = {0.0, 0.0, 0.0, 0.0};
  if (O32) {
    for (int k = 0; k < 4; ++k) {
      ug[k] = U[ig + k];
      uo[k] = U[io + k];
    }
  }
  double v1[4], v2[4];
  wall_load(p, U, S64, S32, mixed_in, c1, r1, v1);
  wall_load(p, U, S64, S32, mixed_in, c2, r2, v2);
  double vy0[4], vy1[4];
  if (e <= 1) {
    wall_load(p, U, S64, S32, mixed_in, c1 - 1, r1, vy0);
    wall_load(p, U, S64, S32, mixed_in, c1 + 1, r1, vy1);
  } else {
    wall_load(p, U, S64, S32, mixed_in, c1, r1 - 1, vy0);
    wall_load(p, U, S64, S32, mixed_in, c1, r1 + 1, vy1);
  }
  // compute_wall_xy_guess from u^n (bc_operators_wall_xy_class.f:898-930)
  const double T_guess = p.c38cv * temperature_eff_at(p, tmp, c1, r1);
  const double md_guess = tmp[((long long)r * p.pitch + c) * 4];
  double g[4], go[4];
  wall_ghost_values(p, e, v1, v2, vy0[0], vy1[0], T_guess, md_guess, g, go);
  for (int k = 0; k < 4; ++k) {
    if (O32) {
      O32[ig + k] = (float)(g[k] - ug[k]);
      O32[io + k] = (float)(go[k] - uo[k]);
    } else {
      O64[ig + k] = g[k];
      O64[io + k] = go[k];
    }
  }
}

__global__ void wall_kernel(const DevParams p, int mixed_in, const double* U, const double* S64,
                            const float* S32, double* O64, float* O32, const double* tmp) {
  const int nS = (p.bc_s == kBcWall) ? p.nx - 2 * kBcSize : 0;
  const int nN = (p.bc_n == kBcWall) ? p.nx - 2 * kBcSize : 0;
  const int nW = (p.bc_w == kBcWall) ? p.ny - 2 * kBcSize : 0;
  const int nE = (p.bc_e == kBcWall) ? p.ny - 2 * kBcSize : 0;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int e;
  if (t < nS) e = 0;
  else if ((t -= nS) < nN) e = 1;
  else if ((t -= nN) < nW) e = 2;
  else if ((t -= nW) < nE) e = 3;
  else return;
  wall_cell(p, e, t, mixed_in, U, S64, S32, O64, O32, tmp);
}

__global__ void wall_batch_kernel(const DevParams* P, int nsim, int mixed_in, const double* U,
                                  const double* S64, const float* S32, double* O64, float* O32,
                                  const double* tmp) {
  const int sim = blockIdx.y;
  if (sim >= nsim) return;
  const DevParams p = P[sim];
  const long long o = p.offset;
  const int nS = (p.bc_s == kBcWall) ? p.nx - 2 * kBcSize : 0;
  const int nN = (p.bc_n == kBcWall) ? p.nx - 2 * kBcSize : 0;
  const int nW = (p.bc_w == kBcWall) ? p.ny - 2 * kBcSize : 0;
  const int nE = (p.bc_e == kBcWall) ? p.ny - 2 * kBcSize : 0;
  for (int t0 = blockIdx.x * blockDim.x + threadIdx.x; t0 < nS + nN + nW + nE;
       t0 += gridDim.x * blockDim.x) {
    int t = t0, e;
    if (t < nS) e = 0;
    else if ((t -= nS) < nN) e = 1;
    else if ((t -= nN) < nW) e = 2;
    else { t -= nW; e = 3; }
    wall_cell(p, e, t, mixed_in, U ? U + o : nullptr, S64 ? S64 + o : nullptr,
              S32 ? S32 + o : nullptr, O64 ? O64 + o : nullptr, O32 ? O32 + o : nullptr,
              tmp + o);
  }
}

// ------------------------------------------------------------------------
// stage pair: RK3-TVD stages 1 and 2 fused (rk3tvd_steps_module.f:52-100,
// 205-269), so the stage-1 state never goes through HBM
// ------------------------------------------------------------------------
// A workgroup of two waves owns a strip of 64 columns, of which the 56 centre
// columns are outputs, and `rows` output rows.  Wave 0, the producer, marches
// up the strip with the stage-1 pipeline from u^n and writes the stage-1 state
// s1 of its 60 valid columns (lanes 2..61) row by row into a ring of LDS rows.
// Wave 1, the consumer, marches kPairLag rows behind with the stage-2
// pipeline, reading s1 from the ring instead of HBM, and writes the stage-2
// state.  One barrier per row keeps them in step; the two waves run on
// different SIMDs of the CU.
// Ghost cells of s1 that the consumer reads:
//  * periodic edges: the producer computes them, its loads reaching two more
//    rows/columns out through the periodic maps, so they are the images'
//    values (same inputs, same instruction stream);
//  * reflection edges: read from the mirror lane/row of the ring with the
//    normal momentum negated, as the stage kernels' virtual ghosts;
//  * S/N walls: the wall root of each ghost cell (wall_ghost_values, as
//    wall_kernel) from s1 of the first and far interior rows, in the tile that
//    holds the wall: the consumer before its first row at the south wall, the
//    producer after its last row at the north wall;
//  * open edges (the 2-cell frame is state): the stage-1 frame kernel's
//    output, read from HBM (pair_gather<MIXED, true>).
// Walls on W/E edges take the separate stage kernels.  The same pair runs the
// low-storage RK3's stages 1 and 2 (LS: q1 rides in the ring's u rows) and,
// persistent over a work list, a batch's (stage_pair_batch_kernel).
// Every output cell sees the same operands in the same order as with the
// separate kernels and the wall kernel between them.
constexpr int kPairOut = kPairOutCols;  // output columns per strip (lanes 4..59)
static_assert(kPairOut == kLanes - 8, "pair strips: 4 halo lanes per side");
constexpr int kPairLag = 6;           // rows the consumer runs behind the producer
constexpr int kPairSlots = 8;         // LDS ring rows (a power of two)

template <int MIXED>
struct PairRing {
  double s[kPairSlots][4][kLanes];                   // s1, one plane per variable
  float d[MIXED ? kPairSlots : 1][4][kLanes];        // mixed: the stage-1 increments
  double u[MIXED ? 1 : kPairSlots][4][kLanes];       // FP64: u^n of the rows, for the
                                                     // consumer's 3/4 u^n term
};

// map_col / map_row reaching two more columns/rows out (c in [-2, nx+1]),
// clamped into the array (the clamped lanes' values are never used)
__device__ __forceinline__ int pair_map_col(const DevParams& p, int c, bool& flip) {
  flip = false;
  int m = c;
  if (c < kBcSize) {
    if (p.bc_w == kBcReflection) { flip = true; m = 2 * kBcSize - 1 - c; }
    else if (p.bc_w == kBcPeriodic) m = c + p.nx - 2 * kBcSize;
  } else if (c >= p.nx - kBcSize) {
    if (p.bc_e == kBcReflection) { flip = true; m = 2 * p.nx - 2 * kBcSize - 1 - c; }
    else if (p.bc_e == kBcPeriodic) m = c - (p.nx - 2 * kBcSize);
  }
  return min(max(m, 0), p.nx - 1);
}
__device__ __forceinline__ int pair_map_row(const DevParams& p, int r, bool& flip) {
  flip = false;
  int m = r;
  if (r < kBcSize) {
    if (p.bc_s == kBcReflection) { flip = true; m = 2 * kBcSize - 1 - r; }
    else if (p.bc_s == kBcPeriodic) m = r + p.ny - 2 * kBcSize;
  } else if (r >= p.ny - kBcSize) {
    if (p.bc_n == kBcReflection) { flip = true; m = 2 * p.ny - 2 * kBcSize - 1 - r; }
    else if (p.bc_n == kBcPeriodic) m = r - (p.ny - 2 * kBcSize);
  }
  return min(max(m, 0), p.ny - 1);
}

// s1 of (row r, lane l) of strip bx from the ring: ghost cells of reflection
// edges come from the mirror lane/row with the normal momentum negated
template <int MIXED, bool OPEN = false>
__device__ __forceinline__ void pair_gather(const DevParams& p, const PairRing<MIXED>& ring,
                                            const StageArgs& a, int bx, int r, int l, double v[4],
                                            float inc[4]) {
  const int cc = kBcSize + bx * kPairOut - 4 + l;
  if (OPEN && (cc < kBcSize || cc >= p.nx - kBcSize || r < kBcSize || r >= p.ny - kBcSize)) {
    // a frame cell of open edges: s1 from the stage-1 frame kernel (a.F64,
    // or u^n + a.F32), which ran before this launch; lanes and rows outside
    // the array read the nearest cell (their values are never used)
    const long long o =
        ((long long)min(max(r, 0), p.ny - 1) * p.pitch + min(max(cc, 0), p.nx - 1)) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (MIXED) {
        inc[k] = a.F32[o + k];
        v[k] = a.U[o + k] + (double)inc[k];
      } else {
        v[k] = a.F64[o + k];
        inc[k] = 0.f;
      }
    }
    return;
  }
  int ll = l;
  bool fx = false, fy = false;
  if (cc < kBcSize && p.bc_w == kBcReflection) {
    ll = l + (2 * kBcSize - 1 - 2 * cc);
    fx = true;
  } else if (cc >= p.nx - kBcSize && p.bc_e == kBcReflection) {
    ll = l + (2 * p.nx - 2 * kBcSize - 1 - 2 * cc);
    fx = true;
  }
  ll = min(max(ll, 0), kLanes - 1);
  int rr = r;
  if (r < kBcSize && p.bc_s == kBcReflection) {
    rr = 2 * kBcSize - 1 - r;
    fy = true;
  } else if (r >= p.ny - kBcSize && p.bc_n == kBcReflection) {
    rr = 2 * p.ny - 2 * kBcSize - 1 - r;
    fy = true;
  }
  const int sl = rr & (kPairSlots - 1);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = ring.s[sl][k][ll];
    inc[k] = MIXED ? ring.d[MIXED ? sl : 0][k][ll] : 0.f;
  }
  if (fx) v[1] = -v[1];
  if (fy) v[2] = -v[2];
}

// The wall ghost rows of s1 of a S (north = false) or N wall into the ring,
// for the lanes whose column lies in [1, nx-2] (the ghost cells the
// consumer's outputs read, and those a reflecting W/E edge mirrors).  The
// values are those wall_kernel stores between the separate stage kernels: in
// mixed precision the ghost state goes through its FP32 increment against
// u^n, as the stage-2 kernel would read it (the low-storage RK3 keeps its
// stage states in FP64 in both modes).  outer: the outer ghost row too (a
// tile whose first gathered row is the inner ghost row needs only that one).
template <int MIXED, bool LS = false>
__device__ void pair_wall_rows(const DevParams& p, const StageArgs& a, PairRing<MIXED>& ring,
                               int bx, bool north, bool outer = true) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int c = kBcSize + bx * kPairOut - 4 + lane;
  const int r = north ? p.ny - kBcSize : kBcSize - 1;  // inner ghost row
  const int r1 = north ? r - 1 : r + 1, r2 = north ? r - 2 : r + 2, ro = north ? r + 1 : r - 1;
  double g[4], go[4];
  if (c >= 1 && c <= p.nx - 2) {
    double v1[4], v2[4], y0[4], y1[4];
    float t[4];
    pair_gather<MIXED>(p, ring, a, bx, r1, lane, v1, t);
    pair_gather<MIXED>(p, ring, a, bx, r2, lane, v2, t);
    pair_gather<MIXED>(p, ring, a, bx, r1, lane - 1, y0, t);
    pair_gather<MIXED>(p, ring, a, bx, r1, lane + 1, y1, t);
    bool fx;
    const int cm = map_col(p, c, fx);
    const double T_guess = p.c38cv * temperature_eff_at(p, a.U, c, r1);
    const double md_guess = a.U[((long long)r * p.pitch + cm) * 4];
    wall_ghost_values(p, north ? 1 : 0, v1, v2, y0[0], y1[0], T_guess, md_guess, g, go);
    if (MIXED && !LS) {
      const double* ug = a.U + ((long long)r * p.pitch + cm) * 4;
      const double* uo = a.U + ((long long)ro * p.pitch + cm) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float dg = (float)(g[k] - ug[k]), dgo = (float)(go[k] - uo[k]);
        g[k] = ug[k] + (double)dg;
        go[k] = uo[k] + (double)dgo;
        ring.d[MIXED ? (r & (kPairSlots - 1)) : 0][k][lane] = dg;
        if (outer) ring.d[MIXED ? (ro & (kPairSlots - 1)) : 0][k][lane] = dgo;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ring.s[r & (kPairSlots - 1)][k][lane] = g[k];
      if (outer) ring.s[ro & (kPairSlots - 1)][k][lane] = go[k];
    }
  }
}

// the producer wave: stage 1 of rows r0-2 .. rlast into the ring, then the
// north wall rows when the tile holds them, then idle barriers up to `steps`
// LS: the low-storage RK3's stages 1 and 2 (the ring's s rows hold u1, its u
// rows q1 in FP64, its d rows q1 in FP32 in mixed precision)
template <typename RD, typename R, int MIXED, bool OPEN, bool LS>
__device__ __forceinline__ void pair_producer(const DevParams& p, const StageArgs& a, double dt,
                                              int bx, int r0, int r1, int steps,
                                              PairRing<MIXED>& ring) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int c = kBcSize + bx * kPairOut - 4 + lane;
  bool flx;
  const int cm = pair_map_col(p, c, flx);
  // open edges: the stage-2 frame kernel reads s1 of the first interior
  // row/column beside each edge, which this wave stores for its own output
  // cells (a.F64, or the increment in a.F32)
  const bool band_col = OPEN && lane >= 4 && lane < kLanes - 4 && c <= p.nx - kBcSize - 1;
  const bool edge_col = c == kBcSize || c == p.nx - kBcSize - 1;
  // The consumer reads s1 up to row r1 + 1, which is a ghost row when
  // r1 + 1 >= ny - 2 (the top tile, and a tile ending one row below it).  The
  // ghost rows of a reflecting or wall edge are the consumer's mirrors and the
  // wall step's; a periodic edge's are the images this wave computes.
  const bool top = r1 + 1 >= p.ny - kBcSize;
  const int rlast = p.bc_n == kBcPeriodic ? r1 + 1 : min(r1 + 1, p.ny - kBcSize - 1);
  const int q0 = r0 - 4;  // two warm-up rows before the first output row r0 - 2
  auto load_raw = [&](int r, double v[4]) {
    bool fly;
    const int rm = pair_map_row(p, r, fly);
    ld_cell(a.U + (long long)rm * p.pitch * 4, (unsigned)cm, v);
    if (flx) v[1] = -v[1];
    if (fly) v[2] = -v[2];
  };
  StripRings<RD, R> sr;
  load_raw(q0, sr.raw[0]);
  load_raw(q0 + 1, sr.raw[1]);
  load_raw(q0 + 2, sr.raw[2]);
  load_raw(q0 + 3, sr.raw[3]);
  strip_start<RD, R, MIXED != 0>(p, sr);
  const int nrows = rlast - q0 + 1;
  int g = 0;
  auto step = [&](auto slot) __attribute__((always_inline)) {
    constexpr int K = decltype(slot)::k;
    constexpr int K4 = (K + 4) % kRing;
    const int r = q0 + g;
    load_raw(min(r + 4, rlast + 2), sr.raw[K4]);
    R L[4];
    strip_time_dev<RD, R, K, MIXED != 0>(p, sr, r, false, false, L);
    if (r >= r0 - 2) {
      const int sl = r & (kPairSlots - 1);
      const double* S = sr.raw[K];
      const bool band = OPEN && band_col && r >= r0 && r < r1 &&
                        (edge_col || r == kBcSize || r == p.ny - kBcSize - 1);
      const long long o = ((long long)r * p.pitch + c) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dl = dt * (double)L[k];
        if (LS) {
          // low-storage stage 1: q1 = dt L, u1 = u0 + B1 q1
          const double q = 0.0 + dl;
          ring.s[sl][k][lane] = S[k] + (1.0 / 3.0) * q;
          if (MIXED) ring.d[MIXED ? sl : 0][k][lane] = (float)q;
          else ring.u[MIXED ? 0 : sl][k][lane] = q;
        } else if (MIXED) {
          const float o32 = (float)dl;
          ring.d[MIXED ? sl : 0][k][lane] = o32;
          ring.s[sl][k][lane] = S[k] + (double)o32;
          if (band) a.F32[o + k] = o32;
        } else {
          ring.s[sl][k][lane] = S[k] + dl;
          ring.u[MIXED ? 0 : sl][k][lane] = S[k];
          if (band) a.F64[o + k] = S[k] + dl;
        }
      }
    }
    __syncthreads();
    return ++g < nrows;
  };
  for (;;) {
    if (!step(Slot<0>{})) break;
    if (!step(Slot<1>{})) break;
    if (!step(Slot<2>{})) break;
    if (!step(Slot<3>{})) break;
    if (!step(Slot<4>{})) break;
  }
  if (top && p.bc_n == kBcWall) {
    pair_wall_rows<MIXED, LS>(p, a, ring, bx, true);
    __syncthreads();
    ++g;
  }
  for (; g < steps; ++g) __syncthreads();
}

// the consumer wave: kPairLag idle barriers, the south wall rows when the
// tile holds them, then stage 2 of rows r0 .. r1-1 from the ring
template <typename RD, typename R, int MIXED, bool OPEN, bool LS>
__device__ __forceinline__ void pair_consumer(const DevParams& p, const StageArgs& a, double dt,
                                              int bx, int r0, int r1, PairRing<MIXED>& ring) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int c0 = kBcSize + bx * kPairOut;  // first output column of the strip
  const int c = c0 - 4 + lane;
  const bool out_col = lane >= 4 && lane < kLanes - 4 && c <= p.nx - kBcSize - 1;
  const unsigned uc = (unsigned)min(max(c, 0), p.nx - 1);
  for (int g = 0; g < kPairLag; ++g) __syncthreads();
  if (r0 - 2 < kBcSize && p.bc_s == kBcWall) {
    // the wall ghost rows this tile gathers: rows 0 and 1, or row 1 alone
    // for a tile starting at row 3 (batch tiles of one row)
    pair_wall_rows<MIXED, LS>(p, a, ring, bx, false, r0 == kBcSize);
    __threadfence_block();
  }
  const int q0 = r0 - 2;
  StripRings<RD, R> sr;
  pair_gather<MIXED, OPEN>(p, ring, a, bx, q0, lane, sr.raw[0], sr.inc[0]);
  pair_gather<MIXED, OPEN>(p, ring, a, bx, q0 + 1, lane, sr.raw[1], sr.inc[1]);
  strip_start<RD, R, MIXED != 0>(p, sr);
  int r = q0;
  auto row = [&](auto slot) __attribute__((always_inline)) {
    constexpr int K = decltype(slot)::k;
    constexpr int K1 = (K + 1) % kRing;
    constexpr int K2 = (K + 2) % kRing;
    // s1 of row r+2 from the ring (written at least one barrier ago)
    pair_gather<MIXED, OPEN>(p, ring, a, bx, r + 2, lane, sr.raw[K2], sr.inc[K2]);
    R L[4];
    strip_time_dev<RD, R, K, MIXED != 0>(p, sr, r, false, false, L);
    if (out_col && r >= r0) {
      const long long row = (long long)r * p.pitch * 4;
      const double* S = sr.raw[K];
      if (LS) {
        // low-storage stage 2: q2 = A2 q1 + dt L, u2 = u1 + B2 q2
        double vout[4], q[4];
        float q32[4];
        const int sl = r & (kPairSlots - 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double q1 = MIXED ? (double)ring.d[MIXED ? sl : 0][k][lane]
                                  : ring.u[MIXED ? 0 : sl][k][lane];
          q[k] = (-5.0 / 9.0) * q1 + dt * (double)L[k];
          vout[k] = S[k] + (15.0 / 16.0) * q[k];
          q32[k] = (float)q[k];
        }
        st_cell(a.O64 + row, uc, vout);
        if (MIXED) st_cell(a.Q32 + row, uc, q32);
        else st_cell(a.Q + row, uc, q);
      } else if (!MIXED) {
        double vout[4];
        const int sl = r & (kPairSlots - 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double inc = S[k] + dt * (double)L[k];
          vout[k] = 0.75 * ring.u[MIXED ? 0 : sl][k][lane] + (1.0 - 0.75) * inc;
        }
        st_cell(a.O64 + row, uc, vout);
      } else {
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double dl = dt * (double)L[k];
          o[k] = (float)(0.25 * ((double)sr.inc[K][k] + dl));
        }
        st_cell(a.O32 + row, uc, o);
      }
    }
    __syncthreads();
    return ++r < r1;
  };
  for (;;) {
    if (!row(Slot<0>{})) break;
    if (!row(Slot<1>{})) break;
    if (!row(Slot<2>{})) break;
    if (!row(Slot<3>{})) break;
    if (!row(Slot<4>{})) break;
  }
}

// One tile of the stage pair: strip bx, output rows r0 .. r1-1 of simulation
// p.  Both waves pass exactly kPairLag + (r1 - r0 + 2) barriers, and the
// consumer's last ring read precedes its last barrier, so a workgroup may run
// tiles back to back on one ring.
template <int MIXED, bool OPEN = false, bool LS = false>
__device__ __forceinline__ void pair_tile(const DevParams& p, const StageArgs& a, double dt,
                                          int bx, int r0, int r1, bool swap,
                                          PairRing<MIXED>& ring) {
  using RD = double;
  using R = typename std::conditional<MIXED, float, double>::type;
  static_assert(!(LS && OPEN), "the low-storage pair runs without open edges");
  // the consumer's rows r0-2 .. r1-1 after kPairLag idle barriers
  const int steps = kPairLag + (r1 - r0 + 2);
  if ((threadIdx.x < kLanes) != swap)
    pair_producer<RD, R, MIXED, OPEN, LS>(p, a, dt, bx, r0, r1, steps, ring);
  else
    pair_consumer<RD, R, MIXED, OPEN, LS>(p, a, dt, bx, r0, r1, ring);
}

// minimum waves per SIMD: 2 in FP64 (up to 256 VGPRs), as the stage kernels;
// 3 in mixed precision (168 VGPRs; its LDS ring fits 6 workgroups per CU).  OPEN:
// all four edges open (hedstrom_xy); the frame cells of s1 come from the
// stage-1 frame kernel, launched before this kernel, and the stage-2 frame
// kernel, launched after it, reads the band of s1 this kernel stores.  LS:
// the low-storage RK3 (no open edges): u0 (a.U) -> u2 (a.O64), q2 (a.Q, or
// a.Q32 in mixed precision).
template <int MIXED, bool OPEN, bool LS>
__global__ void __launch_bounds__(2 * kLanes)
__attribute__((amdgpu_waves_per_eu(MIXED ? 3 : 2, 8)))
stage_pair_kernel(const DevParams p, const StageArgs a, int rows) {
  __shared__ PairRing<MIXED> ring;
  double dt = a.dt_fixed;
  if (a.cfl_mode) dt = cfl_dt_from(a.ctl, a.parity);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    a.ctl->dt = dt;
    clear_slot(a.ctl, a.parity ^ 1);
  }
  const int r0 = kBcSize + blockIdx.y * rows;
  const int r1 = min(r0 + rows, p.ny - kBcSize);
  // FP64 (2 waves per SIMD): the roles of the two waves alternate between
  // neighbouring workgroups, so that each SIMD of a CU holds producers and
  // consumers alike rather than two of one kind (measured 4 % faster at
  // 4096^2; in mixed precision, 3 waves per SIMD, 9 % slower)
  const bool swap = !MIXED && (((blockIdx.x + blockIdx.y) & 1) != 0);
  pair_tile<MIXED, OPEN, LS>(p, a, dt, blockIdx.x, r0, r1, swap, ring);
}

// persistent batch form of the stage pair: one grid walks the pair tiles
// (56-column strips) of all simulations, each with its fixed dt; a.U and
// a.O64/a.O32 are batch bases offset by DevParams::offset
#ifndef DIM2D_PAIR_BATCH_WAVES_MIXED
#define DIM2D_PAIR_BATCH_WAVES_MIXED 3
#endif
template <int MIXED>
__global__ void __launch_bounds__(2 * kLanes)
__attribute__((amdgpu_waves_per_eu(MIXED ? DIM2D_PAIR_BATCH_WAVES_MIXED : 2, 8)))
stage_pair_batch_kernel(const DevParams* __restrict__ P, const WorkItem* __restrict__ items,
                        int nitems, const StageArgs a, const double* __restrict__ dts) {
  __shared__ PairRing<MIXED> ring;
  const bool swap = !MIXED && (blockIdx.x & 1) != 0;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = items[it];
    const DevParams p = P[w.sim];
    StageArgs b = a;
    b.U = a.U + p.offset;
    b.O64 = a.O64 ? a.O64 + p.offset : nullptr;
    b.O32 = a.O32 ? a.O32 + p.offset : nullptr;
    const int r0 = kBcSize + w.by * w.rows;
    const int r1 = min(r0 + w.rows, p.ny - kBcSize);
    pair_tile<MIXED>(p, b, dts[w.sim], w.bx, r0, r1, swap, ring);
  }
}

// The mixed launch of a pipelined batch: the simulations form two groups half
// a step apart, so one launch runs the stage pairs of one group (bound by
// instruction issue) beside the stage 3 of the other (bound by HBM), and the
// two kinds of work share the CUs.  Both waves of a workgroup pass the same
// barriers per item (stage-3 items have none), so a wave that finishes a
// stage-3 strip early simply waits at the next pair item's first barrier.
#ifndef DIM2D_MIX_WAVES_MIXED
#define DIM2D_MIX_WAVES_MIXED 2
#endif
template <int MIXED>
__global__ void __launch_bounds__(2 * kLanes)
__attribute__((amdgpu_waves_per_eu(MIXED ? DIM2D_MIX_WAVES_MIXED : 2, 8)))
stage_mix_batch_kernel(const DevParams* __restrict__ P, const WorkItem* __restrict__ items,
                       int nitems, const MixArgs m, const double* __restrict__ dts) {
  __shared__ PairRing<MIXED> ring;
  using RD = double;
  using R = typename std::conditional<MIXED, float, double>::type;
  const bool swap = !MIXED && (blockIdx.x & 1) != 0;
  RedAcc<false> red;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = items[it];
    const DevParams p = P[w.sim];
    const double dt = dts[w.sim];
    StageArgs b;
    memset(&b, 0, sizeof b);
    if (w.kind == kItemPair) {
      b.U = m.pU + p.offset;
      if (MIXED) b.O32 = m.pO32 + p.offset;
      else b.O64 = m.pO64 + p.offset;
      const int r0 = kBcSize + w.by * w.rows;
      const int r1 = min(r0 + w.rows, p.ny - kBcSize);
      pair_tile<MIXED>(p, b, dt, w.bx, r0, r1, swap, ring);
    } else {
      const int strip = 2 * w.bx + (int)(threadIdx.x / kLanes);
      if (strip * kOutCols >= p.nx - 2 * kBcSize) continue;
      b.U = m.sU + p.offset;
      if (MIXED) b.S32 = m.sS32 + p.offset;
      else b.S64 = m.sS64 + p.offset;
      b.O64 = m.sO64 + p.offset;
      stage_tile<RD, R, MIXED, 3, false>(p, b, dt, strip, w.by, w.rows, red);
    }
  }
}

// ------------------------------------------------------------------------
// open boundaries (hedstrom_xy, bc_operators_hedstrom_xy_class.f:182-496)
// ------------------------------------------------------------------------
// The 2-cell ghost frame of a domain with open edges is state, advanced by the
// same RK stages with its own time derivative (borders widened as
// field_abstract_class.f:244-270 does):
//   N/S edge cells  (1/dx)(Fx(i-1/2) - Fx(i+1/2)) + characteristic y-term + body
//   E/W edge cells  characteristic x-term + (1/dy)(Fy(j-1/2) - Fy(j+1/2)) + body
//   corner cells    characteristic x- and y-terms + body
// The fluxes of the inner ghost row/column use the interior Mattsson
// stencils, those of the outer one the one-sided stencils of
// sd_operators_{y,x}_oneside_{L0,R0}_class.f (bc_operators_openbc_normal_class.f:563-787);
// the characteristic term is hedstrom_xy_module.f:445-558 with the
// eigen-quantities of the point itself (obc_eigenqties_bc).  One thread per
// frame cell evaluates everything pointwise, in FP64, written as plainly as
// the reference; every stencil point is read from a window of the stage input
// that the workgroup stages in LDS first (a frame cell reads ~250 stencil
// points, so reading them from HBM leaves each thread a chain of dependent
// loads).  Indices are 0-based.
namespace frame {

enum Var { kInt = 0, kYL0, kYR0, kXL0, kXR0 };  // stencil set of a face
enum Grad { kGInt = 0, kGL0, kGR0 };            // gradient kind
enum Prim {
  kRho, kQx, kQy, kUx, kUy, kP, kTcl, kQxTx, kQyTx, kQxTy, kQyTy, kEUx, kEUy, kPxw, kPyw,
  kCap, kCapx, kCapy
};

// stage input of cell (i, j): S64, U, or U + S32
__device__ __forceinline__ void load_input(const DevParams& p, const StageArgs& a, int src, int i,
                                           int j, double v[4]) {
  const long long o = ((long long)j * p.pitch + i) * 4;
  for (int k = 0; k < 4; ++k)
    v[k] = src == 0 ? a.S64[o + k] : (src == 1 ? a.U[o + k] : a.U[o + k] + (double)a.S32[o + k]);
}

constexpr int kNumPrim = kCapy + 1;  // point functions per cell

// the stage input over the window [ci0, ci0 + w) x [rj0, ...) held in LDS, and
// the point functions of its cells (tab[pk * stride + cell])
struct View {
  const DevParams* p;
  const double* win;
  const double* tab;
  int ci0, rj0, w, stride;
  __device__ int at(int i, int j) const { return (j - rj0) * w + (i - ci0); }
  __device__ void cell(int i, int j, double v[4]) const {
    const double* c = win + at(i, j) * 4;
    for (int k = 0; k < 4; ++k) v[k] = c[k];
  }
};

// classical_pressure_local (dim2d_prim_module.f:425-453)
__device__ double pressure_of(const DevParams& p, const double v[4]) {
  const double r = v[0];
  return 3.0 / ((3.0 - r) * p.cv) *
             (v[3] - 0.5 * r * ((v[1] / r) * (v[1] / r) + (v[2] / r) * (v[2] / r)) + 3.0 * r * r) -
         3.0 * r * r;
}

// the point functions the fluxes take (dim2d_prim_module.f:99-1150)
__device__ double point_function(const DevParams& p, const double v[4], int pk) {
  const double r = v[0];
  switch (pk) {
    case kRho: return r;
    case kQx: return v[1];
    case kQy: return v[2];
    case kUx: return v[1] / r;
    case kUy: return v[2] / r;
    case kP: return pressure_of(p, v);
    case kTcl: return 1.0 / r * (v[3] - 0.5 * r * ((v[1] / r) * (v[1] / r) + (v[2] / r) * (v[2] / r)) + 3.0 * r * r);
    case kQxTx: return v[1] * v[1] / r;
    case kQyTx: return v[2] * v[1] / r;
    case kQxTy: return v[1] * v[2] / r;
    case kQyTy: return v[2] * v[2] / r;
    case kEUx: return v[3] * v[1] / r;
    case kEUy: return v[3] * v[2] / r;
    case kPxw: return pressure_of(p, v) * v[1] / r;
    case kPyw: return pressure_of(p, v) * v[2] / r;
    case kCap: return 1.0 / (3.0 - r);
    case kCapx: return v[1] / (r * (3.0 - r));
    default: return v[2] / (r * (3.0 - r));
  }
}
// point function pk at (i, j), from the window's table (each one computed
// once per cell, where the stencils read it ~250 times per frame cell)
__device__ __forceinline__ double prim(const View& f, int i, int j, int pk) {
  return f.tab[pk * f.stride + f.at(i, j)];
}

// gradients of a point function (sd_operators_fd_module.f:71-450)
__device__ double grad_x(const View& f, int i, int j, int pk, double dx, int g) {
  if (g == kGL0) return 1.0 / dx * (-prim(f, i, j, pk) + prim(f, i + 1, j, pk));
  if (g == kGR0) return 1.0 / dx * (-prim(f, i - 1, j, pk) + prim(f, i, j, pk));
  return 0.5 / dx * (-prim(f, i - 1, j, pk) + prim(f, i + 1, j, pk));
}
__device__ double grad_y(const View& f, int i, int j, int pk, double dy, int g) {
  if (g == kGL0) return 1.0 / dy * (-prim(f, i, j, pk) + prim(f, i, j + 1, pk));
  if (g == kGR0) return 1.0 / dy * (-prim(f, i, j - 1, pk) + prim(f, i, j, pk));
  return 0.5 / dy * (-prim(f, i, j - 1, pk) + prim(f, i, j + 1, pk));
}
// capillarity_temperature_eff (:582-613) with the given gradient kinds
__device__ double tcap(const View& f, int i, int j, int gx, int gy) {
  const double a = grad_x(f, i, j, kRho, f.p->dx, gx);
  const double b = grad_y(f, i, j, kRho, f.p->dy, gy);
  return 0.5 / prim(f, i, j, kRho) * (a * a + b * b);
}

// x-face stencils at (i-1/2, j) (sd_operators_class.f:185-549, one-sided
// sets sd_operators_y_oneside_{L0,R0}_class.f)
__device__ double st_f(const View& f, int i, int j, int pk) {
  return 0.5 * (+prim(f, i - 1, j, pk) + prim(f, i, j, pk));
}
__device__ double st_dfdx(const View& f, int i, int j, int pk) {
  return 1.0 / f.p->dx * (-prim(f, i - 1, j, pk) + prim(f, i, j, pk));
}
__device__ double st_dfdy(const View& f, int i, int j, int pk, int var) {
  const double dy = f.p->dy;
  if (var == kYL0)
    return 1.0 / dy * (-0.5 * prim(f, i - 1, j, pk) - 0.5 * prim(f, i, j, pk) +
                       0.5 * prim(f, i - 1, j + 1, pk) + 0.5 * prim(f, i, j + 1, pk));
  if (var == kYR0)
    return 1.0 / dy * (-0.5 * prim(f, i - 1, j - 1, pk) - 0.5 * prim(f, i, j - 1, pk) +
                       0.5 * prim(f, i - 1, j, pk) + 0.5 * prim(f, i, j, pk));
  return 1.0 / (2.0 * dy) * (-0.5 * prim(f, i - 1, j - 1, pk) - 0.5 * prim(f, i, j - 1, pk) +
                             0.5 * prim(f, i - 1, j + 1, pk) + 0.5 * prim(f, i, j + 1, pk));
}
__device__ double st_d2fdx2(const View& f, int i, int j, int pk) {
  const double dx = f.p->dx;
  return 0.5 / (dx * dx) *
         (+prim(f, i - 2, j, pk) - prim(f, i - 1, j, pk) - prim(f, i, j, pk) + prim(f, i + 1, j, pk));
}
__device__ double st_d2fdy2(const View& f, int i, int j, int pk, int var) {
  const double dy = f.p->dy;
  const int j0 = var == kYL0 ? j : (var == kYR0 ? j - 2 : j - 1);  // rows j0, j0+1, j0+2
  return 1.0 / (dy * dy) *
         (+0.5 * prim(f, i - 1, j0, pk) + 0.5 * prim(f, i, j0, pk) -
          2.0 * (+0.5 * prim(f, i - 1, j0 + 1, pk) + 0.5 * prim(f, i, j0 + 1, pk)) +
          0.5 * prim(f, i - 1, j0 + 2, pk) + 0.5 * prim(f, i, j0 + 2, pk));
}
__device__ double st_dfdx_nl(const View& f, int i, int j, int var) {
  const int gy = var == kYL0 ? kGL0 : (var == kYR0 ? kGR0 : kGInt);
  return 1.0 / f.p->dx * (-tcap(f, i - 1, j, kGInt, gy) + tcap(f, i, j, kGInt, gy));
}
// y-face stencils at (i, j-1/2) (:622-1003, sd_operators_x_oneside_{L0,R0}_class.f)
__device__ double st_g(const View& f, int i, int j, int pk) {
  return 0.5 * (+prim(f, i, j - 1, pk) + prim(f, i, j, pk));
}
__device__ double st_dgdy(const View& f, int i, int j, int pk) {
  return 1.0 / f.p->dy * (-prim(f, i, j - 1, pk) + prim(f, i, j, pk));
}
__device__ double st_dgdx(const View& f, int i, int j, int pk, int var) {
  const double dx = f.p->dx;
  if (var == kXL0)
    return 1.0 / dx * (-0.5 * prim(f, i, j - 1, pk) - 0.5 * prim(f, i, j, pk) +
                       0.5 * prim(f, i + 1, j - 1, pk) + 0.5 * prim(f, i + 1, j, pk));
  if (var == kXR0)
    return 1.0 / dx * (-0.5 * prim(f, i - 1, j - 1, pk) - 0.5 * prim(f, i - 1, j, pk) +
                       0.5 * prim(f, i, j - 1, pk) + 0.5 * prim(f, i, j, pk));
  return 0.25 / dx * (-prim(f, i - 1, j - 1, pk) - prim(f, i - 1, j, pk) +
                      prim(f, i + 1, j - 1, pk) + prim(f, i + 1, j, pk));
}
__device__ double st_d2gdy2(const View& f, int i, int j, int pk) {
  const double dy = f.p->dy;
  return 0.5 / (dy * dy) *
         (+prim(f, i, j - 2, pk) - prim(f, i, j - 1, pk) - prim(f, i, j, pk) + prim(f, i, j + 1, pk));
}
__device__ double st_d2gdx2(const View& f, int i, int j, int pk, int var) {
  const double dx = f.p->dx;
  if (var == kXL0 || var == kXR0) {
    const int i0 = var == kXL0 ? i : i - 2;  // columns i0, i0+1, i0+2
    return 1.0 / (dx * dx) *
           (0.5 * prim(f, i0, j - 1, pk) + 0.5 * prim(f, i0, j, pk) -
            2.0 * (+0.5 * prim(f, i0 + 1, j - 1, pk) + 0.5 * prim(f, i0 + 1, j, pk)) +
            0.5 * prim(f, i0 + 2, j - 1, pk) + 0.5 * prim(f, i0 + 2, j, pk));
  }
  return 0.5 / (dx * dx) *
         (+prim(f, i - 1, j - 1, pk) + prim(f, i - 1, j, pk) -
          2.0 * (+prim(f, i, j - 1, pk) + prim(f, i, j, pk)) + prim(f, i + 1, j - 1, pk) +
          prim(f, i + 1, j, pk));
}
__device__ double st_dgdy_nl(const View& f, int i, int j, int var) {
  const int gx = var == kXL0 ? kGL0 : (var == kXR0 ? kGR0 : kGInt);
  return 1.0 / f.p->dy * (-tcap(f, i, j - 1, gx, kGInt) + tcap(f, i, j, gx, kGInt));
}

// the four x-face fluxes (dim2d_fluxes_module.f:103-1649): inviscid - visc/Re - cap/We
__device__ void flux_x(const View& f, int i, int j, int var, double F[4]) {
  const DevParams& p = *f.p;
  const double vr = p.vr;
  F[0] = st_f(f, i, j, kQx);
  {
    const double inv = st_f(f, i, j, kQxTx) + st_f(f, i, j, kP);
    const double vis = (2.0 + vr) * st_dfdx(f, i, j, kUx) + vr * st_dfdy(f, i, j, kUy, var);
    const double rx = st_dfdx(f, i, j, kRho), ry = st_dfdy(f, i, j, kRho, var);
    const double cap = 1.5 / p.cv * st_f(f, i, j, kCap) * (rx * rx + ry * ry) +
                       st_f(f, i, j, kRho) * (st_d2fdx2(f, i, j, kRho) + st_d2fdy2(f, i, j, kRho, var)) +
                       0.5 * (-rx * rx + ry * ry);
    F[1] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
  {
    const double inv = st_f(f, i, j, kQyTx);
    const double vis = st_dfdy(f, i, j, kUx, var) + st_dfdx(f, i, j, kUy);
    const double cap = -st_dfdx(f, i, j, kRho) * st_dfdy(f, i, j, kRho, var);
    F[2] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
  {
    const double inv = st_f(f, i, j, kEUx) + st_f(f, i, j, kPxw);
    const double ux = st_f(f, i, j, kUx), uy = st_f(f, i, j, kUy);
    const double vis = ux * ((2.0 + vr) * st_dfdx(f, i, j, kUx) + vr * st_dfdy(f, i, j, kUy, var)) +
                       uy * (st_dfdy(f, i, j, kUx, var) + st_dfdx(f, i, j, kUy)) +
                       1.0 / p.pr * st_dfdx(f, i, j, kTcl);
    const double duxdx = st_dfdx(f, i, j, kUx), duydy = st_dfdy(f, i, j, kUy, var);
    const double rx = st_dfdx(f, i, j, kRho), ry = st_dfdy(f, i, j, kRho, var);
    const double cap = 1.5 / p.cv * st_f(f, i, j, kCapx) * (rx * rx + ry * ry) +
                       st_f(f, i, j, kQx) * (st_d2fdx2(f, i, j, kRho) + st_d2fdy2(f, i, j, kRho, var)) +
                       0.5 * ux * (-rx * rx + ry * ry) - rx * ry * uy -
                       st_f(f, i, j, kRho) * rx * (duxdx + duydy) -
                       1.0 / (p.re * p.pr) * st_dfdx_nl(f, i, j, var);
    F[3] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
}
// the four y-face fluxes
__device__ void flux_y(const View& f, int i, int j, int var, double G[4]) {
  const DevParams& p = *f.p;
  const double vr = p.vr;
  G[0] = st_g(f, i, j, kQy);
  {
    const double inv = st_g(f, i, j, kQxTy);
    const double vis = st_dgdy(f, i, j, kUx) + st_dgdx(f, i, j, kUy, var);
    const double cap = -st_dgdx(f, i, j, kRho, var) * st_dgdy(f, i, j, kRho);
    G[1] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
  {
    const double inv = st_g(f, i, j, kQyTy) + st_g(f, i, j, kP);
    const double vis = vr * st_dgdx(f, i, j, kUx, var) + (2.0 + vr) * st_dgdy(f, i, j, kUy);
    const double rx = st_dgdx(f, i, j, kRho, var), ry = st_dgdy(f, i, j, kRho);
    const double cap = 1.5 / p.cv * st_g(f, i, j, kCap) * (rx * rx + ry * ry) +
                       st_g(f, i, j, kRho) * (st_d2gdx2(f, i, j, kRho, var) + st_d2gdy2(f, i, j, kRho)) +
                       0.5 * (rx * rx - ry * ry);
    G[2] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
  {
    const double inv = st_g(f, i, j, kEUy) + st_g(f, i, j, kPyw);
    const double ux = st_g(f, i, j, kUx), uy = st_g(f, i, j, kUy);
    const double duxdx = st_dgdx(f, i, j, kUx, var), duydy = st_dgdy(f, i, j, kUy);
    const double vis = ux * (st_dgdy(f, i, j, kUx) + st_dgdx(f, i, j, kUy, var)) +
                       uy * (vr * duxdx + (2.0 + vr) * duydy) +
                       1.0 / p.pr * st_dgdy(f, i, j, kTcl);
    const double rx = st_dgdx(f, i, j, kRho, var), ry = st_dgdy(f, i, j, kRho);
    const double cap = 1.5 / p.cv * st_g(f, i, j, kCapy) * (rx * rx + ry * ry) +
                       st_g(f, i, j, kQy) * (st_d2gdx2(f, i, j, kRho, var) + st_d2gdy2(f, i, j, kRho)) +
                       0.5 * uy * (rx * rx - ry * ry) - rx * ry * ux -
                       st_g(f, i, j, kRho) * ry * (duxdx + duydy) -
                       1.0 / (p.re * p.pr) * st_dgdy_nl(f, i, j, var);
    G[3] = inv - 1.0 / p.re * vis - 1.0 / p.we * cap;
  }
}

// compute_timedev_with_openbc (hedstrom_xy_module.f:445-558): amplitudes of
// the outgoing characteristics from one-sided primitive gradients, incoming
// ones zero (openbc_operators_module.f:97-136), mapped back through the right
// eigenvectors and the primitive-to-conservative Jacobian
// (ns_vdw2d_prim_module.f:249-891); eigen-quantities of the point itself.
__device__ void openbc(const View& f, int i, int j, int ydir, int right, double out[4]) {
  const DevParams& p = *f.p;
  double v[4];
  f.cell(i, j, v);
  const double md = v[0], ux = v[1] / v[0], uy = v[2] / v[0];
  const double P = pressure_of(p, v);
  // speed_of_sound (dim2d_prim_module.f:1179-1215)
  const double e = P + md * md * (-3.0 + 2.0 * md);
  const double b = sqrt((P + 3.0 * md * md) / (p.cv * e));
  const double a = 1.0 / sqrt(1.0 + b * b);
  const double d = sqrt(md * md * md) * sqrt((3.0 - md) / e);
  const double c = sqrt(3.0) * md / (a * d);
  const double un = ydir ? uy : ux;
  const double lam[4] = {un, un, un - c, un + c};
  const double q_c = 0.5 * md * c;
  const int g = right ? kGR0 : kGL0;
  const double dn = ydir ? p.dy : p.dx;
  double gr[4];
  for (int k = 0; k < 4; ++k) {
    const int pk = k == 0 ? kRho : (k == 1 ? kUx : (k == 2 ? kUy : kP));
    gr[k] = ydir ? grad_y(f, i, j, pk, dn, g) : grad_x(f, i, j, pk, dn, g);
  }
  // left eigenvectors l_k (columns of the reference's left eigenmatrix)
  const int t = ydir ? 2 : 1;  // velocity normal to the edge in (rho, ux, uy, P)
  const int s = ydir ? 1 : 2;  // transverse velocity
  double amp[4];
  for (int k = 0; k < 4; ++k) {
    const bool incoming = right ? (lam[k] <= 0.0) : (lam[k] >= 0.0);
    double dot;
    if (k == 0) dot = gr[s];
    else if (k == 1) dot = gr[0] + (-1.0 / (c * c)) * gr[3];
    else if (k == 2) dot = -q_c * gr[t] + 0.5 * gr[3];
    else dot = q_c * gr[t] + 0.5 * gr[3];
    amp[k] = incoming ? 0.0 : -lam[k] * dot;
  }
  // primitive time derivative: sum_k amp_k r_k
  const double inv_c = 1.0 / (c * c), inv_q_c = 1.0 / (md * c);
  double tp[4];
  tp[0] = amp[1] + amp[2] * inv_c + amp[3] * inv_c;
  tp[t] = amp[2] * -inv_q_c + amp[3] * inv_q_c;
  tp[s] = amp[0];
  tp[3] = amp[2] + amp[3];
  // conservative: J(l, m) = d cons_m / d prim_l
  const double dEdrho = 0.5 * (ux * ux + uy * uy) - 6.0 * md -
                        1.0 / 3.0 * p.cv * (P + 9.0 * (md - 2.0) * md);
  out[0] = tp[0];
  out[1] = tp[0] * ux + tp[1] * md;
  out[2] = tp[0] * uy + tp[2] * md;
  out[3] = tp[0] * dEdrho + tp[1] * (md * ux) + tp[2] * (md * uy) + tp[3] * (p.cv * (1.0 - md / 3.0));
}

__device__ double body(const DevParams& p, const double v[4], int k) {
  if (!p.gravity_ac) return 0.0;
  if (k == 2) return -v[0] * p.gravity;
  if (k == 3) return -v[2] * p.gravity;
  return 0.0;
}

// time derivative of frame cell (i, j)
__device__ void timedev(const View& f, int i, int j, double L[4]) {
  const DevParams& p = *f.p;
  const int nx = p.nx, ny = p.ny;
  double v[4];
  f.cell(i, j, v);
  const bool xg = i < kBcSize || i >= nx - kBcSize;
  const bool yg = j < kBcSize || j >= ny - kBcSize;
  double o[4], F[4], F2[4];
  if (xg && yg) {  // corner
    double oy[4];
    openbc(f, i, j, 0, i >= nx - kBcSize, o);
    openbc(f, i, j, 1, j >= ny - kBcSize, oy);
    for (int k = 0; k < 4; ++k) L[k] = o[k] + oy[k] + body(p, v, k);
  } else if (yg) {  // N or S edge
    const int var = j == 0 ? kYL0 : (j == ny - 1 ? kYR0 : kInt);
    flux_x(f, i, j, var, F);
    flux_x(f, i + 1, j, var, F2);
    openbc(f, i, j, 1, j >= ny - kBcSize, o);
    for (int k = 0; k < 4; ++k) L[k] = 1.0 / p.dx * (F[k] - F2[k]) + o[k] + body(p, v, k);
  } else {  // E or W edge
    const int var = i == 0 ? kXL0 : (i == nx - 1 ? kXR0 : kInt);
    flux_y(f, i, j, var, F);
    flux_y(f, i, j + 1, var, F2);
    openbc(f, i, j, 0, i >= nx - kBcSize, o);
    for (int k = 0; k < 4; ++k) L[k] = o[k] + 1.0 / p.dy * (F[k] - F2[k]) + body(p, v, k);
  }
}

}  // namespace frame

// Workgroups of 128 threads, each two frame lines of 64 cells:
//   S, N: rows {0, 1} / {ny-2, ny-1}, columns i0 .. i0+63 of [2, nx-3];
//         window columns i0-2 .. i0+65, rows 0..2 / ny-3..ny-1
//   W, E: columns {0, 1} / {nx-2, nx-1}, rows j0 .. j0+63 of [2, ny-3];
//         window rows j0-2 .. j0+65, columns 0..2 / nx-3..nx-1
//   the last workgroup: the 16 corner cells, four 3x3 windows
// The stencils of a frame cell reach two cells along its edge and stay within
// the three outer rows (columns) across it, so every point it reads lies in
// its window.
constexpr int kFrameSeg = 64;
constexpr int kFrameWin = (kFrameSeg + 4) * 3;  // cells per window

template <int MIXED, int STAGE>
__global__ void __launch_bounds__(2 * kFrameSeg) frame_kernel(const DevParams p, const StageArgs a) {
  __shared__ double win[kFrameWin * 4];
  __shared__ double tab[frame::kNumPrim * kFrameWin];
  constexpr int SRC = (MIXED && STAGE < 4) ? ((STAGE == 2 || STAGE == 3) ? 2 : 1) : 0;
  const int nxi = p.nx - 2 * kBcSize, nyi = p.ny - 2 * kBcSize;
  const int segx = (nxi + kFrameSeg - 1) / kFrameSeg, segy = (nyi + kFrameSeg - 1) / kFrameSeg;
  const int tid = threadIdx.x, line = tid / kFrameSeg, lt = tid % kFrameSeg;
  int b = blockIdx.x;
  int i, j, ci0, rj0, w, h;
  bool active;
  const bool corners = b == 2 * segx + 2 * segy;
  if (b < 2 * segx) {  // S (b < segx) or N
    const bool north = b >= segx;
    const int i0 = kBcSize + (north ? b - segx : b) * kFrameSeg;
    i = i0 + lt;
    j = north ? p.ny - kBcSize + line : line;
    active = i <= p.nx - kBcSize - 1;
    ci0 = i0 - 2;
    rj0 = north ? p.ny - 3 : 0;
    w = kFrameSeg + 4;
    h = 3;
  } else if ((b -= 2 * segx) < 2 * segy) {  // W or E
    const bool east = b >= segy;
    const int j0 = kBcSize + (east ? b - segy : b) * kFrameSeg;
    j = j0 + lt;
    i = east ? p.nx - kBcSize + line : line;
    active = j <= p.ny - kBcSize - 1;
    ci0 = east ? p.nx - 3 : 0;
    rj0 = j0 - 2;
    w = 3;
    h = kFrameSeg + 4;
  } else {  // corners: thread t < 16, corner q = t / 4
    const int q = (tid >> 2) & 3, cx = q & 1, cy = q >> 1;
    i = (cx ? p.nx - kBcSize : 0) + (tid & 1);
    j = (cy ? p.ny - kBcSize : 0) + ((tid >> 1) & 1);
    active = tid < 16;
    ci0 = cx ? p.nx - 3 : 0;
    rj0 = cy ? p.ny - 3 : 0;
    w = 3;
    h = 3;
  }
  // stage the window(s) (cells outside the array are never read)
  if (corners) {
    for (int t = tid; t < 36; t += blockDim.x) {
      const int q = t / 9, cx = q & 1, cy = q >> 1;
      const int ci = (cx ? p.nx - 3 : 0) + t % 3, rj = (cy ? p.ny - 3 : 0) + (t % 9) / 3;
      frame::load_input(p, a, SRC, ci, rj, win + t * 4);
    }
  } else {
    for (int t = tid; t < w * h; t += blockDim.x) {
      const int ci = ci0 + t % w, rj = rj0 + t / w;
      if (ci >= 0 && ci < p.nx && rj >= 0 && rj < p.ny)
        frame::load_input(p, a, SRC, ci, rj, win + t * 4);
    }
  }
  __syncthreads();
  // the point functions of every window cell (cells outside the array hold
  // unused values)
  const int ncell = corners ? 36 : w * h;
  for (int t = tid; t < ncell; t += blockDim.x)
    for (int pk = 0; pk < frame::kNumPrim; ++pk)
      tab[pk * kFrameWin + t] = frame::point_function(p, win + t * 4, pk);
  __syncthreads();
  if (!active) return;
  const int wo = corners ? ((tid >> 2) & 3) * 9 : 0;  // first cell of this thread's window
  const frame::View f{&p, win + wo * 4, tab + wo, ci0, rj0, w, kFrameWin};
  double L[4];
  frame::timedev(f, i, j, L);
  const long long o = ((long long)j * p.pitch + i) * 4;
  if (STAGE == 0) {
    for (int k = 0; k < 4; ++k) a.O64[o + k] = L[k];
    return;
  }
  // the frame of a step's first stage runs beside its stage kernel, which
  // writes ctl->dt: it evaluates the same rule itself (later stages read it)
  const double dt = (STAGE == 1 || STAGE == 4)
                        ? (a.cfl_mode ? cfl_dt_from(a.ctl, a.parity) : a.dt_fixed)
                        : a.ctl->dt;
  double S[4];
  f.cell(i, j, S);
  if (STAGE >= 4) {
    constexpr int s = STAGE - 4;
    constexpr double LA = s == 0 ? 0.0 : (s == 1 ? -5.0 / 9.0 : -153.0 / 128.0);
    constexpr double LB = s == 0 ? 1.0 / 3.0 : (s == 1 ? 15.0 / 16.0 : 8.0 / 15.0);
    for (int k = 0; k < 4; ++k) {
      const double qin = s == 0 ? 0.0 : (MIXED ? (double)a.Q32[o + k] : a.Q[o + k]);
      const double q = (s == 0 ? 0.0 : LA * qin) + dt * L[k];
      a.O64[o + k] = S[k] + LB * q;
      if (s < 2) {
        if (MIXED) a.Q32[o + k] = (float)q;
        else a.Q[o + k] = q;
      }
    }
  } else if (!MIXED) {
    for (int k = 0; k < 4; ++k) {
      const double inc = S[k] + dt * L[k];
      double val;
      if (STAGE == 1) val = inc;
      else if (STAGE == 2) val = 0.75 * a.U[o + k] + (1.0 - 0.75) * inc;
      else val = (1.0 / 3.0) * a.U[o + k] + (1.0 - 1.0 / 3.0) * inc;
      a.O64[o + k] = val;
    }
  } else {
    for (int k = 0; k < 4; ++k) {
      const double dl = dt * L[k];
      if (STAGE == 1) a.O32[o + k] = (float)dl;
      else if (STAGE == 2) a.O32[o + k] = (float)(0.25 * ((double)a.S32[o + k] + dl));
      else a.O64[o + k] = S[k] + ((2.0 / 3.0) * dl - (1.0 / 3.0) * (double)a.S32[o + k]);
    }
  }
}

// ------------------------------------------------------------------------
// ghost materialisation, reductions
// ------------------------------------------------------------------------
__global__ void fill_ghosts_kernel(const DevParams p, double* F) {
  const int nrow = 2 * kBcSize * p.nx;
  const int ncol = 2 * kBcSize * (p.ny - 2 * kBcSize);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int c, r;
  if (t < nrow) {
    const int q = t / p.nx;
    c = t % p.nx;
    r = q < kBcSize ? q : p.ny - 2 * kBcSize + q;
  } else if (t < nrow + ncol) {
    const int u = t - nrow;
    const int q = u / (p.ny - 2 * kBcSize);
    r = kBcSize + u % (p.ny - 2 * kBcSize);
    c = q < kBcSize ? q : p.nx - 2 * kBcSize + q;
  } else {
    return;
  }
  bool fx, fy;
  const int cm = map_col(p, c, fx), rm = map_row(p, r, fy);
  if (cm == c && rm == r) return;  // stored (wall) ghost
  const long long src = ((long long)rm * p.pitch + cm) * 4, dst = ((long long)r * p.pitch + c) * 4;
  for (int k = 0; k < 4; ++k) {
    double v = F[src + k];
    if ((k == 1 && fx) || (k == 2 && fy)) v = -v;
    F[dst + k] = v;
  }
}

__global__ void clear_slot_kernel(StepCtl* ctl, int slot) { clear_slot(ctl, slot); }

// Reductions of the interior of F into slot `slot` (cleared before).  A
// workgroup of 256 lanes takes segments of 256 consecutive cells of a row
// (32-bit index arithmetic, uniform per segment) and commits once.
constexpr int kRedThreads = 256;
__global__ void __launch_bounds__(kRedThreads) reduce_kernel(const DevParams p, const double* F,
                                                             StepCtl* ctl, int slot) {
  RedAcc<true> red;
  red.init();
  const int ni = p.nx - 2 * kBcSize, nj = p.ny - 2 * kBcSize;
  const int segs = (ni + kRedThreads - 1) / kRedThreads;
  const int total = segs * nj;
  for (int sg = blockIdx.x; sg < total; sg += gridDim.x) {
    const int j = kBcSize + sg / segs;
    const int i = kBcSize + (sg % segs) * kRedThreads + (int)threadIdx.x;
    if (i < kBcSize + ni) {
      double v[4];
      ld_cell(F + (long long)j * p.pitch * 4, (unsigned)i, v);
      red.add(p, v);
    }
  }
  red.flush_block(ctl, slot);
}

// planar nodes(nx,ny,4) <-> interleaved cells, one thread per cell
__global__ void pack_kernel(const DevParams p, const double* __restrict__ planar,
                            double* __restrict__ F) {
  const long long n = (long long)p.nx * p.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % p.nx), j = (int)(t / p.nx);
  double v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = planar[k * n + t];
  st_cell(F + (long long)j * p.pitch * 4, (unsigned)i, v);
}

__global__ void unpack_kernel(const DevParams p, const double* __restrict__ F,
                              double* __restrict__ planar, int fill_ghosts, double ghost_value) {
  const long long n = (long long)p.nx * p.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % p.nx), j = (int)(t / p.nx);
  double v[4];
  ld_cell(F + (long long)j * p.pitch * 4, (unsigned)i, v);
  if (fill_ghosts &&
      (i < kBcSize || i >= p.nx - kBcSize || j < kBcSize || j >= p.ny - kBcSize)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ghost_value;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) planar[k * n + t] = v[k];
}

// ------------------------------------------------------------------------
// Cockburn and Gau operators (sd_operators/cg_operators/sd_operators_class.f)
// ------------------------------------------------------------------------
// FP64 time derivative and RK3-TVD stages with the fourth-order face stencils
// of Cockburn and Gau: f = (-p[i-2] + 7p[i-1] + 7p[i] - p[i+1])/12, dfdy over
// four columns of rows j+-1, d2fdy2 over the f of three rows, dfdx and d2fdx2
// as Mattsson's; g, dgdx, dgdy, d2gdx2, d2gdy2 the same across y.  An x-face
// reads a 4x3 block of cells and a y-face a 3x4 block, so this kernel does not
// use the register pipeline of the Mattsson stencils (whose faces read a cross
// of cells): a 16x15 tile of output cells stages the per-cell quantities of
// the tile and its 2-cell halo in LDS (state, velocities, pressure, classical
// and capillary temperatures, 1/(3 - rho)), computes each face of the tile
// once into LDS, then the divergence and the RK update.  The DIM fluxes follow
// dim2d_fluxes_module.f:103-1649 term by term (F = inv - visc/Re - cap/We).
namespace cg {

// 16 x 15 output cells: the tile's 17 x 15 x-faces and 16 x 16 y-faces each
// take one pass of the 256 threads.  A 32-wide tile (512 threads, two tiles
// per CU) measured the same as the 16-wide one.
constexpr int kTX = 16, kTY = 15;  // output cells per tile along x and y
constexpr int kThreads = kTX * (kTY + 1);   // one y-face per thread
static_assert(kTX == 16 && kTY == 15, "row_of and the face maps assume 16 x 15 tiles");
constexpr int kH = 2;              // halo (bc_size)
constexpr int kWX = kTX + 2 * kH;  // staged cells along x (20) and y (19)
constexpr int kWY = kTY + 2 * kH;
// Staged quantities besides the density.  Two planes carry sums that the
// fluxes only use through linear stencils: H = E + P, since the energy flux
// averages E u + P u; and the heat-flux potential
// Q_T = -T_cl/(Re Pr) + T_cap/(We Re Pr), whose difference across a face is
// the flux's two temperature-gradient terms together.  In FP64 the momenta
// are not staged either: they are rho u and rho v at the point where the
// fluxes use them, from the density and velocity planes, which keeps the FP64
// tile at 29.4 KB (five tiles per CU; the mixed tile, 19.3 KB with the
// momenta, is bound by registers).
enum { Q_H = 0, Q_U, Q_V, Q_P, Q_T, Q_W, Q_QX, Q_QY };
template <typename R>
constexpr bool kMomFromRho = std::is_same<R, double>::value;
template <typename R>
constexpr int kQ = kMomFromRho<R> ? 6 : 8;

// Face rows of the thread groups.  A group of 16 threads computes one row of
// 16 faces, and the LDS serves FP32 reads 32 lanes at a time: with the row
// stride of kWX = 20 words, the two rows of a half-wave are placed 4 rows
// apart (4 * 20 = 16 mod 32 banks), so that their 16-word segments fall on
// disjoint banks.  row_of(g) is a permutation of 0..15, and of 0..14 over
// g < 15.
__device__ __forceinline__ int row_of(int g) { return ((g >> 1) & 3) + 8 * (g >> 3) + 4 * (g & 1); }

// one face buffer, for the x-faces (kTX+1 per row, kTY rows) and then the
// y-faces (kTX per row, kTY+1 rows).  The density stays FP64 in both modes
// (its differences feed the capillary second differences); the other
// quantities and the faces are in the flux precision R: 32.5 KB per tile in
// FP64 (five tiles per CU), 19.3 KB with FP32 fluxes
constexpr int kFaces = (kTX + 1) * kTY > kTX * (kTY + 1) ? (kTX + 1) * kTY : kTX * (kTY + 1);
template <typename R>
struct Tile {
  double rho[kWX * kWY];
  R q[kQ<R>][kWX * kWY];
  R f[4][kFaces];
};

// staged value of quantity Q (or the density) at tile-local cell (x, y)
#define CG_AT(Q, x, y) (t.q[(Q)][(y) * kWX + (x)])
#define CG_RHO(x, y) (t.rho[(y) * kWX + (x)])

// staged quantity Q at (x, y); in FP64 the momenta from the density and the
// velocity planes
template <typename R>
__device__ __forceinline__ R qv(const Tile<R>& t, int Q, int x, int y) {
  if (kMomFromRho<R> && (Q == Q_QX || Q == Q_QY))
    return (R)CG_RHO(x, y) * CG_AT(Q == Q_QX ? Q_U : Q_V, x, y);
  return CG_AT(Q, x, y);
}

// the four-point face average of a product A*B (B = 1 when absent)
template <typename R>
__device__ __forceinline__ R fx4(const Tile<R>& t, int qa, int qb, int x, int y) {
  auto v = [&](int xx) { return qb < 0 ? qv(t, qa, xx, y) : qv(t, qa, xx, y) * qv(t, qb, xx, y); };
  return (R)(1.0 / 12.0) * (-v(x - 2) + (R)7 * v(x - 1) + (R)7 * v(x) - v(x + 1));
}
template <typename R>
__device__ __forceinline__ R gy4(const Tile<R>& t, int qa, int qb, int x, int y) {
  auto v = [&](int yy) { return qb < 0 ? qv(t, qa, x, yy) : qv(t, qa, x, yy) * qv(t, qb, x, yy); };
  return (R)(1.0 / 12.0) * (-v(y - 2) + (R)7 * v(y - 1) + (R)7 * v(y) - v(y + 1));
}

// x-face (i-1/2, j) at tile-local (x, y) = cell i's position; the density's
// face value and derivatives in FP64, rounded to R once
template <typename R>
__device__ void x_face(const DevParams& p, const Tile<R>& t, int x, int y, R F[4]) {
  const double idx = p.idx, idy = p.idy;
  const R ridx = PC(R, idx), ridy = PC(R, idy);
  auto dfdx = [&](int Q) { return ridx * (-CG_AT(Q, x - 1, y) + CG_AT(Q, x, y)); };
  auto dfdy = [&](int Q) {
    return (R)(1.0 / 24.0) * ridy *
           (+CG_AT(Q, x - 2, y - 1) - (R)7 * CG_AT(Q, x - 1, y - 1) - (R)7 * CG_AT(Q, x, y - 1) +
            CG_AT(Q, x + 1, y - 1) - CG_AT(Q, x - 2, y + 1) + (R)7 * CG_AT(Q, x - 1, y + 1) +
            (R)7 * CG_AT(Q, x, y + 1) - CG_AT(Q, x + 1, y + 1));
  };
  auto row4 = [&](int yy) {
    return -CG_RHO(x - 2, yy) + 7.0 * CG_RHO(x - 1, yy) + 7.0 * CG_RHO(x, yy) - CG_RHO(x + 1, yy);
  };
  const double d2x = 0.5 * idx * idx *
                     (CG_RHO(x - 2, y) - CG_RHO(x - 1, y) - CG_RHO(x, y) + CG_RHO(x + 1, y));
  const double d2y = 1.0 / 12.0 * idy * idy * (row4(y - 1) - 2.0 * row4(y) + row4(y + 1));
  const double rxd = idx * (-CG_RHO(x - 1, y) + CG_RHO(x, y));
  const double ryd = 1.0 / 24.0 * idy *
                     (+CG_RHO(x - 2, y - 1) - 7.0 * CG_RHO(x - 1, y - 1) - 7.0 * CG_RHO(x, y - 1) +
                      CG_RHO(x + 1, y - 1) - CG_RHO(x - 2, y + 1) + 7.0 * CG_RHO(x - 1, y + 1) +
                      7.0 * CG_RHO(x, y + 1) - CG_RHO(x + 1, y + 1));
  const double frd = 1.0 / 12.0 * (-CG_RHO(x - 2, y) + 7.0 * CG_RHO(x - 1, y) +
                                   7.0 * CG_RHO(x, y) - CG_RHO(x + 1, y));
  const R rx = (R)rxd, ry = (R)ryd, lap = (R)(d2x + d2y), fr = (R)frd;
  const R dux = dfdx(Q_U), duy = dfdy(Q_U), dvx = dfdx(Q_V), dvy = dfdy(Q_V);
  const R fu = fx4(t, Q_U, -1, x, y), fv = fx4(t, Q_V, -1, x, y);
  const R fqx = fx4(t, Q_QX, -1, x, y);
  const R visc = PC(R, two_vr) * dux + PC(R, vr) * dvy;
  const R shear = duy + dvx;
  F[0] = fqx;
  F[1] = fx4(t, Q_QX, Q_U, x, y) + fx4(t, Q_P, -1, x, y) - PC(R, inv_re) * visc -
         PC(R, inv_we) * (PC(R, c15cv) * fx4(t, Q_W, -1, x, y) * (rx * rx + ry * ry) + fr * lap +
                          (R)0.5 * (-rx * rx + ry * ry));
  F[2] = fx4(t, Q_QY, Q_U, x, y) - PC(R, inv_re) * shear - PC(R, inv_we) * (-rx * ry);
  const R visc_e = fu * visc + fv * shear;
  const R cap_e = PC(R, c15cv) * fx4(t, Q_W, Q_U, x, y) * (rx * rx + ry * ry) + fqx * lap +
                  (R)0.5 * fu * (-rx * rx + ry * ry) - rx * ry * fv - fr * rx * (dux + dvy);
  F[3] = fx4(t, Q_H, Q_U, x, y) - PC(R, inv_re) * visc_e - PC(R, inv_we) * cap_e + dfdx(Q_T);
}

// y-face (i, j-1/2) at tile-local (x, y) = cell j's position
template <typename R>
__device__ void y_face(const DevParams& p, const Tile<R>& t, int x, int y, R G[4]) {
  const double idx = p.idx, idy = p.idy;
  const R ridx = PC(R, idx), ridy = PC(R, idy);
  auto dgdy = [&](int Q) { return ridy * (-CG_AT(Q, x, y - 1) + CG_AT(Q, x, y)); };
  auto dgdx = [&](int Q) {
    return (R)(1.0 / 24.0) * ridx *
           (+CG_AT(Q, x - 1, y - 2) - (R)7 * CG_AT(Q, x - 1, y - 1) - (R)7 * CG_AT(Q, x - 1, y) +
            CG_AT(Q, x - 1, y + 1) - CG_AT(Q, x + 1, y - 2) + (R)7 * CG_AT(Q, x + 1, y - 1) +
            (R)7 * CG_AT(Q, x + 1, y) - CG_AT(Q, x + 1, y + 1));
  };
  auto col4 = [&](int xx) {
    return -CG_RHO(xx, y - 2) + 7.0 * CG_RHO(xx, y - 1) + 7.0 * CG_RHO(xx, y) - CG_RHO(xx, y + 1);
  };
  const double d2x = 1.0 / 12.0 * idx * idx * (col4(x - 1) - 2.0 * col4(x) + col4(x + 1));
  const double d2y = 0.5 * idy * idy *
                     (CG_RHO(x, y - 2) - CG_RHO(x, y - 1) - CG_RHO(x, y) + CG_RHO(x, y + 1));
  const double rxd = 1.0 / 24.0 * idx *
                     (+CG_RHO(x - 1, y - 2) - 7.0 * CG_RHO(x - 1, y - 1) - 7.0 * CG_RHO(x - 1, y) +
                      CG_RHO(x - 1, y + 1) - CG_RHO(x + 1, y - 2) + 7.0 * CG_RHO(x + 1, y - 1) +
                      7.0 * CG_RHO(x + 1, y) - CG_RHO(x + 1, y + 1));
  const double ryd = idy * (-CG_RHO(x, y - 1) + CG_RHO(x, y));
  const double grd = 1.0 / 12.0 * (-CG_RHO(x, y - 2) + 7.0 * CG_RHO(x, y - 1) +
                                   7.0 * CG_RHO(x, y) - CG_RHO(x, y + 1));
  const R rx = (R)rxd, ry = (R)ryd, lap = (R)(d2x + d2y), gr = (R)grd;
  const R dux = dgdx(Q_U), duy = dgdy(Q_U), dvx = dgdx(Q_V), dvy = dgdy(Q_V);
  const R gu = gy4(t, Q_U, -1, x, y), gv = gy4(t, Q_V, -1, x, y);
  const R gqy = gy4(t, Q_QY, -1, x, y);
  const R shear = duy + dvx;
  const R visc = PC(R, vr) * dux + PC(R, two_vr) * dvy;
  G[0] = gqy;
  G[1] = gy4(t, Q_QX, Q_V, x, y) - PC(R, inv_re) * shear - PC(R, inv_we) * (-rx * ry);
  G[2] = gy4(t, Q_QY, Q_V, x, y) + gy4(t, Q_P, -1, x, y) - PC(R, inv_re) * visc -
         PC(R, inv_we) * (PC(R, c15cv) * gy4(t, Q_W, -1, x, y) * (rx * rx + ry * ry) + gr * lap +
                          (R)0.5 * (rx * rx - ry * ry));
  const R visc_e = gu * shear + gv * visc;
  const R cap_e = PC(R, c15cv) * gy4(t, Q_W, Q_V, x, y) * (rx * rx + ry * ry) + gqy * lap +
                  (R)0.5 * gv * (rx * rx - ry * ry) - rx * ry * gu - gr * ry * (dux + dvy);
  G[3] = gy4(t, Q_H, Q_V, x, y) - PC(R, inv_re) * visc_e - PC(R, inv_we) * cap_e + dgdy(Q_T);
}

#undef CG_AT
#undef CG_RHO

}  // namespace cg

// FP64: five tiles per CU fit the LDS, so the registers are held to five
// waves per SIMD (stage 3 would otherwise take 103 VGPRs and four)
#ifndef DIM2D_CG_WAVES_FP64
#define DIM2D_CG_WAVES_FP64 5
#endif
// STAGE 0: time derivative into O64; 1, 2, 3: RK3-TVD stages, as stage_kernel's:
// FP64 (MIXED = 0) from S64 (stage input) and U (u^n), or mixed precision
// (FP32 fluxes and stage increments: stage 1 U -> O32, stage 2 U + S32 ->
// O32, stage 3 U + S32 -> O64), stage 3 with the CFL reduction.  Grid: one
// workgroup of kThreads threads per tile.
template <int MIXED, int STAGE>
__global__ void __launch_bounds__(cg::kThreads)
__attribute__((amdgpu_waves_per_eu(MIXED ? 1 : DIM2D_CG_WAVES_FP64, 8)))
cg_stage_kernel(const DevParams p, const StageArgs a) {
  using namespace cg;
  using R = typename std::conditional<MIXED, float, double>::type;
  __shared__ Tile<R> t;
  constexpr bool FIRST = STAGE == 1, LAST = STAGE == 3;
  // the stage input: S64, U alone (mixed stage 0/1) or U + S32 (mixed 2/3)
  constexpr int SRC = MIXED ? ((STAGE == 2 || STAGE == 3) ? 2 : 1) : 0;
  double dt = a.dt_fixed;
  if (FIRST) {
    if (a.cfl_mode) dt = cfl_dt_from(a.ctl, a.parity);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      a.ctl->dt = dt;
      clear_slot(a.ctl, a.parity ^ 1);
    }
  } else if (STAGE != 0) {
    dt = a.ctl->dt;
  }
  const int tid = threadIdx.x;
  const int i0 = kBcSize + blockIdx.x * kTX, j0 = kBcSize + blockIdx.y * kTY;  // first output cell
  // stage 3: the reduction words, read now so that the read's latency hides
  // behind the tile's work (the slot was cleared by stage 1)
  unsigned long long pk[3] = {0ull, 0ull, 0ull};
  if (LAST && tid == 0) {
    pk[0] = peek(&a.ctl->amax[a.parity ^ 1]);
    pk[1] = peek(&a.ctl->rmin[a.parity ^ 1]);
    pk[2] = peek(&a.ctl->rmax[a.parity ^ 1]);
  }
  // the output cell of this thread; in stages 2 and 3 its u^n (and, mixed,
  // its stage increment) are loaded now, so that the loads' latency hides
  // behind the tile's work instead of delaying the update at its end
  const int ti = tid % kTX, tj = tid / kTX;
  const int i = i0 + ti, j = j0 + tj;
  const bool out_cell = tj < kTY && i <= p.nx - kBcSize - 1 && j <= p.ny - kBcSize - 1;
  const long long o = ((long long)j * p.pitch + i) * 4;
  constexpr bool NEED_U = (STAGE == 2 || STAGE == 3) && !(MIXED && STAGE == 2);
  constexpr bool NEED_S32 = MIXED && (STAGE == 2 || STAGE == 3);
  // FP64 stages: the stage input of the output cell (its momenta and energy
  // are not staged on their own)
  constexpr bool NEED_S = !MIXED && STAGE != 0;
  double un[4] = {0.0, 0.0, 0.0, 0.0};
  float s32v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  double s_in[4] = {0.0, 0.0, 0.0, 0.0};
  if (out_cell) {
    if (NEED_S) ld_cell(a.S64 + (long long)j * p.pitch * 4, (unsigned)i, s_in);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (NEED_U) un[k] = a.U[o + k];
      if (NEED_S32) s32v[k] = a.S32[o + k];
    }
  }
  // 1. state of the tile and its halo (virtual ghosts through the index
  // maps), the per-cell quantities in FP64, stored in the flux precision
  for (int c = tid; c < kWX * kWY; c += kThreads) {
    const int x = c % kWX, y = c / kWX;
    const int i = min(max(i0 - kH + x, 0), p.nx - 1), j = min(max(j0 - kH + y, 0), p.ny - 1);
    bool fx, fy;
    const int cm = map_col(p, i, fx), rm = map_row(p, j, fy);
    double s[4];
    float inc[4];
    load_state<SRC>(p, a.U, a.S64, a.S32, rm, (unsigned)cm, fx, fy, s, inc);
    const double r = s[0], qx = s[1], qy = s[2], E = s[3];
    // reciprocals as the Mattsson stage kernels take them (hardware estimate
    // and two Newton steps, ~1 ulp) instead of five divisions
    const double ir = fast_rcp(r), w = fast_rcp(3.0 - r);
    const double u = qx * ir, v = qy * ir;
    const double Tn = E - 0.5 * r * (u * u + v * v) + 3.0 * r * r;
    t.rho[c] = r;
    if (!kMomFromRho<R>) {
      t.q[Q_QX][c] = (R)qx;
      t.q[Q_QY][c] = (R)qy;
    }
    const double P = p.c3cv * w * Tn - 3.0 * r * r;  // classical_pressure_local
    t.q[Q_H][c] = (R)(E + P);
    t.q[Q_U][c] = (R)u;
    t.q[Q_V][c] = (R)v;
    t.q[Q_P][c] = (R)P;
    t.q[Q_T][c] = (R)(Tn * ir);  // classical_temperature_eff, completed to Q_T in step 2
    t.q[Q_W][c] = (R)w;                                // capillarity_pressure
  }
  __syncthreads();
  // 2. capillarity_temperature_eff (central gradients) on the tile and a 1-cell
  // halo, combined with the classical temperature into the heat-flux potential
  const double kt_cl = -p.inv_re * p.inv_pr, kt_cap = p.inv_we * p.inv_repr;
  // (rows of 16 cells on consecutive threads, then the last two columns)
  for (int c = tid; c < (kWX - 2) * (kWY - 2); c += kThreads) {
    constexpr int kMain = kTX * (kWY - 2);
    const int e = c - kMain;
    const int x = c < kMain ? 1 + c % kTX : kTX + 1 + (e & 1);
    const int y = c < kMain ? 1 + c / kTX : 1 + (e >> 1);
    const int k = y * kWX + x;
    const double gx = 0.5 * p.idx * (-t.rho[k - 1] + t.rho[k + 1]);
    const double gy = 0.5 * p.idy * (-t.rho[k - kWX] + t.rho[k + kWX]);
    const double tcap = 0.5 * fast_rcp(t.rho[k]) * (gx * gx + gy * gy);
    t.q[Q_T][k] = (R)(kt_cl * (double)t.q[Q_T][k] + kt_cap * tcap);
  }
  __syncthreads();
  // 3. the faces of the tile, one of each kind per thread, and their
  // divergence (td_operators_class.f:181-194): x-faces i-1/2 for
  // i = i0..i0+kTX, then y-faces j-1/2 for j = j0..j0+kTY, through one buffer
  if (tid < (kTX + 1) * kTY) {
    // rows of kTX faces on consecutive threads, so that each group of 16
    // lanes reads one contiguous row segment of the staged planes (no LDS
    // bank conflicts); the last face column on the kTY threads after them
    R F[4];
    const bool last_col = tid >= kTX * kTY;
    const int fi = last_col ? kTX : tid % kTX, fj = last_col ? tid - kTX * kTY : row_of(tid / kTX);
    cg::x_face<R>(p, t, kH + fi, kH + fj, F);
#pragma unroll
    for (int k = 0; k < 4; ++k) t.f[k][fj * (kTX + 1) + fi] = F[k];
  }
  __syncthreads();
  R L[4];
  const R hidx = PC(R, idx), hidy = PC(R, idy);
  if (out_cell) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      L[k] = hidx * (t.f[k][tj * (kTX + 1) + ti] - t.f[k][tj * (kTX + 1) + ti + 1]);
  }
  __syncthreads();
  {
    R G[4];
    const int fi = tid % kTX, fj = row_of(tid / kTX);
    cg::y_face<R>(p, t, kH + fi, kH + fj, G);
#pragma unroll
    for (int k = 0; k < 4; ++k) t.f[k][fj * kTX + fi] = G[k];
  }
  __syncthreads();
  // 4. body force and RK update
  RedAcc<false> red;
  red.init();
  if (out_cell) {
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] += hidy * (t.f[k][tj * kTX + ti] - t.f[k][(tj + 1) * kTX + ti]);
    const int c = (kH + tj) * kWX + kH + ti;
    if (p.gravity_ac) {
      // compute_body_forces (pmodel_eq_class.f:1840-1918)
      L[2] -= (R)(t.rho[c] * p.gravity);
      L[3] -= (R)((double)qv(t, Q_QY, kH + ti, kH + tj) * p.gravity);
    }
    double out[4];
    if (STAGE == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) out[k] = (double)L[k];
#pragma unroll
      for (int k = 0; k < 4; ++k) a.O64[o + k] = out[k];
    } else if (!MIXED) {
      // the stage input at an output cell is interior: no map or sign applies
      const double* S = s_in;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double inc = S[k] + dt * (double)L[k];
        if (STAGE == 1) {
          out[k] = inc;
        } else {
          if (STAGE == 2) out[k] = 0.75 * un[k] + (1.0 - 0.75) * inc;
          else out[k] = (1.0 / 3.0) * un[k] + (1.0 - 1.0 / 3.0) * inc;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) a.O64[o + k] = out[k];
    } else if (STAGE == 1 || STAGE == 2) {
      // FP32 increments against u^n, as the Mattsson mixed stages
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dl = dt * (double)L[k];
        a.O32[o + k] = STAGE == 1 ? (float)dl : (float)(0.25 * ((double)s32v[k] + dl));
      }
    } else {
      // u^n + 2/3 (S32 + dt L) = (u^n + S32) + (2/3 dt L - 1/3 S32)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double s32 = (double)s32v[k];
        out[k] = (un[k] + s32) + ((2.0 / 3.0) * (dt * (double)L[k]) - (1.0 / 3.0) * s32);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) a.O64[o + k] = out[k];
    }
    if (LAST) red.template add<MIXED != 0>(p, out);
  }
  if (LAST) {
    // the four waves combine in LDS; one thread commits the workgroup's values
    constexpr int kWaves = kThreads / 64;
    __shared__ double rl[3][kWaves];
    __shared__ unsigned rnf[kWaves];
    const double am = wave_max(red.amax), rn = wave_min(red.rmin), rx = wave_max(red.rmax);
    const unsigned nf = wave_sum_u(red.nonfinite);
    if ((tid & 63) == 0) {
      rl[0][tid >> 6] = am;
      rl[1][tid >> 6] = rn;
      rl[2][tid >> 6] = rx;
      rnf[tid >> 6] = nf;
    }
    __syncthreads();
    if (tid == 0) {
      double A = rl[0][0], N = rl[1][0], X = rl[2][0];
      unsigned F = rnf[0];
      for (int w = 1; w < kWaves; ++w) {
        A = fmax(A, rl[0][w]);
        N = fmin(N, rl[1][w]);
        X = fmax(X, rl[2][w]);
        F += rnf[w];
      }
      RedAcc<false>::commit(a.ctl, a.parity ^ 1, A, N, X, F, pk);
      if (blockIdx.x == 0 && blockIdx.y == 0) a.ctl->t += dt;
    }
  }
}

// ------------------------------------------------------------------------
// steady-state check (field_abstract_class.f:629-667)
// ------------------------------------------------------------------------
// max |F| per variable over the interior cells, and with `frame` the ghost
// frame too (open edges, where it carries a time derivative); a non-finite
// value counts as +inf.  out[0..3]: bits of non-negative doubles, combined
// with atomicMax.
__global__ void __launch_bounds__(kRedThreads) absmax_kernel(const DevParams p, int frame,
                                                             const double* __restrict__ F,
                                                             unsigned long long* __restrict__ out) {
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  const int i0 = frame ? 0 : kBcSize, j0 = frame ? 0 : kBcSize;
  const int ni = p.nx - 2 * i0, nj = p.ny - 2 * j0;
  const int segs = (ni + kRedThreads - 1) / kRedThreads;
  const int total = segs * nj;
  for (int sg = blockIdx.x; sg < total; sg += gridDim.x) {
    const int j = j0 + sg / segs;
    const int i = i0 + (sg % segs) * kRedThreads + (int)threadIdx.x;
    if (i >= i0 + ni) continue;
    const double* c = F + ((long long)j * p.pitch + i) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double a = fabs(c[k]);
      if (!(a <= m[k])) m[k] = isnan(a) ? INFINITY : a;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double w = wave_max(m[k]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(w);
    if ((threadIdx.x & 63) == 0 && w > 0.0 && b > peek(&out[k])) atomicMax(&out[k], b);
  }
}

// ------------------------------------------------------------------------
// domain extension (field_extended_class.f:427-452, bf_interface_icr_class.f:251-386)
// ------------------------------------------------------------------------
// are_openbc_undermined (pmodel_eq_class.f:1991-2070): the sound speed of the
// point is imaginary, P + rho^2 (2 rho - 3) < 0; with md_thr >= 0 (the
// reference's bf_openbc_md_threshold_ac with bf_openbc_md_threshold = md_thr,
// parameters_input.f:833-846) also a density inside the two-phase band
// [rho_vap + eps (rho_mid - rho_vap), rho_liq - eps (rho_liq - rho_mid)] of the
// saturation densities at the point's temperature 3/(8 cv) temperature_eff
// (dim2d_prim_module.f:652-691, central density gradients; fits of
// dim2d_state_eq_module.f:145-215).  Counted over the `dist` interior rows or
// columns nearest each edge (the band up to the increase detectors, which sit
// obc_dct_distance points inside the edge, parameters_bf_layer.f:138-142);
// counts[0..3] = W, E, S, N.  Corner points count for both of their edges.
__device__ bool openbc_undermined_at(const DevParams& p, const double* __restrict__ U, int i,
                                     int j, double md_thr) {
  auto cell = [&](int ii, int jj) { return U + ((long long)jj * p.pitch + ii) * 4; };
  const double* c = cell(i, j);
  const double v[4] = {c[0], c[1], c[2], c[3]};
  const double P = frame::pressure_of(p, v);
  if (P + v[0] * v[0] * (-3.0 + 2.0 * v[0]) < 0.0) return true;
  if (md_thr < 0.0) return false;
  const double gx = (cell(i + 1, j)[0] - cell(i - 1, j)[0]) / (2.0 * p.dx);
  const double gy = (cell(i, j + 1)[0] - cell(i, j - 1)[0]) / (2.0 * p.dy);
  const double r = v[0];
  const double Teff = 1.0 / r *
                      (v[3] - 0.5 * r * ((v[1] / r) * (v[1] / r) + (v[2] / r) * (v[2] / r)) -
                       0.5 / p.we * (gx * gx + gy * gy) + 3.0 * r * r);
  const double T = 3.0 / (8.0 * p.cv) * Teff;
  const double md_liq = 1.0 + 2.06 * sqrt(1.0 - T);
  const double md_vap = 1.0 - 1.91 * sqrt(1.0 - T);
  const double md_mid = 0.5 * (md_vap + md_liq);
  const double vap_thr = md_vap + md_thr * (md_mid - md_vap);
  const double liq_thr = md_liq - md_thr * (md_liq - md_mid);
  return r >= vap_thr && r <= liq_thr;
}

__global__ void undermined_kernel(const DevParams p, int dist, double md_thr,
                                  const double* __restrict__ U, int* __restrict__ counts) {
  const int nxi = p.nx - 2 * kBcSize, nyi = p.ny - 2 * kBcSize;
  const int dx = min(dist, nxi), dy = min(dist, nyi);
  const long long nx_band = (long long)dx * nyi;  // one x-band (W or E)
  const long long n = 2 * nx_band + 2LL * dy * nxi;
  int local[4] = {0, 0, 0, 0};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int edge, i, j;
    if (t < 2 * nx_band) {
      edge = t < nx_band ? 0 : 1;
      const long long q = t - edge * nx_band;
      const int off = (int)(q % dx);
      j = kBcSize + (int)(q / dx);
      i = edge == 0 ? kBcSize + off : p.nx - kBcSize - 1 - off;
    } else {
      const long long q0 = t - 2 * nx_band;
      const long long y_band = (long long)dy * nxi;
      edge = q0 < y_band ? 2 : 3;
      const long long q = q0 - (edge - 2) * y_band;
      const int off = (int)(q / nxi);
      i = kBcSize + (int)(q % nxi);
      j = edge == 2 ? kBcSize + off : p.ny - kBcSize - 1 - off;
    }
    if (openbc_undermined_at(p, U, i, j, md_thr)) local[edge] += 1;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const unsigned c = wave_sum_u((unsigned)local[e]);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&counts[e], (int)c);
  }
}

// undermined points in up to four rectangles [i0, i1) x [j0, j1) (0-based,
// every point with its 3x3 neighbourhood in the array), counted per rectangle
__global__ void undermined_rects_kernel(const DevParams p, const Rects4 r, double md_thr,
                                        const double* __restrict__ U, int* __restrict__ counts) {
  long long area[4], total = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    area[e] = (long long)max(r.i1[e] - r.i0[e], 0) * max(r.j1[e] - r.j0[e], 0);
    total += area[e];
  }
  int local[4] = {0, 0, 0, 0};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int e = 0;
    long long q = t;
    while (q >= area[e]) q -= area[e++];
    const int w = r.i1[e] - r.i0[e];
    const int i = r.i0[e] + (int)(q % w), j = r.j0[e] + (int)(q / w);
    if (openbc_undermined_at(p, U, i, j, md_thr)) local[e] += 1;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const unsigned c = wave_sum_u((unsigned)local[e]);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&counts[e], (int)c);
  }
}

// copy of the state into a grid grown by add_w columns on the west and add_s
// rows on the south (and any number on the east and north): every point of
// the new grid takes the value of the nearest point of the old grid (the old
// points keep theirs, new points repeat the old edge point along the normal,
// new corner points the old corner point).  Negative counts crop: the new
// grid is a sub-rectangle of the old one.
__global__ void extend_kernel(const DevParams np, const DevParams op, int add_w, int add_s,
                              const double* __restrict__ oldU, double* __restrict__ newU) {
  const long long n = (long long)np.nx * np.ny;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % np.nx), j = (int)(t / np.nx);
    const int io = min(max(i - add_w, 0), op.nx - 1);
    const int jo = min(max(j - add_s, 0), op.ny - 1);
    const double2* src = reinterpret_cast<const double2*>(oldU + ((long long)jo * op.pitch + io) * 4);
    double2* dst = reinterpret_cast<double2*>(newU + ((long long)j * np.pitch + i) * 4);
    dst[0] = src[0];
    dst[1] = src[1];
  }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
template <int MIXED, int STAGE>
static hipError_t launch_one(const DevParams& p, const StageArgs& a, int block_x, int rows,
                             hipStream_t s) {
  const int ni = p.nx - 2 * kBcSize, nj = p.ny - 2 * kBcSize;
  const int per_wg = block_x == 2 * kLanes ? 2 : 1;  // strips per workgroup
  const int strips = (ni + kOutCols - 1) / kOutCols;
  dim3 grid((strips + per_wg - 1) / per_wg, (nj + rows - 1) / rows);
  hipLaunchKernelGGL((stage_kernel<MIXED, STAGE>), grid, dim3(per_wg * kLanes), 0, s, p, a, rows);
  return hipGetLastError();
}

hipError_t launch_stage(const DevParams& p, int mixed, int stage, const StageArgs& a, int block_x,
                        int rows, hipStream_t s) {
  if (!mixed) {
    switch (stage) {
      case 0: return launch_one<0, 0>(p, a, block_x, rows, s);
      case 1: return launch_one<0, 1>(p, a, block_x, rows, s);
      case 2: return launch_one<0, 2>(p, a, block_x, rows, s);
      case 3: return launch_one<0, 3>(p, a, block_x, rows, s);
      case 4: return launch_one<0, 4>(p, a, block_x, rows, s);
      case 5: return launch_one<0, 5>(p, a, block_x, rows, s);
      default: return launch_one<0, 6>(p, a, block_x, rows, s);
    }
  }
  switch (stage) {
    case 0: return launch_one<1, 0>(p, a, block_x, rows, s);
    case 1: return launch_one<1, 1>(p, a, block_x, rows, s);
    case 2: return launch_one<1, 2>(p, a, block_x, rows, s);
    case 3: return launch_one<1, 3>(p, a, block_x, rows, s);
    case 4: return launch_one<1, 4>(p, a, block_x, rows, s);
    case 5: return launch_one<1, 5>(p, a, block_x, rows, s);
    default: return launch_one<1, 6>(p, a, block_x, rows, s);
  }
}

bool pair_supported(const DevParams& p) {
  if (p.bc_w == kBcWall || p.bc_e == kBcWall) return false;
  // the producer's maps reach 4 rows/columns past an edge: the images must be
  // interior cells
  return p.nx >= 2 * kBcSize + 8 && p.ny >= 2 * kBcSize + 8;
}

hipError_t launch_stage_pair(const DevParams& p, int mixed, int lsrk, const StageArgs& a, int rows,
                             hipStream_t s) {
  const int ni = p.nx - 2 * kBcSize, nj = p.ny - 2 * kBcSize;
  dim3 grid((ni + kPairOut - 1) / kPairOut, (nj + rows - 1) / rows);
  const bool open = p.bc_n == kBcOpen;
#define DIM2D_PK(M, O, L) \
  hipLaunchKernelGGL((stage_pair_kernel<M, O, L>), grid, dim3(2 * kLanes), 0, s, p, a, rows)
  if (lsrk) {
    if (open) return hipErrorInvalidValue;
    if (mixed) DIM2D_PK(1, false, true);
    else DIM2D_PK(0, false, true);
  } else if (mixed) {
    if (open) DIM2D_PK(1, true, false);
    else DIM2D_PK(1, false, false);
  } else {
    if (open) DIM2D_PK(0, true, false);
    else DIM2D_PK(0, false, false);
  }
#undef DIM2D_PK
  return hipGetLastError();
}

hipError_t launch_stage_pair_batch(const DevParams* d_params, const WorkItem* d_items,
                                   int nitems, int mixed, const StageArgs& a, int workgroups,
                                   hipStream_t s) {
  if (mixed)
    hipLaunchKernelGGL(stage_pair_batch_kernel<1>, dim3(workgroups), dim3(2 * kLanes), 0, s,
                       d_params, d_items, nitems, a, a.dts);
  else
    hipLaunchKernelGGL(stage_pair_batch_kernel<0>, dim3(workgroups), dim3(2 * kLanes), 0, s,
                       d_params, d_items, nitems, a, a.dts);
  return hipGetLastError();
}

hipError_t launch_stage_mix_batch(const DevParams* d_params, const WorkItem* d_items, int nitems,
                                  int mixed, const MixArgs& m, const double* dts, int workgroups,
                                  hipStream_t s) {
  if (mixed)
    hipLaunchKernelGGL(stage_mix_batch_kernel<1>, dim3(workgroups), dim3(2 * kLanes), 0, s,
                       d_params, d_items, nitems, m, dts);
  else
    hipLaunchKernelGGL(stage_mix_batch_kernel<0>, dim3(workgroups), dim3(2 * kLanes), 0, s,
                       d_params, d_items, nitems, m, dts);
  return hipGetLastError();
}

hipError_t launch_stage_batch(const DevParams* d_params, const WorkItem* d_items, int nitems,
                              int mixed, int stage, const StageArgs& a, int workgroups,
                              hipStream_t s) {
  const int wg = workgroups;
#define DIM2D_BL(M, S)                                                                     \
  hipLaunchKernelGGL((stage_batch_kernel<M, S>), dim3(wg), dim3(kLanes), 0, s, d_params,  \
                     d_items, nitems, a, a.dts)
  if (!mixed) {
    if (stage == 1) DIM2D_BL(0, 1);
    else if (stage == 2) DIM2D_BL(0, 2);
    else DIM2D_BL(0, 3);
  } else {
    if (stage == 1) DIM2D_BL(1, 1);
    else if (stage == 2) DIM2D_BL(1, 2);
    else DIM2D_BL(1, 3);
  }
#undef DIM2D_BL
  return hipGetLastError();
}

hipError_t launch_wall(const DevParams& p, int mixed_in, const double* U, const double* S64,
                       const float* S32, double* O64, float* O32, const double* tmp,
                       hipStream_t s) {
  const int n = ((p.bc_s == kBcWall) + (p.bc_n == kBcWall)) * (p.nx - 2 * kBcSize) +
                ((p.bc_w == kBcWall) + (p.bc_e == kBcWall)) * (p.ny - 2 * kBcSize);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(wall_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, mixed_in, U, S64, S32,
                     O64, O32, tmp);
  return hipGetLastError();
}

hipError_t launch_wall_batch(const DevParams* h_params, const DevParams* d_params, int nsim,
                             int mixed_in, const double* U, const double* S64, const float* S32,
                             double* O64, float* O32, const double* tmp, hipStream_t s) {
  int nmax = 0;
  for (int i = 0; i < nsim; ++i) {
    const DevParams& p = h_params[i];
    const int n = ((p.bc_s == kBcWall) + (p.bc_n == kBcWall)) * (p.nx - 2 * kBcSize) +
                  ((p.bc_w == kBcWall) + (p.bc_e == kBcWall)) * (p.ny - 2 * kBcSize);
    nmax = n > nmax ? n : nmax;
  }
  if (nmax == 0) return hipSuccess;
  hipLaunchKernelGGL(wall_batch_kernel, dim3((nmax + 63) / 64, nsim), dim3(64), 0, s, d_params,
                     nsim, mixed_in, U, S64, S32, O64, O32, tmp);
  return hipGetLastError();
}

hipError_t launch_fill_ghosts(const DevParams& p, double* F, hipStream_t s) {
  const int n = 2 * kBcSize * p.nx + 2 * kBcSize * (p.ny - 2 * kBcSize);
  hipLaunchKernelGGL(fill_ghosts_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, F);
  return hipGetLastError();
}

hipError_t launch_frame(const DevParams& p, int mixed, int stage, const StageArgs& a,
                        hipStream_t s) {
  const int segx = (p.nx - 2 * kBcSize + kFrameSeg - 1) / kFrameSeg;
  const int segy = (p.ny - 2 * kBcSize + kFrameSeg - 1) / kFrameSeg;
  const dim3 g(2 * segx + 2 * segy + 1), b(2 * kFrameSeg);
#define DIM2D_FR(M, S) hipLaunchKernelGGL((frame_kernel<M, S>), g, b, 0, s, p, a)
  if (!mixed) {
    switch (stage) {
      case 0: DIM2D_FR(0, 0); break;
      case 1: DIM2D_FR(0, 1); break;
      case 2: DIM2D_FR(0, 2); break;
      case 3: DIM2D_FR(0, 3); break;
      case 4: DIM2D_FR(0, 4); break;
      case 5: DIM2D_FR(0, 5); break;
      default: DIM2D_FR(0, 6); break;
    }
  } else {
    switch (stage) {
      case 0: DIM2D_FR(1, 0); break;
      case 1: DIM2D_FR(1, 1); break;
      case 2: DIM2D_FR(1, 2); break;
      case 3: DIM2D_FR(1, 3); break;
      case 4: DIM2D_FR(1, 4); break;
      case 5: DIM2D_FR(1, 5); break;
      default: DIM2D_FR(1, 6); break;
    }
  }
#undef DIM2D_FR
  return hipGetLastError();
}

hipError_t launch_pack(const DevParams& p, const double* planar, double* F, hipStream_t s) {
  const long long n = (long long)p.nx * p.ny;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, planar, F);
  return hipGetLastError();
}

hipError_t launch_unpack(const DevParams& p, const double* F, double* planar, int fill_ghosts,
                         double ghost_value, hipStream_t s) {
  const long long n = (long long)p.nx * p.ny;
  hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, F,
                     planar, fill_ghosts, ghost_value);
  return hipGetLastError();
}

hipError_t launch_reduce(const DevParams& p, const double* F, StepCtl* ctl, int slot,
                         hipStream_t s) {
  hipLaunchKernelGGL(clear_slot_kernel, dim3(1), dim3(1), 0, s, ctl, slot);
  const int segs = (p.nx - 2 * kBcSize + kRedThreads - 1) / kRedThreads;
  const int blocks = std::min(segs * (p.ny - 2 * kBcSize), 2048);
  hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)blocks), dim3(kRedThreads), 0, s, p, F, ctl,
                     slot);
  return hipGetLastError();
}

hipError_t launch_absmax(const DevParams& p, int frame, const double* F, unsigned long long* out,
                         hipStream_t s) {
  const int ni = frame ? p.nx : p.nx - 2 * kBcSize, nj = frame ? p.ny : p.ny - 2 * kBcSize;
  const int blocks = std::min((ni + kRedThreads - 1) / kRedThreads * nj, 2048);
  hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)blocks), dim3(kRedThreads), 0, s, p, frame, F,
                     out);
  return hipGetLastError();
}

hipError_t launch_cg_stage(const DevParams& p, int mixed, int stage, const StageArgs& a,
                           hipStream_t s) {
  if (mixed && stage == 0) return hipErrorInvalidValue;  // time derivatives are FP64
  const int ni = p.nx - 2 * kBcSize, nj = p.ny - 2 * kBcSize;
  const dim3 grid((ni + cg::kTX - 1) / cg::kTX, (nj + cg::kTY - 1) / cg::kTY);
#define DIM2D_CG(M, S) hipLaunchKernelGGL((cg_stage_kernel<M, S>), grid, dim3(cg::kThreads), 0, s, p, a)
  if (!mixed) {
    switch (stage) {
      case 0: DIM2D_CG(0, 0); break;
      case 1: DIM2D_CG(0, 1); break;
      case 2: DIM2D_CG(0, 2); break;
      default: DIM2D_CG(0, 3); break;
    }
  } else {
    switch (stage) {
      case 1: DIM2D_CG(1, 1); break;
      case 2: DIM2D_CG(1, 2); break;
      default: DIM2D_CG(1, 3); break;
    }
  }
#undef DIM2D_CG
  return hipGetLastError();
}

hipError_t launch_undermined(const DevParams& p, int dist, double md_thr, const double* U,
                             int* counts, hipStream_t s) {
  hipLaunchKernelGGL(undermined_kernel, dim3(1024), dim3(256), 0, s, p, dist, md_thr, U, counts);
  return hipGetLastError();
}

hipError_t launch_undermined_rects(const DevParams& p, const Rects4& r, double md_thr,
                                   const double* U, int* counts, hipStream_t s) {
  hipLaunchKernelGGL(undermined_rects_kernel, dim3(1024), dim3(256), 0, s, p, r, md_thr, U,
                     counts);
  return hipGetLastError();
}

hipError_t launch_extend(const DevParams& np, const DevParams& op, int add_w, int add_s,
                         const double* oldU, double* newU, hipStream_t s) {
  const long long n = (long long)np.nx * np.ny;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(extend_kernel, dim3((unsigned)blocks), dim3(256), 0, s, np, op, add_w, add_s,
                     oldU, newU);
  return hipGetLastError();
}

int max_resident_pair_workgroups(int mixed) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  int per = 0;
  hipError_t e = mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_pair_batch_kernel<1>, 2 * kLanes, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_pair_batch_kernel<0>, 2 * kLanes, 0);
  if (e != hipSuccess || per <= 0) per = 4;
  return ncu * per;
}

int max_resident_mix_workgroups(int mixed) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  int per = 0;
  hipError_t e = mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_mix_batch_kernel<1>, 2 * kLanes, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_mix_batch_kernel<0>, 2 * kLanes, 0);
  if (e != hipSuccess || per <= 0) per = 4;
  return ncu * per;
}

int max_resident_workgroups(int mixed) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  // the stage with the most registers bounds the persistent grid of all three
  int per = 0;
  hipError_t e = mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_batch_kernel<1, 3>, kLanes, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per, stage_batch_kernel<0, 2>, kLanes, 0);
  if (e != hipSuccess || per <= 0) per = 4;
  return ncu * per;
}

}  // namespace dim2d
