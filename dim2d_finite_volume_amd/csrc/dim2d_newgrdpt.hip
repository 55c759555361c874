// dim2d_newgrdpt.hip -- new grid points of an open grid grown by one layer:
// the characteristic extrapolation in time of the reference's buffer layers
// (boundary_conditions/open_bc/buffer_layer/bf_layer_newgrdpt/
// bf_newgrdpt_class.f), one thread per new point, FP64.
//
// Inputs of one extension: the state at t (the old points, already moved to
// their place in the new grid's field) and, at t - dt, the two outermost
// rows and columns of the old grid (the frame snapshot written at the start
// of the last step).  Every procedure reads the points of the old grid only:
// the transverse gradients at t - dt sit in the old frame, those at t in the
// old grid.  The procedure of a new point is the reference's procedure for
// its configuration of old points (bf_newgrdpt_procedure_module.f,
// get_config_id / get_config_procedures); the configurations of a rectangle
// grown by one layer are the four edges (x or y direction, centered or
// one-sided transverse gradient) and the four corners (diagonal direction).
#include "dim2d_internal.hpp"

namespace dim2d {
namespace {

constexpr int kGradI = 0, kGradL0 = 1, kGradR0 = 2;

struct NgArgs {
  int pitch;            // cells per row of the new field
  int nxn, nyn;         // new extents
  int nxo, nyo;         // old extents
  int aw, as;           // layers added west and south (0 or 1)
  int gravity_ac;
  double gravity, cv;
  double x0, y0, dx, dy;  // coordinates of the new grid's point (0, 0), spacings
  double ff[4];         // conservative far-field state
  const StepCtl* ctl;   // dt of the last step
  const double* U;      // new field: the state at t (old points in place)
  const double* F0;     // old grid's frame at t - dt
  double* out;          // new field (new points written)
};

// index of an old point (io, jo) in the frame snapshot, -1 outside the frame:
// rows 0, 1 (nx each), rows ny-2, ny-1, then columns 0, 1 and nx-2, nx-1 of
// rows 2 .. ny-3
__host__ __device__ inline long long frame_cell(int nx, int ny, int io, int jo) {
  if (io < 0 || io >= nx || jo < 0 || jo >= ny) return -1;
  if (jo < 2) return (long long)jo * nx + io;
  if (jo >= ny - 2) return 2LL * nx + (long long)(jo - (ny - 2)) * nx + io;
  if (io < 2) return 4LL * nx + (long long)(jo - 2) * 2 + io;
  if (io >= nx - 2) return 4LL * nx + 2LL * (ny - 4) + (long long)(jo - 2) * 2 + (io - (nx - 2));
  return -1;
}

// the state of new-grid point (i, j) at t (T = 1) or t - dt (T = 0); NaN for
// points the snapshot or the old grid do not hold, so a stencil that strays
// shows as NaN instead of reading out of bounds
template <int T>
__device__ void load(const NgArgs& a, int i, int j, double v[4]) {
  const int io = i - a.aw, jo = j - a.as;
  long long c = -1;
  if (T == 0) {
    c = frame_cell(a.nxo, a.nyo, io, jo);
    if (c >= 0) c *= 4;
  } else if (io >= 0 && io < a.nxo && jo >= 0 && jo < a.nyo) {
    c = ((long long)j * a.pitch + i) * 4;
  }
  const double* src = T == 0 ? a.F0 : a.U;
  for (int k = 0; k < 4; ++k) v[k] = c >= 0 ? src[c + k] : __builtin_nan("");
}

__device__ inline double sq(double x) { return x * x; }

// classical_pressure_local (dim2d_prim_module.f)
__device__ inline double pressure(const NgArgs& a, const double v[4]) {
  return 3.0 / ((3.0 - v[0]) * a.cv) *
             (v[3] - 0.5 * v[0] * (sq(v[1] / v[0]) + sq(v[2] / v[0])) + 3.0 * sq(v[0])) -
         3.0 * sq(v[0]);
}

// speed_of_sound (dim2d_prim_module.f)
__device__ inline double sound_speed(const NgArgs& a, const double v[4]) {
  const double P = pressure(a, v);
  const double e = P + sq(v[0]) * (-3.0 + 2.0 * v[0]);
  const double b = sqrt((P + 3.0 * sq(v[0])) / (a.cv * e));
  const double al = 1.0 / sqrt(1.0 + sq(b));
  const double d = sqrt(v[0] * v[0] * v[0]) * sqrt((3.0 - v[0]) / e);
  return sqrt(3.0) * v[0] / (al * d);
}

// compute_prim_var_ns_vdw2d: [rho, u, v, P]
__device__ inline void prim_var(const NgArgs& a, const double v[4], double p[4]) {
  p[0] = v[0];
  p[1] = v[1] / v[0];
  p[2] = v[2] / v[0];
  p[3] = pressure(a, v);
}

// compute_cons_var_ns_vdw2d
__device__ inline void cons_var(const NgArgs& a, const double p[4], double v[4]) {
  const double T = (p[3] + 3.0 * sq(p[0])) * (3.0 - p[0]) / (8.0 * p[0]);
  v[0] = p[0];
  v[1] = p[0] * p[1];
  v[2] = p[0] * p[2];
  v[3] = 0.5 * p[0] * (sq(p[1]) + sq(p[2])) + p[0] * (8.0 / 3.0 * a.cv * T - 3.0 * p[0]);
}

// compute_body_forces(prim=.true.): -g on the y velocity
__device__ inline void body_prim(const NgArgs& a, double b[4]) {
  b[0] = 0.0;
  b[1] = 0.0;
  b[2] = a.gravity_ac ? -a.gravity : 0.0;
  b[3] = 0.0;
}

// get_prim_obc_eigenqties (obc_eigenqties_bc): [rho, ux, uy, P, c] of the point
__device__ inline void eigenqties(const NgArgs& a, const double v[4], double pe[5]) {
  pe[0] = v[0];
  pe[1] = v[1] / v[0];
  pe[2] = v[2] / v[0];
  pe[3] = pressure(a, v);
  pe[4] = sound_speed(a, v);
}

// compute_xy_to_n_var / compute_n_to_xy_var
__device__ inline void xy_to_n(double b[4]) {
  const double a1 = b[1], a2 = b[2];
  b[1] = 0.5 * sqrt(2.0) * (a1 - a2);
  b[2] = 0.5 * sqrt(2.0) * (a1 + a2);
}
__device__ inline void n_to_xy(double b[4]) {
  const double a1 = b[1], a2 = b[2];
  b[1] = 0.5 * sqrt(2.0) * (a1 + a2);
  b[2] = 0.5 * sqrt(2.0) * (a2 - a1);
}

// eigenvalues and left/right eigenvectors of the primitive variables in x
// (ydir 0) or y (ydir 1) (ns_vdw2d_prim_module.f:504-891); L[r][k] is the
// component r of left eigenvector k, R[k][c] the component c of right
// eigenvector k
__device__ void eigen(int ydir, double md, double un, double c, double lam[4], double L[4][4],
                      double R[4][4]) {
  const double q_c = 0.5 * md * c, inv_c = 1.0 / (c * c), inv_q_c = 1.0 / (md * c);
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) L[r][k] = R[r][k] = 0.0;
  lam[0] = un;
  lam[1] = un;
  lam[2] = un - c;
  lam[3] = un + c;
  const int nv = ydir ? 2 : 1, tv = ydir ? 1 : 2;  // normal and transverse velocity slots
  L[tv][0] = 1.0;
  L[0][1] = 1.0;
  L[3][1] = -1.0 / (c * c);
  L[nv][2] = -q_c;
  L[3][2] = 0.5;
  L[nv][3] = q_c;
  L[3][3] = 0.5;
  R[0][tv] = 1.0;
  R[1][0] = 1.0;
  R[2][0] = inv_c;
  R[3][0] = inv_c;
  R[2][nv] = -inv_q_c;
  R[3][nv] = inv_q_c;
  R[2][3] = 1.0;
  R[3][3] = 1.0;
}

// transverse term MATMUL(gradient, transM) of compute_x_transM (ydir 0:
// gradient along y, advected by uy) or compute_y_transM (ydir 1)
__device__ inline void trans_term(int ydir, const double pe[5], const double g[4],
                                  double out[4]) {
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

// compute_gradient_prim with gradient_{x,y}_interior or the one-sided
// gradient_{x_x,y_y}_oneside_{L0,R0} at (i, j)
template <int T>
__device__ void gradient_prim(const NgArgs& a, int i, int j, int ydir, int type, double dn,
                              double g[4]) {
  const int di = ydir ? 0 : 1, dj = ydir ? 1 : 0;
  int lo, hi;
  double w;
  if (type == kGradI) {
    lo = -1;
    hi = 1;
    w = 0.5 / dn;
  } else if (type == kGradL0) {
    lo = 0;
    hi = 1;
    w = 1.0 / dn;
  } else {
    lo = -1;
    hi = 0;
    w = 1.0 / dn;
  }
  double va[4], vb[4], pa[4], pb[4];
  load<T>(a, i + lo * di, j + lo * dj, va);
  load<T>(a, i + hi * di, j + hi * dj, vb);
  prim_var(a, va, pa);
  prim_var(a, vb, pb);
  for (int k = 0; k < 4; ++k) g[k] = w * (-pa[k] + pb[k]);
}

__device__ inline void line_coeff(double xa, double xb, const double na[4], const double nb[4],
                                  double c[8]) {
  for (int k = 0; k < 4; ++k) {
    c[k] = (nb[k] - na[k]) / (xb - xa);
    c[4 + k] = na[k] - c[k] * xa;
  }
}

__device__ inline void line_at(double x, const double c[8], double o[4]) {
  for (int k = 0; k < 4; ++k) o[k] = x * c[k] + c[4 + k];
}

// get_interpolation_coeff_2D: the plane through (x_l, y_l, n_l[k])
__device__ void plane_coeff(const double x[3], const double y[3], const double n[3][4],
                            double c[12]) {
  for (int k = 0; k < 4; ++k) {
    const double ab0 = x[1] - x[0], ab1 = y[1] - y[0], ab2 = n[1][k] - n[0][k];
    const double ac0 = x[2] - x[0], ac1 = y[2] - y[0], ac2 = n[2][k] - n[0][k];
    const double n0 = ab1 * ac2 - ab2 * ac1;
    const double n1 = -ab0 * ac2 + ab2 * ac0;
    const double n2 = ab0 * ac1 - ab1 * ac0;
    c[k] = -n0 / n2;
    c[4 + k] = -n1 / n2;
    c[8 + k] = -x[0] * c[k] - y[0] * c[4 + k] + n[0][k];
  }
}

__device__ inline void plane_at(double x, double y, const double c[12], double o[4]) {
  for (int k = 0; k < 4; ++k) o[k] = x * c[k] + y * c[4 + k] + c[8 + k];
}

// characteristic amplitudes -> new point: amp = n_amp0 - trapezoid(t - b)
// projected on the left eigenvectors, recombined with the right ones
__device__ inline double char_amp(const double n0[4], const double t0[4], const double b0[4],
                                  const double t1[4], const double b1[4], double dt,
                                  const double L[4][4], int k) {
  double acc = 0.0;
  for (int l = 0; l < 4; ++l) {
    const double d0 = t0[l] - b0[l], d1 = t1[l] - b1[l];
    const double integ = (d0 + 0.5 * (d1 - d0)) * dt;
    acc += (n0[l] - integ) * L[l][k];
  }
  return acc;
}

// compute_newgrdpt_x (ydir 0) / compute_newgrdpt_y (ydir 1) at new point
// (i1, j1); side_right: the grid lies on the low side of the point
__device__ void newgrdpt_dir(const NgArgs& a, double dt, int i1, int j1, int ydir,
                             int side_right, int grad, double out[4]) {
  const int p1 = ydir ? j1 : i1;
  const int pe_ = side_right ? p1 - 1 : p1 + 1;
  const int pa = side_right ? pe_ - 1 : pe_, pb = side_right ? pe_ : pe_ + 1;
  const double org = ydir ? a.y0 : a.x0, h = ydir ? a.dy : a.dx;
  const double dn_t = ydir ? a.dx : a.dy;
  auto at = [&](int p, int& i, int& j) {
    i = ydir ? i1 : p;
    j = ydir ? p : j1;
  };
  const double s1 = org + p1 * h, sa = org + pa * h, sb = org + pb * h;
  double va[4], vb[4], pra[4], prb[4], pe[5], g[4], ta[4], tb[4];
  double c_n0[8], c_t0[8], c_b0[8], c_t1[8], c_b1[8];
  int ia, ja, ib, jb;
  at(pa, ia, ja);
  at(pb, ib, jb);
  // t - dt
  load<0>(a, ia, ja, va);
  load<0>(a, ib, jb, vb);
  prim_var(a, va, pra);
  prim_var(a, vb, prb);
  line_coeff(sa, sb, pra, prb, c_n0);
  eigenqties(a, va, pe);
  gradient_prim<0>(a, ia, ja, !ydir, grad, dn_t, g);
  trans_term(ydir, pe, g, ta);
  eigenqties(a, vb, pe);
  gradient_prim<0>(a, ib, jb, !ydir, grad, dn_t, g);
  trans_term(ydir, pe, g, tb);
  line_coeff(sa, sb, ta, tb, c_t0);
  double bf[4];
  body_prim(a, bf);
  line_coeff(sa, sb, bf, bf, c_b0);
  // t
  load<1>(a, ia, ja, va);
  load<1>(a, ib, jb, vb);
  eigenqties(a, va, pe);
  gradient_prim<1>(a, ia, ja, !ydir, grad, dn_t, g);
  trans_term(ydir, pe, g, ta);
  eigenqties(a, vb, pe);
  gradient_prim<1>(a, ib, jb, !ydir, grad, dn_t, g);
  trans_term(ydir, pe, g, tb);
  line_coeff(sa, sb, ta, tb, c_t1);
  line_coeff(sa, sb, bf, bf, c_b1);
  double t1[4], b1[4];
  line_at(s1, c_t1, t1);
  line_at(s1, c_b1, b1);
  // eigen-decomposition at the neighbour at t
  int ie, je;
  at(pe_, ie, je);
  double ve[4];
  load<1>(a, ie, je, ve);
  eigenqties(a, ve, pe);
  double lam[4], L[4][4], R[4][4];
  eigen(ydir, pe[0], ydir ? pe[2] : pe[1], pe[4], lam, L, R);
  double ffp[4], zero[4] = {0.0, 0.0, 0.0, 0.0};
  prim_var(a, a.ff, ffp);
  double ca[4];
  for (int k = 0; k < 4; ++k) {
    const double s0 = s1 - lam[k] * dt;
    const bool inside = side_right ? (lam[k] >= 0) : !(lam[k] > 0);
    if (inside) {
      double n0[4], t0[4], b0[4];
      line_at(s0, c_n0, n0);
      line_at(s0, c_t0, t0);
      line_at(s0, c_b0, b0);
      ca[k] = char_amp(n0, t0, b0, t1, b1, dt, L, k);
    } else {
      ca[k] = char_amp(ffp, zero, bf, t1, b1, dt, L, k);
    }
  }
  double p[4];
  for (int c = 0; c < 4; ++c) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += ca[k] * R[k][c];
    p[c] = acc;
  }
  cons_var(a, p, out);
}

// get_n_transverse_data_for_interpolation at the three points
template <int T>
__device__ void n_transverse(const NgArgs& a, const int idx[6], int n2dir, const int grads[6],
                             double o[3][4]) {
  for (int k = 0; k < 3; ++k) {
    const int i = idx[2 * k], j = idx[2 * k + 1];
    double xg[4], yg[4], ng[4], v[4], pe[5];
    gradient_prim<T>(a, i, j, 0, grads[2 * k], a.dx, xg);
    gradient_prim<T>(a, i, j, 1, grads[2 * k + 1], a.dy, yg);
    // the gradient across the direction: d/dn2 for n1, d/dn1 for n2
    for (int l = 0; l < 4; ++l)
      ng[l] = n2dir ? sqrt(2.0) * 0.5 * (xg[l] - yg[l]) : sqrt(2.0) * 0.5 * (xg[l] + yg[l]);
    xy_to_n(ng);
    load<T>(a, i, j, v);
    eigenqties(a, v, pe);
    xy_to_n(pe);
    trans_term(n2dir, pe, ng, o[k]);
  }
}

// compute_newgrdpt_xy at new point (i1, j1): direction n2 (else n1), side,
// the gradient procedures (x, y) of the three points, the eigen point and
// the three points (i, j)
__device__ void newgrdpt_xy(const NgArgs& a, double dt, int i1, int j1, int n2dir,
                            int side_right, const int grads[6], int ie, int je,
                            const int idx[6], double out[4]) {
  const double r = 0.5 * sqrt(2.0);
  auto n1c = [&](double x, double y) { return r * (x - y); };
  auto n2c = [&](double x, double y) { return r * (x + y); };
  const double x1 = a.x0 + i1 * a.dx, y1 = a.y0 + j1 * a.dy;
  const double n1_1 = n1c(x1, y1), n2_1 = n2c(x1, y1);
  double xs[3], ys[3], nodes[3][4], c_n0[12], c_t0[12], c_b0[12], c_t1[12], c_b1[12];
  for (int k = 0; k < 3; ++k) {
    const double x = a.x0 + idx[2 * k] * a.dx, y = a.y0 + idx[2 * k + 1] * a.dy;
    xs[k] = n1c(x, y);
    ys[k] = n2c(x, y);
    double v[4];
    load<0>(a, idx[2 * k], idx[2 * k + 1], v);
    prim_var(a, v, nodes[k]);
    xy_to_n(nodes[k]);
  }
  plane_coeff(xs, ys, nodes, c_n0);
  n_transverse<0>(a, idx, n2dir, grads, nodes);
  plane_coeff(xs, ys, nodes, c_t0);
  double bf[4];
  body_prim(a, bf);
  xy_to_n(bf);
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 4; ++l) nodes[k][l] = bf[l];
  plane_coeff(xs, ys, nodes, c_b0);
  plane_coeff(xs, ys, nodes, c_b1);
  n_transverse<1>(a, idx, n2dir, grads, nodes);
  plane_coeff(xs, ys, nodes, c_t1);
  double t1[4], b1[4];
  plane_at(n1_1, n2_1, c_t1, t1);
  plane_at(n1_1, n2_1, c_b1, b1);
  double ve[4], pe[5];
  load<1>(a, ie, je, ve);
  eigenqties(a, ve, pe);
  xy_to_n(pe);
  double lam[4], L[4][4], R[4][4];
  eigen(n2dir, pe[0], n2dir ? pe[2] : pe[1], pe[4], lam, L, R);
  double ffp[4], zero[4] = {0.0, 0.0, 0.0, 0.0};
  prim_var(a, a.ff, ffp);
  xy_to_n(ffp);
  double ca[4];
  for (int k = 0; k < 4; ++k) {
    const double n1_0 = n2dir ? n1_1 : n1_1 - lam[k] * dt;
    const double n2_0 = n2dir ? n2_1 - lam[k] * dt : n2_1;
    const bool inside = side_right ? (lam[k] >= 0) : !(lam[k] > 0);
    if (inside) {
      double n0[4], t0[4], b0[4];
      plane_at(n1_0, n2_0, c_n0, n0);
      plane_at(n1_0, n2_0, c_t0, t0);
      plane_at(n1_0, n2_0, c_b0, b0);
      ca[k] = char_amp(n0, t0, b0, t1, b1, dt, L, k);
    } else {
      ca[k] = char_amp(ffp, zero, bf, t1, b1, dt, L, k);
    }
  }
  double p[4];
  for (int c = 0; c < 4; ++c) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += ca[k] * R[k][c];
    p[c] = acc;
  }
  n_to_xy(p);
  cons_var(a, p, out);
}

__global__ void __launch_bounds__(64) newgrdpt_kernel(const NgArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.nxn, ny = a.nyn;
  int i, j;
  if (t < nx) {
    i = t;
    j = 0;
  } else if (t < 2 * nx) {
    i = t - nx;
    j = ny - 1;
  } else if (t < 2 * nx + ny) {
    i = 0;
    j = t - 2 * nx;
  } else if (t < 2 * (nx + ny)) {
    i = nx - 1;
    j = t - 2 * nx - ny;
  } else {
    return;
  }
  const int io = i - a.aw, jo = j - a.as;
  if (io >= 0 && io < a.nxo && jo >= 0 && jo < a.nyo) return;  // an old point
  // rows own the corner points; a column thread on a new row returns
  if (t >= 2 * nx && (jo < 0 || jo >= a.nyo)) return;
  // configuration of old points around (get_config_id: SW, S, SE, W, E, NW, N, NE)
  int config = 0, bit = 0;
  for (int dj = -1; dj <= 1; ++dj)
    for (int di = -1; di <= 1; ++di) {
      if (!di && !dj) continue;
      const int a0 = io + di, b0 = jo + dj;
      if (a0 >= 0 && a0 < a.nxo && b0 >= 0 && b0 < a.nyo) config |= 1 << bit;
      ++bit;
    }
  const double dt = a.ctl->dt;
  double v[4];
  switch (config) {
    // edges (get_config_procedures: E 41/40/9, W 148/144/20, N 7/6/3, S 224/192/96)
    case 41: newgrdpt_dir(a, dt, i, j, 0, 1, kGradI, v); break;
    case 40: newgrdpt_dir(a, dt, i, j, 0, 1, kGradL0, v); break;
    case 9: newgrdpt_dir(a, dt, i, j, 0, 1, kGradR0, v); break;
    case 148: newgrdpt_dir(a, dt, i, j, 0, 0, kGradI, v); break;
    case 144: newgrdpt_dir(a, dt, i, j, 0, 0, kGradL0, v); break;
    case 20: newgrdpt_dir(a, dt, i, j, 0, 0, kGradR0, v); break;
    case 7: newgrdpt_dir(a, dt, i, j, 1, 1, kGradI, v); break;
    case 6: newgrdpt_dir(a, dt, i, j, 1, 1, kGradL0, v); break;
    case 3: newgrdpt_dir(a, dt, i, j, 1, 1, kGradR0, v); break;
    case 224: newgrdpt_dir(a, dt, i, j, 1, 0, kGradI, v); break;
    case 192: newgrdpt_dir(a, dt, i, j, 1, 0, kGradL0, v); break;
    case 96: newgrdpt_dir(a, dt, i, j, 1, 0, kGradR0, v); break;
    // corners (compute_newgrdpt_local, bf_newgrdpt_class.f:245-420)
    case 128: {  // SW: n2, left
      const int gr[6] = {kGradL0, kGradL0, kGradI, kGradL0, kGradL0, kGradI};
      const int idx[6] = {i + 1, j + 1, i + 2, j + 1, i + 1, j + 2};
      newgrdpt_xy(a, dt, i, j, 1, 0, gr, i + 1, j + 1, idx, v);
      break;
    }
    case 32: {  // SE: n1, right
      const int gr[6] = {kGradI, kGradL0, kGradR0, kGradL0, kGradR0, kGradI};
      const int idx[6] = {i - 2, j + 1, i - 1, j + 1, i - 1, j + 2};
      newgrdpt_xy(a, dt, i, j, 0, 1, gr, i - 1, j + 1, idx, v);
      break;
    }
    case 4: {  // NW: n1, left
      const int gr[6] = {kGradL0, kGradI, kGradL0, kGradR0, kGradI, kGradR0};
      const int idx[6] = {i + 1, j - 2, i + 1, j - 1, i + 2, j - 1};
      newgrdpt_xy(a, dt, i, j, 0, 0, gr, i + 1, j - 1, idx, v);
      break;
    }
    case 1: {  // NE: n2, right
      const int gr[6] = {kGradR0, kGradI, kGradI, kGradR0, kGradR0, kGradR0};
      const int idx[6] = {i - 1, j - 2, i - 2, j - 1, i - 1, j - 1};
      newgrdpt_xy(a, dt, i, j, 1, 1, gr, i - 1, j - 1, idx, v);
      break;
    }
    default:
      for (int k = 0; k < 4; ++k) v[k] = __builtin_nan("");
  }
  double* dst = a.out + ((long long)j * a.pitch + i) * 4;
  for (int k = 0; k < 4; ++k) dst[k] = v[k];
}

// the two outermost rows and columns of U into the frame snapshot
__global__ void __launch_bounds__(256) save_frame_kernel(int nx, int ny, int pitch,
                                                         const double* U, double* F0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long rows = 4LL * nx, cols = 4LL * (ny - 4);
  int io, jo;
  if (t < rows) {
    const int r = (int)(t / nx);
    io = (int)(t % nx);
    jo = r < 2 ? r : ny - 4 + r;
  } else if (t < rows + cols) {
    const long long q = t - rows;
    const int side = q < 2LL * (ny - 4) ? 0 : 1;
    const long long q2 = side ? q - 2LL * (ny - 4) : q;
    jo = 2 + (int)(q2 / 2);
    io = side ? nx - 2 + (int)(q2 % 2) : (int)(q2 % 2);
  } else {
    return;
  }
  const double4* s = reinterpret_cast<const double4*>(U) + ((long long)jo * pitch + io);
  reinterpret_cast<double4*>(F0)[frame_cell(nx, ny, io, jo)] = *s;
}

}  // namespace

long long frame_cells(int nx, int ny) { return 4LL * nx + 4LL * (ny - 4); }

hipError_t launch_save_frame(const DevParams& p, const double* U, double* F0, hipStream_t s) {
  const long long n = frame_cells(p.nx, p.ny);
  const int block = 256;
  hipLaunchKernelGGL(save_frame_kernel, dim3((unsigned)((n + block - 1) / block)), dim3(block),
                     0, s, p.nx, p.ny, p.pitch, U, F0);
  return hipGetLastError();
}

hipError_t launch_newgrdpt(const DevParams& np, const DevParams& op, int aw, int as, double x0,
                           double y0, const double ff[4], const StepCtl* ctl, const double* F0,
                           double* U, hipStream_t s) {
  NgArgs a;
  a.pitch = np.pitch;
  a.nxn = np.nx;
  a.nyn = np.ny;
  a.nxo = op.nx;
  a.nyo = op.ny;
  a.aw = aw;
  a.as = as;
  a.gravity_ac = np.gravity_ac;
  a.gravity = np.gravity;
  a.cv = np.cv;
  a.x0 = x0;
  a.y0 = y0;
  a.dx = np.dx;
  a.dy = np.dy;
  for (int k = 0; k < 4; ++k) a.ff[k] = ff[k];
  a.ctl = ctl;
  a.U = U;
  a.F0 = F0;
  a.out = U;
  const int n = 2 * (np.nx + np.ny);
  const int block = 64;
  hipLaunchKernelGGL(newgrdpt_kernel, dim3((n + block - 1) / block), dim3(block), 0, s, a);
  return hipGetLastError();
}

}  // namespace dim2d
