// dim2d_capi.hip -- C ABI of the DIM RK3-TVD library (include/dim2d_hip.h).
//
// A context owns one simulation: the HBM-resident state u^n, the stage
// buffers, a step-control block and one HIP stream.  Every RK stage is one
// fused kernel launch (plus a small launch for wall ghost cells when a wall
// edge exists); a step never synchronises with the host.
#include "dim2d_hip.h"
#include "dim2d_internal.hpp"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using namespace dim2d;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(DIM2D_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));          \
  } while (0)

int validate(const dim2d_params* p) {
  if (!p) return fail(DIM2D_EINVAL, "params is NULL");
  if (p->bc_size != kBcSize)
    return fail(DIM2D_EINVAL, "bc_size must be 2 (Mattsson operators), got %d", p->bc_size);
  if (p->nx < 2 * kBcSize + 3 || p->ny < 2 * kBcSize + 3)
    return fail(DIM2D_EINVAL, "grid %dx%d too small: need at least 3 interior points per axis",
                p->nx, p->ny);
  if (p->nx > kMaxExtent || p->ny > kMaxExtent)
    return fail(DIM2D_EINVAL, "grid %dx%d too large: at most %d points per axis", p->nx, p->ny,
                kMaxExtent);
  const int b[4] = {p->bc_w, p->bc_e, p->bc_s, p->bc_n};
  int nopen = 0;
  for (int k = 0; k < 4; ++k) {
    if (b[k] < DIM2D_BC_REFLECTION || b[k] > DIM2D_BC_OPEN)
      return fail(DIM2D_EINVAL, "unknown boundary kind %d", b[k]);
    nopen += b[k] == DIM2D_BC_OPEN;
  }
  if (nopen != 0 && nopen != 4)
    return fail(DIM2D_EINVAL,
                "open (Hedstrom) boundaries are supported on all four edges at once "
                "(hedstrom_xy_choice)");
  if ((p->bc_w == DIM2D_BC_PERIODIC) != (p->bc_e == DIM2D_BC_PERIODIC) ||
      (p->bc_s == DIM2D_BC_PERIODIC) != (p->bc_n == DIM2D_BC_PERIODIC))
    return fail(DIM2D_EINVAL, "periodic boundaries must pair W/E and S/N");
  const bool wall_x = p->bc_w == DIM2D_BC_WALL || p->bc_e == DIM2D_BC_WALL;
  const bool wall_y = p->bc_s == DIM2D_BC_WALL || p->bc_n == DIM2D_BC_WALL;
  if (wall_x && wall_y)
    return fail(DIM2D_EINVAL,
                "walls on both axes (wall_xy) are not supported, as in the reference's "
                "library_config.py");
  if (p->precision != DIM2D_PREC_FP64 && p->precision != DIM2D_PREC_MIXED)
    return fail(DIM2D_EINVAL, "unknown precision mode %d", p->precision);
  if (p->integrator != DIM2D_INT_RK3TVD && p->integrator != DIM2D_INT_LSRK3)
    return fail(DIM2D_EINVAL, "unknown integrator %d", p->integrator);
  if (p->operators != DIM2D_OPS_MATTSSON && p->operators != DIM2D_OPS_COCKBURN_GAU)
    return fail(DIM2D_EINVAL, "unknown operator set %d", p->operators);
  if (p->operators == DIM2D_OPS_COCKBURN_GAU) {
    if (p->integrator != DIM2D_INT_RK3TVD)
      return fail(DIM2D_EINVAL, "the Cockburn-Gau operators run with RK3-TVD only");
    for (int k = 0; k < 4; ++k)
      if (b[k] == DIM2D_BC_WALL || b[k] == DIM2D_BC_OPEN)
        return fail(DIM2D_EINVAL,
                    "the Cockburn-Gau operators support reflection and periodic edges only");
  }
  if (!(p->x_max > p->x_min) || !(p->y_max > p->y_min))
    return fail(DIM2D_EINVAL, "empty coordinate range");
  if (!(p->re > 0) || !(p->we > 0) || !(p->pr > 0) || !(p->cv_r > 0))
    return fail(DIM2D_EINVAL, "Re, We, Pr and cv_r must be positive");
  return DIM2D_OK;
}

// coordinates as field_abstract_class.f:286-315 and dx, dy as td_operators_class.f:156-157
DevParams make_dev_params(const dim2d_params& q) {
  DevParams p;
  memset(&p, 0, sizeof p);
  p.nx = q.nx;
  p.ny = q.ny;
  p.pitch = (q.nx + 31) / 32 * 32;
  p.ncell = (long long)p.pitch * q.ny;
  p.bc_w = q.bc_w;
  p.bc_e = q.bc_e;
  p.bc_s = q.bc_s;
  p.bc_n = q.bc_n;
  p.gravity_ac = q.gravity_ac;
  const double dxf = (q.x_max - q.x_min) / (q.nx - 1 - 2 * kBcSize);
  const double dyf = (q.y_max - q.y_min) / (q.ny - 1 - 2 * kBcSize);
  const double x1 = q.x_min + (0 - kBcSize) * dxf, x2 = q.x_min + (1 - kBcSize) * dxf;
  const double y1 = q.y_min + (0 - kBcSize) * dyf, y2 = q.y_min + (1 - kBcSize) * dyf;
  p.dx = x2 - x1;
  p.dy = y2 - y1;
  p.idx = 1.0 / p.dx;
  p.idy = 1.0 / p.dy;
  p.half_idx = 0.5 / p.dx;
  p.half_idy = 0.5 / p.dy;
  p.q_idx = 0.25 / p.dx;
  p.q_idy = 0.25 / p.dy;
  p.h2x = 0.5 / (p.dx * p.dx);
  p.h2y = 0.5 / (p.dy * p.dy);
  p.re = q.re;
  p.we = q.we;
  p.pr = q.pr;
  p.vr = q.viscous_r;
  p.cv = q.cv_r;
  p.inv_re = 1.0 / q.re;
  p.inv_we = 1.0 / q.we;
  p.inv_pr = 1.0 / q.pr;
  p.inv_repr = 1.0 / (q.re * q.pr);
  p.c15cv = 1.5 / q.cv_r;
  p.c3cv = 3.0 / q.cv_r;
  p.c3cv_3 = p.c3cv / 3.0;  // rounded once here, not divided per cell
  p.c38cv = 3.0 / (8.0 * q.cv_r);
  p.two_vr = 2.0 + q.viscous_r;
  p.two_inv_re = 2.0 / q.re;
  p.two_inv_we = 2.0 / q.we;
  p.k_heat_x = 2.0 / q.pr * p.idx;
  p.k_heat_y = 2.0 / q.pr * p.idy;
  p.k_cap_x = 2.0 * p.inv_repr * p.idx;
  p.k_cap_y = 2.0 * p.inv_repr * p.idy;
  p.gravity = q.gravity;
  p.wall_cos_angle = cos(q.wall_contact_angle * acos(-1.0) / 180.0);
  p.wall_q = q.wall_heat_flux;
  p.sigma_pref = 2.0 * sqrt(2.0) / (3.0 * sqrt(q.we));
  p.f = DevParams::F32{(float)p.idx,        (float)p.idy,        (float)p.half_idx,
                       (float)p.half_idy,   (float)p.q_idx,      (float)p.q_idy,
                       (float)p.vr,         (float)p.two_vr,     (float)p.inv_re,
                       (float)p.inv_we,     (float)p.two_inv_re, (float)p.two_inv_we,
                       (float)p.c15cv,      (float)p.k_heat_x,   (float)p.k_heat_y,
                       (float)p.k_cap_x,    (float)p.k_cap_y,    (float)p.inv_pr,
                       (float)p.inv_repr,   (float)p.c3cv / 3.0f};
  return p;
}

bool has_wall(const DevParams& p) {
  return p.bc_w == kBcWall || p.bc_e == kBcWall || p.bc_s == kBcWall || p.bc_n == kBcWall;
}

bool is_open(const DevParams& p) { return p.bc_n == kBcOpen; }

// Field buffers start `shift` elements past their allocation so that the
// first interior column of every row (column 2; rows start on 1 KiB
// boundaries) begins a 128-byte line: the 60 output cells of a column strip
// are then written as whole lines (2 cells for FP64, 6 for FP32 increments).
constexpr int kShift64 = 2 * 4;
constexpr int kShift32 = 6 * 4;
template <typename T>
hipError_t field_alloc(T** p, size_t n) {
  constexpr int shift = sizeof(T) == 8 ? kShift64 : kShift32;
  T* raw = nullptr;
  const hipError_t e = hipMalloc(&raw, (n + shift) * sizeof(T));
  *p = raw ? raw + shift : nullptr;
  return e;
}
template <typename T>
void field_free(T* p) {
  constexpr int shift = sizeof(T) == 8 ? kShift64 : kShift32;
  if (p) (void)hipFree(p - shift);
}

}  // namespace

struct dim2d_ctx {
  dim2d_params prm;
  DevParams dp;
  int device = 0;
  int mixed = 0;
  int lsrk = 0;          // low-storage RK3 instead of RK3-TVD
  int cg = 0;            // Cockburn-Gau face stencils (FP64 RK3-TVD, own kernel)
  hipStream_t stream = nullptr;
  // open edges: the ghost-frame kernel of a stage runs on `side`, forked from
  // and joined to `stream` around the stage kernel
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  double* U = nullptr;   // u^n
  double* U2 = nullptr;  // u^{n+1}
  double* A = nullptr;   // FP64 stage 1 output
  double* B = nullptr;   // FP64 stage 2 output
  double* Q = nullptr;   // low-storage RK register q
  float* Q32 = nullptr;  // the same in mixed precision
  float* D1 = nullptr;   // mixed stage increments
  float* D2 = nullptr;
  double* T = nullptr;   // planar staging of uploads, downloads and time derivatives
  StepCtl* ctl = nullptr;
  // open edges with a far field (dim2d_set_far_field): every step first
  // copies the two outermost rows and columns of u^n into F0, the state at
  // t - dt that the new points of a domain extension are extrapolated from
  double ff[4] = {0.0, 0.0, 0.0, 0.0};
  bool ff_set = false;
  double* F0 = nullptr;
  bool f0_valid = false;  // F0 holds the frame of the last step's u^n
  // columns/rows added per edge (W, E, S, N) since creation: the grid the
  // context was created on is the reference's interior domain, the rest its
  // buffer layers
  int grown[4] = {0, 0, 0, 0};
  int block_x = 64;
  int rows = 16;
  // RK3-TVD stages 1 and 2 in one launch (stage pair) where the edges allow
  int fuse = 0;
  int pair_rows = 32;
  bool pair_rows_set = false;
  bool block_x_set = false;  // block_x chosen by dim2d_set_tiles, else by default_block_x()
  bool rows_set = false;     // rows chosen by dim2d_set_tiles, else by default_rows()
  bool uploaded = false;
  long long steps = 0;
  // the reductions of u^n (CFL quantities) sit in slot (steps & 1): written by
  // the last stage of the step that produced u^n, cleared by an upload
  bool red_valid = false;
  // kernel timing: event pairs around launches of the kinds in profile_mask
  // (bit k: kind k = stage 1, 2, 3, boundary kernels)
  bool profiling = false;
  int profile_mask = 0;
  // dim2d_integrate replays captured steps when not profiling: graphs of 8
  // steps, then of 2 steps for the remainder ([0]: 2 steps, [1]: 8 steps);
  // one graph per step parity (which buffer holds u^n and which reduction
  // slot a step reads alternate with it), valid for the dt/cfl/tiles it was
  // made with
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  double g_dt[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  int g_cfl[2][2] = {{0, 0}, {0, 0}};
  struct Timed {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> pool;
};

namespace {

hipEvent_t take_event(dim2d_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// record an event pair around `launch` when profiling (kind 0..3)
template <typename F>
hipError_t timed_launch(dim2d_ctx* c, int kind, F&& launch) {
  if (!c->profiling || !(c->profile_mask & (1 << kind))) return launch();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, c->stream);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(b, c->stream);
  if (e == hipSuccess) c->timed.push_back({a, b, kind});
  return e;
}

void drop_graphs(dim2d_ctx* c) {
  for (int g = 0; g < 2; ++g)
    for (int k = 0; k < 2; ++k)
      if (c->gexec[g][k]) {
        (void)hipGraphExecDestroy(c->gexec[g][k]);
        c->gexec[g][k] = nullptr;
      }
}

void release_timed(dim2d_ctx* c) {
  for (auto& t : c->timed) {
    c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->timed.clear();
}

}  // namespace

extern "C" {

const char* dim2d_last_error(void) { return g_err.c_str(); }

const char* dim2d_version(void) { return "dim2d-mi355x 0.1 (gfx950)"; }

// Default tile height from the grid: a wave runs two warm-up rows before its
// first output row, so tall tiles waste least, but a small grid cut into
// 16-row tiles leaves most of the GPU idle and every wave's row latency on the
// critical path.  Measured on MI355X (two-bubble state, FP64 and mixed):
// 128^2 is fastest with 2 rows (1.8x the 16-row step rate), 256^2 with 4
// (1.37x), 512^2 and 1024^2 with 8 (2-7 %), 2048^2 and up with 16.  The rule
// counts the waves of 16-row tiles: fewer than 50, 200, 2000 take 2, 4, 8.
static long long waves16(const DevParams& p) {
  const int strips = (p.nx - 2 * kBcSize + 59) / 60, nj = p.ny - 2 * kBcSize;
  return (long long)strips * ((nj + 15) / 16);
}
static int default_rows(const DevParams& p) {
  const long long w = waves16(p);
  return w < 50 ? 2 : w < 200 ? 4 : w < 2000 ? 8 : 16;
}
// Two strips per workgroup (block_x 128) share the halo line between them in
// one CU's cache, which pays on large grids (-0.7 to -1.2 % per step at
// 4096^2); on small grids the fewer, larger workgroups leave CUs idle
// (+5 to +13 % at 128^2 to 512^2).
constexpr long long kPairStripsWaves = 8000;
static int default_block_x(const DevParams& p) { return waves16(p) >= kPairStripsWaves ? 128 : 64; }
// Tile height of the stage pair: its producer runs 4 rows past each end of the
// tile, so tall tiles waste least; small grids take shorter tiles to fill the
// GPU (workgroups of 32-row tiles: fewer than 500 take 8 rows, fewer than 2000
// take 16).  Measured at 4096^2 (two bubbles, ms per step FP64 / mixed): 32
// rows 0.718 / 0.623, 48 rows 0.699 / 0.608, 60 rows 0.681 / 0.599.
static int default_pair_rows(const DevParams& p) {
  const long long strips = (p.nx - 2 * kBcSize + 55) / 56, nj = p.ny - 2 * kBcSize;
  const long long wg32 = strips * ((nj + 31) / 32);
  return wg32 < 500 ? 8 : wg32 < 2000 ? 16 : 60;
}

// stage fusion (the stage pair) applies to the Mattsson operators on the
// edges pair_supported takes; the low-storage RK3 (FP64) not with open edges
static bool fusion_applies(const dim2d_ctx* c) {
  return !c->cg && pair_supported(c->dp) && !(c->lsrk && is_open(c->dp));
}

int dim2d_ctx_create(const dim2d_params* params, dim2d_ctx** out) {
  if (!out) return fail(DIM2D_EINVAL, "out is NULL");
  *out = nullptr;
  int st = validate(params);
  if (st) return st;
  dim2d_ctx* c = new (std::nothrow) dim2d_ctx();
  if (!c) return fail(DIM2D_ENOMEM, "host allocation failed");
  c->prm = *params;
  c->dp = make_dev_params(*params);
  c->mixed = params->precision == DIM2D_PREC_MIXED;
  c->lsrk = params->integrator == DIM2D_INT_LSRK3;
  c->cg = params->operators == DIM2D_OPS_COCKBURN_GAU;
  c->rows = default_rows(c->dp);
  c->block_x = default_block_x(c->dp);
  c->fuse = fusion_applies(c);
  c->pair_rows = default_pair_rows(c->dp);
  const size_t n = (size_t)c->dp.ncell * 4;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = field_alloc(&c->U, n);
  if (e == hipSuccess) e = field_alloc(&c->U2, n);
  if (e == hipSuccess && !c->mixed && !c->lsrk) e = field_alloc(&c->A, n);
  if (e == hipSuccess && !c->mixed && !c->lsrk) e = field_alloc(&c->B, n);
  if (e == hipSuccess && c->lsrk && !c->mixed) e = field_alloc(&c->Q, n);
  if (e == hipSuccess && c->lsrk && c->mixed) e = field_alloc(&c->Q32, n);
  // the low-storage RK3 keeps its stage states in FP64 in both modes
  if (e == hipSuccess && c->mixed && !c->lsrk) e = field_alloc(&c->D1, n);
  if (e == hipSuccess && c->mixed && !c->lsrk) e = field_alloc(&c->D2, n);
  if (e == hipSuccess) e = hipMalloc(&c->T, n * sizeof(double));
  if (e == hipSuccess && params->bc_n == DIM2D_BC_OPEN) {
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipMalloc(&c->ctl, sizeof(StepCtl));
  if (e == hipSuccess) e = hipMemsetAsync(c->U, 0, n * sizeof(double), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->U2, 0, n * sizeof(double), c->stream);
  if (e == hipSuccess && c->A) e = hipMemsetAsync(c->A, 0, n * sizeof(double), c->stream);
  if (e == hipSuccess && c->B) e = hipMemsetAsync(c->B, 0, n * sizeof(double), c->stream);
  if (e == hipSuccess && c->Q) e = hipMemsetAsync(c->Q, 0, n * sizeof(double), c->stream);
  if (e == hipSuccess && c->Q32) e = hipMemsetAsync(c->Q32, 0, n * sizeof(float), c->stream);
  if (e == hipSuccess && c->D1) e = hipMemsetAsync(c->D1, 0, n * sizeof(float), c->stream);
  if (e == hipSuccess && c->D2) e = hipMemsetAsync(c->D2, 0, n * sizeof(float), c->stream);
  if (e == hipSuccess) {
    StepCtl h;
    memset(&h, 0, sizeof h);
    h.h = std::min(c->dp.dx, c->dp.dy);
    h.re = params->re;
    h.we = params->we;
    e = hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory) ? DIM2D_ENOMEM : DIM2D_EDEVICE;
    dim2d_ctx_destroy(c);
    return fail(code, "context allocation failed: %s", hipGetErrorString(e));
  }
  *out = c;
  return DIM2D_OK;
}

int dim2d_ctx_destroy(dim2d_ctx* c) {
  if (!c) return DIM2D_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  field_free(c->U);
  field_free(c->U2);
  field_free(c->A);
  field_free(c->B);
  field_free(c->Q);
  field_free(c->Q32);
  field_free(c->D1);
  field_free(c->D2);
  (void)hipFree(c->T);
  (void)hipFree(c->ctl);
  (void)hipFree(c->F0);
  release_timed(c);
  drop_graphs(c);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DIM2D_OK;
}

int dim2d_set_tiles(dim2d_ctx* c, int32_t block_x, int32_t rows) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (block_x != 0 && block_x != 64 && block_x != 128)
    return fail(DIM2D_EINVAL,
                "block_x must be 64 or 128 (one or two column strips per workgroup), got %d",
                block_x);
  if (rows < 0 || rows > 4096) return fail(DIM2D_EINVAL, "rows_per_block out of range: %d", rows);
  if (block_x) {
    c->block_x = block_x;
    c->block_x_set = true;
  }
  if (rows) {
    c->rows = rows;
    c->rows_set = true;
  }
  drop_graphs(c);
  return DIM2D_OK;
}

int dim2d_set_stage_fusion(dim2d_ctx* c, int32_t fuse, int32_t rows_per_block) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (fuse != 0 && fuse != 1) return fail(DIM2D_EINVAL, "fuse must be 0 or 1, got %d", fuse);
  if (rows_per_block != 0 && (rows_per_block < 4 || rows_per_block > 4096))
    return fail(DIM2D_EINVAL, "rows_per_block must be 0 or in [4, 4096], got %d", rows_per_block);
  if (fuse && !fusion_applies(c))
    return fail(DIM2D_EINVAL,
                "stage fusion needs the Mattsson operators, reflection, periodic, S/N wall or "
                "(RK3-TVD only) open edges and at least 8 interior points per axis");
  c->fuse = fuse;
  if (rows_per_block) {
    c->pair_rows = rows_per_block;
    c->pair_rows_set = true;
  }
  drop_graphs(c);
  return DIM2D_OK;
}

int dim2d_profile_enable(dim2d_ctx* c, int32_t enable) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (enable < 0 || enable > 15) return fail(DIM2D_EINVAL, "profile mask out of range: %d", enable);
  HIP_TRY(hipStreamSynchronize(c->stream));
  release_timed(c);
  c->profiling = enable != 0;
  c->profile_mask = enable;
  return DIM2D_OK;
}

int dim2d_profile_read(dim2d_ctx* c, double out[8]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  double sum[4] = {0, 0, 0, 0};
  double cnt[4] = {0, 0, 0, 0};
  for (auto& t : c->timed) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    sum[t.kind] += ms;
    cnt[t.kind] += 1;
  }
  for (int k = 0; k < 4; ++k) {
    out[2 * k] = cnt[k] > 0 ? sum[k] / cnt[k] : 0.0;
    out[2 * k + 1] = cnt[k];
  }
  return DIM2D_OK;
}

// host or device planar nodes(nx,ny,4) -> interleaved cells of U
static int copy_in(dim2d_ctx* c, const double* src, hipMemcpyKind kind) {
  const DevParams& p = c->dp;
  const double* planar = src;
  if (kind == hipMemcpyHostToDevice) {
    HIP_TRY(hipMemcpyAsync(c->T, src, (size_t)p.nx * p.ny * 4 * sizeof(double), kind, c->stream));
    planar = c->T;
  }
  HIP_TRY(launch_pack(p, planar, c->U, c->stream));
  c->uploaded = true;
  c->red_valid = false;
  c->f0_valid = false;
  return DIM2D_OK;
}

// U with its ghost cells stored -> host or device planar nodes(nx,ny,4)
static int copy_out(dim2d_ctx* c, double* dst, hipMemcpyKind kind) {
  const DevParams& p = c->dp;
  HIP_TRY(launch_fill_ghosts(p, c->U, c->stream));
  if (kind == hipMemcpyDeviceToHost) {
    HIP_TRY(launch_unpack(p, c->U, c->T, 0, 0.0, c->stream));
    HIP_TRY(hipMemcpyAsync(dst, c->T, (size_t)p.nx * p.ny * 4 * sizeof(double), kind, c->stream));
  } else {
    HIP_TRY(launch_unpack(p, c->U, dst, 0, 0.0, c->stream));
  }
  return DIM2D_OK;
}

int dim2d_upload(dim2d_ctx* c, const double* nodes) {
  if (!c || !nodes) return fail(DIM2D_EINVAL, "NULL argument");
  int st = copy_in(c, nodes, hipMemcpyHostToDevice);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_upload_device(dim2d_ctx* c, const double* d_nodes) {
  if (!c || !d_nodes) return fail(DIM2D_EINVAL, "NULL argument");
  return copy_in(c, d_nodes, hipMemcpyDeviceToDevice);
}

int dim2d_download(dim2d_ctx* c, double* nodes) {
  if (!c || !nodes) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "download before upload");
  int st = copy_out(c, nodes, hipMemcpyDeviceToHost);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_download_device(dim2d_ctx* c, double* d_nodes) {
  if (!c || !d_nodes) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "download before upload");
  return copy_out(c, d_nodes, hipMemcpyDeviceToDevice);
}

int dim2d_apply_bc(dim2d_ctx* c, double t) {
  (void)t;
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "apply_bc before upload");
  HIP_TRY(launch_fill_ghosts(c->dp, c->U, c->stream));
  if (has_wall(c->dp)) {
    // bracket guess from the same state, as at initialisation (field_abstract_class.f:447-463)
    const size_t n = (size_t)c->dp.ncell * 4;
    HIP_TRY(hipMemcpyAsync(c->U2, c->U, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(launch_wall(c->dp, 0, nullptr, c->U, nullptr, c->U, nullptr, c->U2, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

// The time derivative of u^n into U2 (free between steps), in FP64 whatever
// the context's precision: the host-facing derivative replaces
// compute_time_dev, and the steady-state check compares it with limits such
// as 1e-12 (parameters_input.f:913) that FP32 fluxes, ~1e-7 |F|/dx, could
// never meet.
static hipError_t launch_time_dev_fp64(dim2d_ctx* c) {
  const DevParams& p = c->dp;
  StageArgs a;
  memset(&a, 0, sizeof a);
  a.U = c->U;
  a.S64 = c->U;
  a.O64 = c->U2;
  hipError_t e = c->cg ? launch_cg_stage(p, 0, 0, a, c->stream)
                       : launch_stage(p, 0, 0, a, c->block_x, c->rows, c->stream);
  // open edges: the ghost frame has a time derivative of its own
  if (e == hipSuccess && is_open(p)) e = launch_frame(p, 0, 0, a, c->stream);
  return e;
}

int dim2d_time_dev(dim2d_ctx* c, double t, double* time_dev) {
  (void)t;
  if (!c || !time_dev) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "time_dev before upload");
  const DevParams& p = c->dp;
  // U2 is free between steps: the derivative goes there, then to planar order
  // with the ghost cells set to debug_real (td_operators_class.f:150-154)
  // unless the edges are open
  HIP_TRY(launch_time_dev_fp64(c));
  HIP_TRY(launch_unpack(p, c->U2, c->T, is_open(p) ? 0 : 1, kDebugReal, c->stream));
  HIP_TRY(hipMemcpyAsync(time_dev, c->T, (size_t)p.nx * p.ny * 4 * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_check_steady_state(dim2d_ctx* c, double limit, double out[4], int32_t* steady) {
  if (!c || !steady) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "check_steady_state before upload");
  if (!(limit >= 0.0)) return fail(DIM2D_EINVAL, "limit must be >= 0, got %g", limit);
  const DevParams& p = c->dp;
  // the time derivative into U2 (free between steps), as dim2d_time_dev
  HIP_TRY(launch_time_dev_fp64(c));
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(c->T);
  HIP_TRY(hipMemsetAsync(bits, 0, 4 * sizeof(unsigned long long), c->stream));
  HIP_TRY(launch_absmax(p, is_open(p) ? 1 : 0, c->U2, bits, c->stream));
  unsigned long long h[4];
  HIP_TRY(hipMemcpyAsync(h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  bool ok = true;
  for (int k = 0; k < 4; ++k) {
    double v;
    memcpy(&v, &h[k], sizeof v);
    if (out) out[k] = v;
    ok = ok && v <= limit;
  }
  *steady = ok ? 1 : 0;
  return DIM2D_OK;
}

// the three stages of one step; dt from `dt_fixed` or from the CFL reductions
static int run_step(dim2d_ctx* c, double dt_fixed, int cfl_mode) {
  const DevParams& p = c->dp;
  StageArgs a;
  memset(&a, 0, sizeof a);
  a.ctl = c->ctl;
  a.dt_fixed = dt_fixed;
  a.cfl_mode = cfl_mode;
  a.parity = (int)(c->steps & 1);
  const bool wall = has_wall(p);
  const bool open = is_open(p);
  // the ghost frame of open edges: its kernel reads only the stage input and
  // writes only frame cells of the stage output, so it runs on the side stream
  // beside the stage kernel (forked before it, joined after it)
  auto fork_frame = [&](int mixed, int stage) -> int {
    if (!open) return DIM2D_OK;
    HIP_TRY(hipEventRecord(c->fork_ev, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    HIP_TRY(launch_frame(p, mixed, stage, a, c->side));
    HIP_TRY(hipEventRecord(c->join_ev, c->side));
    return DIM2D_OK;
  };
  auto frame = [&](int, int) -> int {
    if (open) HIP_TRY(hipStreamWaitEvent(c->stream, c->join_ev, 0));
    return DIM2D_OK;
  };
  int st = DIM2D_OK;
  bool result_in_u = false;  // the new state is in U (else in U2, swapped below)
  if (c->ff_set) HIP_TRY(launch_save_frame(p, c->U, c->F0, c->stream));
  if (c->cg && !c->mixed) {
    // Cockburn-Gau stencils: FP64 RK3-TVD stages u^n -> A -> B -> U2, no walls
    a.U = c->U;
    double* in[3] = {c->U, c->A, c->B};
    double* out[3] = {c->A, c->B, c->U2};
    for (int s = 0; s < 3; ++s) {
      a.S64 = in[s];
      a.O64 = out[s];
      HIP_TRY(timed_launch(c, s, [&] { return launch_cg_stage(p, 0, 1 + s, a, c->stream); }));
    }
  } else if (c->cg) {
    // in mixed precision: u^n -> D1 -> D2 (FP32 increments) -> U2
    a.U = c->U;
    a.O32 = c->D1;
    HIP_TRY(timed_launch(c, 0, [&] { return launch_cg_stage(p, 1, 1, a, c->stream); }));
    a.S32 = c->D1;
    a.O32 = c->D2;
    HIP_TRY(timed_launch(c, 1, [&] { return launch_cg_stage(p, 1, 2, a, c->stream); }));
    a.S32 = c->D2;
    a.O32 = nullptr;
    a.O64 = c->U2;
    HIP_TRY(timed_launch(c, 2, [&] { return launch_cg_stage(p, 1, 3, a, c->stream); }));
  } else if (c->lsrk && c->fuse) {
    // low-storage stages 1 and 2 in one launch, u1 and q1 on chip: u0 (U) ->
    // u2 (U2) and q2 (Q, or Q32 in mixed precision); stage 3 (U2, Q) -> U, so
    // the new state is in U and the buffers are not swapped below.  The wall
    // root guess is each stage's input state, except for the ghosts of u2
    // (u1 is not stored): u0.  The states are FP64 in both modes, so are the
    // wall launches.
    a.U = c->U;
    a.Q = c->Q;
    a.Q32 = c->Q32;
    a.O64 = c->U2;
    HIP_TRY(timed_launch(c, 0, [&] { return launch_stage_pair(p, c->mixed, 1, a, c->pair_rows, c->stream); }));
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U2, nullptr, c->U2, nullptr, c->U, c->stream); }));
    a.S64 = c->U2;
    a.O64 = c->U;
    HIP_TRY(timed_launch(c, 2, [&] { return launch_stage(p, c->mixed, 6, a, c->block_x, c->rows, c->stream); }));
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U, nullptr, c->U, nullptr, c->U2, c->stream); }));
    result_in_u = true;
  } else if (c->lsrk) {
    // low-storage RK3: the state ping-pongs U -> U2 -> U -> U2, q stays in Q;
    // the wall root guess is each stage's input state
    double* in = c->U;
    double* out = c->U2;
    a.Q = c->Q;
    a.Q32 = c->Q32;
    for (int s = 0; s < 3; ++s) {
      a.S64 = in;
      a.O64 = out;
      if ((st = fork_frame(c->mixed, 4 + s))) return st;
      HIP_TRY(timed_launch(c, s, [&] { return launch_stage(p, c->mixed, 4 + s, a, c->block_x, c->rows, c->stream); }));
      if ((st = frame(c->mixed, 4 + s))) return st;
      if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, out, nullptr, out, nullptr, in, c->stream); }));
      std::swap(in, out);
    }
    // three swaps: the new state is in U2, as for RK3-TVD
  } else if (c->fuse) {
    // stages 1 and 2 in one launch: u^n -> stage-2 state (B, or its increment
    // D2 in mixed precision), the stage-1 state and its wall ghosts inside.
    // Open edges: the stage-1 frame (into A, or the increments D1) before it,
    // since the pair reads it, and the stage-2 frame after it, since that
    // reads the band of s1 beside the frame the pair stores there.
    a.U = c->U;
    if (open) {
      a.S64 = c->U;
      if (!c->mixed) a.O64 = c->A;
      else a.O32 = c->D1;
      HIP_TRY(timed_launch(c, 3, [&] { return launch_frame(p, c->mixed, 1, a, c->stream); }));
      a.F64 = c->mixed ? nullptr : c->A;
      a.F32 = c->mixed ? c->D1 : nullptr;
    }
    a.O64 = c->mixed ? nullptr : c->B;
    a.O32 = c->mixed ? c->D2 : nullptr;
    HIP_TRY(timed_launch(c, 0, [&] { return launch_stage_pair(p, c->mixed, 0, a, c->pair_rows, c->stream); }));
    if (open) {
      a.S64 = c->mixed ? nullptr : c->A;
      a.S32 = c->mixed ? c->D1 : nullptr;
      HIP_TRY(timed_launch(c, 3, [&] { return launch_frame(p, c->mixed, 2, a, c->stream); }));
    }
    if (!c->mixed) {
      if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->B, nullptr, c->B, nullptr, c->U, c->stream); }));
      a.O64 = c->U2;
      a.S64 = c->B;
      if ((st = fork_frame(0, 3))) return st;
      HIP_TRY(timed_launch(c, 2, [&] { return launch_stage(p, 0, 3, a, c->block_x, c->rows, c->stream); }));
      if ((st = frame(0, 3))) return st;
      if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U2, nullptr, c->U2, nullptr, c->U, c->stream); }));
    } else {
      if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 1, c->U, nullptr, c->D2, nullptr, c->D2, c->U, c->stream); }));
      a.S32 = c->D2;
      a.O32 = nullptr;
      a.O64 = c->U2;
      if ((st = fork_frame(1, 3))) return st;
      HIP_TRY(timed_launch(c, 2, [&] { return launch_stage(p, 1, 3, a, c->block_x, c->rows, c->stream); }));
      if ((st = frame(1, 3))) return st;
      if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U2, nullptr, c->U2, nullptr, c->U, c->stream); }));
    }
  } else if (!c->mixed) {
    a.U = c->U;
    a.S64 = c->U;
    a.O64 = c->A;
    if ((st = fork_frame(0, 1))) return st;
    HIP_TRY(timed_launch(c, 0, [&] { return launch_stage(p, 0, 1, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(0, 1))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->A, nullptr, c->A, nullptr, c->U, c->stream); }));
    a.S64 = c->A;
    a.O64 = c->B;
    if ((st = fork_frame(0, 2))) return st;
    HIP_TRY(timed_launch(c, 1, [&] { return launch_stage(p, 0, 2, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(0, 2))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->B, nullptr, c->B, nullptr, c->U, c->stream); }));
    a.S64 = c->B;
    a.O64 = c->U2;
    if ((st = fork_frame(0, 3))) return st;
    HIP_TRY(timed_launch(c, 2, [&] { return launch_stage(p, 0, 3, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(0, 3))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U2, nullptr, c->U2, nullptr, c->U, c->stream); }));
  } else {
    a.U = c->U;
    a.O32 = c->D1;
    if ((st = fork_frame(1, 1))) return st;
    HIP_TRY(timed_launch(c, 0, [&] { return launch_stage(p, 1, 1, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(1, 1))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 1, c->U, nullptr, c->D1, nullptr, c->D1, c->U, c->stream); }));
    a.S32 = c->D1;
    a.O32 = c->D2;
    if ((st = fork_frame(1, 2))) return st;
    HIP_TRY(timed_launch(c, 1, [&] { return launch_stage(p, 1, 2, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(1, 2))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 1, c->U, nullptr, c->D2, nullptr, c->D2, c->U, c->stream); }));
    a.S32 = c->D2;
    a.O32 = nullptr;
    a.O64 = c->U2;
    if ((st = fork_frame(1, 3))) return st;
    HIP_TRY(timed_launch(c, 2, [&] { return launch_stage(p, 1, 3, a, c->block_x, c->rows, c->stream); }));
    if ((st = frame(1, 3))) return st;
    if (wall) HIP_TRY(timed_launch(c, 3, [&] { return launch_wall(p, 0, nullptr, c->U2, nullptr, c->U2, nullptr, c->U, c->stream); }));
  }
  if (!result_in_u) std::swap(c->U, c->U2);
  c->steps++;
  c->red_valid = true;
  if (c->ff_set) c->f0_valid = true;
  return DIM2D_OK;
}

static int set_time(dim2d_ctx* c, double t) {
  HIP_TRY(hipMemcpyAsync(&c->ctl->t, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
  return DIM2D_OK;
}

int dim2d_rk3_step(dim2d_ctx* c, double t, double dt) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "step before upload");
  if (!(dt > 0) || !isfinite(dt)) return fail(DIM2D_EINVAL, "dt must be positive and finite");
  int st = set_time(c, t);
  if (st) return st;
  st = run_step(c, dt, 0);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_integrate(dim2d_ctx* c, double t0, double dt_fixed, double cfl, int32_t nsteps) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "integrate before upload");
  if (nsteps < 0) return fail(DIM2D_EINVAL, "nsteps must be >= 0");
  const int cfl_mode = cfl > 0;
  if (!cfl_mode && (!(dt_fixed > 0) || !isfinite(dt_fixed)))
    return fail(DIM2D_EINVAL, "dt must be positive and finite when cfl <= 0");
  int st = set_time(c, t0);
  if (st) return st;
  if (cfl_mode) {
    HIP_TRY(hipMemcpyAsync(&c->ctl->cfl, &cfl, sizeof(double), hipMemcpyHostToDevice, c->stream));
    // the first step's dt needs the reductions of u^n; the last stage of the
    // previous step left them in its slot unless the state was uploaded since
    if (!c->red_valid) {
      HIP_TRY(launch_reduce(c->dp, c->U, c->ctl, (int)(c->steps & 1), c->stream));
      c->red_valid = true;
    }
  }
  // steps replayed from captured graphs, 8 at a time, then 2 at a time: no
  // host launch gaps inside a graph, one (~8 us) between graphs.  Not while
  // profiling (event pairs around launches).  With open edges the frame
  // kernels' fork and join events take the side stream into the capture, so
  // the graph holds them as parallel branches.
  int s = 0;
  const bool graphs = !c->profiling;
  constexpr int kGraphSteps[2] = {2, 8};
  while (graphs) {
    const int gi = nsteps - s >= kGraphSteps[1] ? 1 : nsteps - s >= kGraphSteps[0] ? 0 : -1;
    if (gi < 0) break;
    const int len = kGraphSteps[gi];
    const int k = (int)(c->steps & 1);
    hipGraphExec_t& gx = c->gexec[gi][k];
    if (gx && (c->g_dt[gi][k] != dt_fixed || c->g_cfl[gi][k] != cfl_mode)) {
      (void)hipGraphExecDestroy(gx);
      gx = nullptr;
    }
    if (!gx) {
      // capture `len` steps; run_step's bookkeeping (buffer swap, step count)
      // is undone afterwards, the graph launch below makes the steps happen
      double* u = c->U;
      double* u2 = c->U2;
      const long long steps0 = c->steps;
      const bool rv0 = c->red_valid;
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int q = 0; q < len && !st; ++q) st = run_step(c, dt_fixed, cfl_mode);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      c->U = u;
      c->U2 = u2;
      c->steps = steps0;
      c->red_valid = rv0;
      if (st || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (st) return st;
        HIP_TRY(e);
      }
      const hipError_t ei = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ei);
      c->g_dt[gi][k] = dt_fixed;
      c->g_cfl[gi][k] = cfl_mode;
    }
    HIP_TRY(hipGraphLaunch(gx, c->stream));
    c->steps += len;  // an even number of buffer swaps: U and U2 are where they were
    c->red_valid = true;
    s += len;
  }
  for (; s < nsteps; ++s) {
    st = run_step(c, dt_fixed, cfl_mode);
    if (st) return st;
  }
  return DIM2D_OK;
}

int dim2d_get_params(dim2d_ctx* c, dim2d_params* out) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  *out = c->prm;
  return DIM2D_OK;
}

int dim2d_openbc_undermined(dim2d_ctx* c, int32_t distance, double md_threshold,
                            int32_t out[4]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "openbc_undermined before upload");
  if (!is_open(c->dp)) return fail(DIM2D_EINVAL, "the edges are not open (DIM2D_BC_OPEN)");
  if (distance < 1) return fail(DIM2D_EINVAL, "distance must be >= 1, got %d", distance);
  // the planar staging buffer is free between calls: four counters in it
  int* counts = reinterpret_cast<int*>(c->T);
  HIP_TRY(hipMemsetAsync(counts, 0, 4 * sizeof(int), c->stream));
  if (!(md_threshold < 0.0) && !(md_threshold <= 1.0))
    return fail(DIM2D_EINVAL, "md_threshold must be < 0 (off) or in [0, 1], got %g", md_threshold);
  HIP_TRY(launch_undermined(c->dp, distance, md_threshold, c->U, counts, c->stream));
  HIP_TRY(hipMemcpyAsync(out, counts, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_set_far_field(dim2d_ctx* c, const double far_field[4]) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!is_open(c->dp)) return fail(DIM2D_EINVAL, "the edges are not open (DIM2D_BC_OPEN)");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (!far_field) {
    (void)hipFree(c->F0);
    c->F0 = nullptr;
    c->ff_set = false;
    c->f0_valid = false;
    drop_graphs(c);  // captured steps hold the frame snapshot
    return DIM2D_OK;
  }
  for (int k = 0; k < 4; ++k)
    if (!isfinite(far_field[k])) return fail(DIM2D_EINVAL, "far_field[%d] is not finite", k);
  if (!(far_field[0] > 0.0 && far_field[0] < 3.0))
    return fail(DIM2D_EINVAL, "far-field density %g outside (0, 3)", far_field[0]);
  if (!c->F0) {
    const size_t n = (size_t)frame_cells(c->dp.nx, c->dp.ny) * 4;
    HIP_TRY(hipMalloc(&c->F0, n * sizeof(double)));
    c->f0_valid = false;
    drop_graphs(c);
  }
  for (int k = 0; k < 4; ++k) c->ff[k] = far_field[k];
  c->ff_set = true;
  return DIM2D_OK;
}

// The context on the grid grown by add_w .. add_n points per edge (negative:
// cropped), the state moved to its place: the caller's handle keeps pointing
// at `c`.  New points (add > 0) by the characteristic rule with a far field
// set, else the copy rule; a crop only copies.
static int resize_domain(dim2d_ctx* c, int add_w, int add_e, int add_s, int add_n) {
  const bool characteristic = c->ff_set && (add_w > 0 || add_e > 0 || add_s > 0 || add_n > 0);
  const long long nx = (long long)c->prm.nx + add_w + add_e;
  const long long ny = (long long)c->prm.ny + add_s + add_n;
  if (nx > kMaxExtent || ny > kMaxExtent)
    return fail(DIM2D_EINVAL, "extended grid %lldx%lld exceeds %d points per axis", nx, ny,
                kMaxExtent);
  if (nx == c->prm.nx && ny == c->prm.ny) return DIM2D_OK;
  // the new grid keeps dx and dy: x_map(bc+1) = x_min moves add_w points west
  // (field_abstract_class.f:286-315 spacing)
  dim2d_params q = c->prm;
  q.nx = (int32_t)nx;
  q.ny = (int32_t)ny;
  q.x_min = c->prm.x_min - add_w * c->dp.dx;
  q.x_max = c->prm.x_max + add_e * c->dp.dx;
  q.y_min = c->prm.y_min - add_s * c->dp.dy;
  q.y_max = c->prm.y_max + add_n * c->dp.dy;
  dim2d_ctx* n = nullptr;
  int st = dim2d_ctx_create(&q, &n);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(c->stream));
  // state and step control (t, dt, CFL number, reduction slots) move over
  hipError_t e = launch_extend(n->dp, c->dp, add_w, add_s, c->U, n->U, n->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(n->ctl, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToDevice, n->stream);
  const double h = std::min(n->dp.dx, n->dp.dy);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&n->ctl->h, &h, sizeof h, hipMemcpyHostToDevice, n->stream);
  // new points by the characteristic extrapolation from the last step's
  // u^n frame (t - dt) and the state (t), over the copy rule's values
  if (e == hipSuccess && characteristic) {
    const double x0 = q.x_min - kBcSize * n->dp.dx, y0 = q.y_min - kBcSize * n->dp.dy;
    e = launch_newgrdpt(n->dp, c->dp, add_w, add_s, x0, y0, c->ff, n->ctl, c->F0, n->U,
                        n->stream);
  }
  if (e == hipSuccess && c->ff_set)
    e = hipMalloc(&n->F0, (size_t)frame_cells(n->dp.nx, n->dp.ny) * 4 * sizeof(double));
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) {
    dim2d_ctx_destroy(n);
    return fail(DIM2D_EDEVICE, "domain extension failed: %s", hipGetErrorString(e));
  }
  if (c->ff_set) {
    for (int k = 0; k < 4; ++k) n->ff[k] = c->ff[k];
    n->ff_set = true;
    n->f0_valid = false;
  }
  n->grown[0] = c->grown[0] + add_w;
  n->grown[1] = c->grown[1] + add_e;
  n->grown[2] = c->grown[2] + add_s;
  n->grown[3] = c->grown[3] + add_n;
  if (c->block_x_set) {  // else the new grid keeps the default of its own size
    n->block_x = c->block_x;
    n->block_x_set = true;
  }
  if (c->rows_set) {
    n->rows = c->rows;
    n->rows_set = true;
  }
  if (c->pair_rows_set) {
    n->pair_rows = c->pair_rows;
    n->pair_rows_set = true;
  }
  n->fuse = c->fuse && fusion_applies(n);
  n->uploaded = true;
  n->steps = c->steps;
  n->profiling = c->profiling;
  n->profile_mask = c->profile_mask;
  // the caller's handle keeps pointing at `c`: swap the contents, then free
  // the old buffers (now held by `n`)
  std::swap(*c, *n);
  dim2d_ctx_destroy(n);
  return DIM2D_OK;
}

int dim2d_extend(dim2d_ctx* c, int32_t add_w, int32_t add_e, int32_t add_s, int32_t add_n) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "extend before upload");
  if (!is_open(c->dp)) return fail(DIM2D_EINVAL, "only open (DIM2D_BC_OPEN) edges are extended");
  if (add_w < 0 || add_e < 0 || add_s < 0 || add_n < 0)
    return fail(DIM2D_EINVAL, "grid points to add must be >= 0");
  if (c->ff_set && (add_w > 1 || add_e > 1 || add_s > 1 || add_n > 1))
    return fail(DIM2D_EINVAL,
                "with a far field the grid grows by at most one layer per edge and call");
  if (add_w + add_e + add_s + add_n == 0) return DIM2D_OK;
  if (c->ff_set && !c->f0_valid)
    return fail(DIM2D_ESTATE,
                "no step since the last upload or extension: the state at t - dt that new "
                "points are extrapolated from is missing");
  return resize_domain(c, add_w, add_e, add_s, add_n);
}

int dim2d_get_growth(dim2d_ctx* c, int32_t out[4]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  for (int k = 0; k < 4; ++k) out[k] = c->grown[k];
  return DIM2D_OK;
}

int dim2d_shrink(dim2d_ctx* c, int32_t del_w, int32_t del_e, int32_t del_s, int32_t del_n) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "shrink before upload");
  if (!is_open(c->dp)) return fail(DIM2D_EINVAL, "only open (DIM2D_BC_OPEN) edges are shrunk");
  const int32_t del[4] = {del_w, del_e, del_s, del_n};
  for (int k = 0; k < 4; ++k)
    if (del[k] < 0 || del[k] > c->grown[k])
      return fail(DIM2D_EINVAL,
                  "edge %d: %d points to remove, but %d were added since the context was "
                  "created", k, del[k], c->grown[k]);
  if (del_w + del_e + del_s + del_n == 0) return DIM2D_OK;
  return resize_domain(c, -del_w, -del_e, -del_s, -del_n);
}

// the check bands of the decrease: for each edge grown since creation, the
// `distance` rows (S, N) or columns (W, E) on each side of the edge's place
// at creation, along the created grid's extent and `distance` - 1 points
// beyond it (the ends of the S/N bands; W/E bands one more point south),
// clipped to the points whose 3x3 neighbourhood lies in the grid
// (get_check_line_param and check_layer_interior / check_layer_bf,
// bf_decrease_module.f:252-330, 448-600); an edge not grown gets an empty band
static Rects4 decrease_bands(const dim2d_ctx* c, int d) {
  const int nx = c->dp.nx, ny = c->dp.ny;
  const int gw = c->grown[0], ge = c->grown[1], gs = c->grown[2], gn = c->grown[3];
  // 0-based columns/rows of the created grid's W/S ghost cells next to the
  // interior (align_W, align_S) and of its E/N ones (align_E, align_N)
  const int aw = 1 + gw, ae = nx - 2 - ge, as = 1 + gs, an = ny - 2 - gn;
  Rects4 r;
  auto set = [&](int k, int i0, int i1, int j0, int j1) {  // inclusive bounds
    i0 = std::max(i0, 1);
    i1 = std::min(i1, nx - 2);
    j0 = std::max(j0, 1);
    j1 = std::min(j1, ny - 2);
    const bool on = c->grown[k] > 0 && i0 <= i1 && j0 <= j1;
    r.i0[k] = on ? i0 : 0;
    r.i1[k] = on ? i1 + 1 : 0;
    r.j0[k] = on ? j0 : 0;
    r.j1[k] = on ? j1 + 1 : 0;
  };
  set(0, aw - d + 1, aw + d, as - d, an + d - 1);  // W
  set(1, ae - d, ae + d - 1, as - d, an + d - 1);  // E
  set(2, aw - d + 1, ae + d - 1, as - d + 1, as + d);  // S
  set(3, aw - d + 1, ae + d - 1, an - d, an + d - 1);  // N
  return r;
}

int dim2d_openbc_undermined_decrease(dim2d_ctx* c, int32_t distance, double md_threshold,
                                     int32_t out[4]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "openbc_undermined_decrease before upload");
  if (!is_open(c->dp)) return fail(DIM2D_EINVAL, "the edges are not open (DIM2D_BC_OPEN)");
  if (distance < 1) return fail(DIM2D_EINVAL, "distance must be >= 1, got %d", distance);
  if (!(md_threshold < 0.0) && !(md_threshold <= 1.0))
    return fail(DIM2D_EINVAL, "md_threshold must be < 0 (off) or in [0, 1], got %g", md_threshold);
  int* counts = reinterpret_cast<int*>(c->T);
  HIP_TRY(hipMemsetAsync(counts, 0, 4 * sizeof(int), c->stream));
  HIP_TRY(launch_undermined_rects(c->dp, decrease_bands(c, distance), md_threshold, c->U, counts,
                                  c->stream));
  HIP_TRY(hipMemcpyAsync(out, counts, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_adapt_domain(dim2d_ctx* c, int32_t adapt_mask, int32_t distance, double md_threshold,
                       int32_t grow, int32_t out[4]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  if (adapt_mask < 0 || adapt_mask > 255)
    return fail(DIM2D_EINVAL, "adapt_mask out of range: %d", adapt_mask);
  if (grow < 1) return fail(DIM2D_EINVAL, "grow must be >= 1, got %d", grow);
  if (c->ff_set && grow != 1)
    return fail(DIM2D_EINVAL, "with a far field the grid grows by one layer per call, got %d",
                grow);
  int32_t cnt[4], del[4] = {0, 0, 0, 0};
  int st;
  if (adapt_mask >> 4) {
    // bf_interface_dcr%adapt_domain_extension: the decrease first
    // (remove_inactivated_bf_layers, bf_interface_dcr_class.f:101-216)
    st = dim2d_openbc_undermined_decrease(c, distance, md_threshold, cnt);
    if (st) return st;
    for (int k = 0; k < 4; ++k)
      del[k] = ((adapt_mask >> (4 + k)) & 1) && c->grown[k] > 0 && cnt[k] == 0 ? c->grown[k] : 0;
    st = dim2d_shrink(c, del[0], del[1], del[2], del[3]);
    if (st) return st;
  }
  // then the increase, on the grid as it now is
  st = dim2d_openbc_undermined(c, distance, md_threshold, cnt);
  if (st) return st;
  int32_t add[4];
  for (int k = 0; k < 4; ++k) add[k] = (adapt_mask >> k) & 1 && cnt[k] > 0 ? grow : 0;
  if (c->ff_set && del[0] + del[1] + del[2] + del[3] > 0) {
    // a crop in this call: the extrapolation's state at t - dt is gone, the
    // increase waits for the next call
    for (int k = 0; k < 4; ++k) add[k] = 0;
  }
  st = dim2d_extend(c, add[0], add[1], add[2], add[3]);
  for (int k = 0; k < 4; ++k) out[k] = add[k] - del[k];
  return st;
}

int dim2d_synchronize(dim2d_ctx* c) {
  if (!c) return fail(DIM2D_EINVAL, "ctx is NULL");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DIM2D_OK;
}

int dim2d_time(dim2d_ctx* c, double out[2]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  StepCtl h;
  HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  out[0] = h.t;
  out[1] = h.dt;
  return DIM2D_OK;
}

static int reduce_now(dim2d_ctx* c, StepCtl* h) {
  const int slot = (int)((c->steps + 1) & 1);  // the slot the next step does not read
  HIP_TRY(launch_reduce(c->dp, c->U, c->ctl, slot, c->stream));
  HIP_TRY(hipMemcpyAsync(h, c->ctl, sizeof *h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return slot;
}

int dim2d_cfl_dt(dim2d_ctx* c, double cfl, double out[4]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "cfl_dt before upload");
  StepCtl h;
  const int slot = reduce_now(c, &h);
  if (slot < 0) return slot;
  double amax, rmin, rmax;
  memcpy(&amax, &h.amax[slot], 8);
  memcpy(&rmin, &h.rmin[slot], 8);
  memcpy(&rmax, &h.rmax[slot], 8);
  const double hh = std::min(c->dp.dx, c->dp.dy);
  double dt = hh / amax;
  dt = std::min(dt, c->prm.re * rmin * hh * hh);
  dt = std::min(dt, sqrt(c->prm.we / rmax) * hh * hh);
  out[0] = cfl * dt;
  out[1] = amax;
  out[2] = rmin;
  out[3] = rmax;
  if (h.nonfinite[slot]) return fail(DIM2D_ENONFINITE, "%u non-finite cells", h.nonfinite[slot]);
  return DIM2D_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// allocation-free time derivative
// ------------------------------------------------------------------------
// Device scratch of dim2d_time_dev_noalloc: the interleaved state and its
// time derivative (two fields of pitch x ny cells, 32 B per cell, each
// starting kShift64 doubles past a 128-byte line like the contexts' fields).
// Every GPU the library runs on holds one copy for the life of the process.
// In host code `scratch` names a host-side stand-in, never device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// the current device's copy of `scratch`, looked up on the first call made
// on that device
int scratch_for_current_device(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(DIM2D_EINVAL, "device %d beyond the %d devices the scratch table holds", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(DIM2D_EDEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return DIM2D_OK;
}

}  // namespace

extern "C" {

int dim2d_time_dev_noalloc(const dim2d_params* params, const double* d_nodes, double* d_time_dev,
                           void* stream) {
  int st = validate(params);
  if (st) return st;
  if (!d_nodes || !d_time_dev) return fail(DIM2D_EINVAL, "NULL device array");
  const DevParams p = make_dev_params(*params);
  const size_t n = (size_t)p.ncell * 4;  // doubles per interleaved field
  const size_t field = (n + kShift64 + 15) / 16 * 16;  // whole 128-byte lines
  if (2 * field > kScratchDoubles)
    return fail(DIM2D_EINVAL,
                "grid %dx%d needs %zu bytes of scratch, more than the %zu bytes "
                "dim2d_time_dev_noalloc holds (at most about %zu cells of pitch x ny)",
                params->nx, params->ny, 2 * field * sizeof(double), kScratchBytes,
                kScratchDoubles / 8);
  double* base = nullptr;
  if ((st = scratch_for_current_device(&base))) return st;
  double* U = base + kShift64;
  double* D = base + field + kShift64;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // every cell of U is written by the pack, every interior cell of D by the
  // stage kernel (the frame too with open edges), every cell of the output
  // by the unpack: nothing is read that this call did not write
  HIP_TRY(launch_pack(p, d_nodes, U, s));
  StageArgs a;
  memset(&a, 0, sizeof a);
  a.U = U;
  a.S64 = U;
  a.O64 = D;
  const bool open = is_open(p);
  if (params->operators == DIM2D_OPS_COCKBURN_GAU) HIP_TRY(launch_cg_stage(p, 0, 0, a, s));
  else HIP_TRY(launch_stage(p, 0, 0, a, default_block_x(p), default_rows(p), s));
  if (open) HIP_TRY(launch_frame(p, 0, 0, a, s));
  HIP_TRY(launch_unpack(p, D, d_time_dev, open ? 0 : 1, kDebugReal, s));
  return DIM2D_OK;
}

int dim2d_diagnostics(dim2d_ctx* c, double out[7]) {
  if (!c || !out) return fail(DIM2D_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(DIM2D_ESTATE, "diagnostics before upload");
  StepCtl h;
  const int slot = reduce_now(c, &h);
  if (slot < 0) return slot;
  for (int k = 0; k < 4; ++k) out[k] = h.sum[slot][k];
  memcpy(&out[4], &h.rmin[slot], 8);
  memcpy(&out[5], &h.rmax[slot], 8);
  out[6] = (double)h.nonfinite[slot];
  return DIM2D_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// batches
// ------------------------------------------------------------------------
struct dim2d_batch {
  int n = 0;
  int mixed = 0;
  int device = 0;
  std::vector<dim2d_params> prm;
  std::vector<DevParams> hp;
  DevParams* dp = nullptr;
  size_t total = 0;  // elements per variable-set buffer (sum of 4*ncell)
  double *U = nullptr, *U2 = nullptr, *A = nullptr, *B = nullptr;
  double* stage = nullptr;  // planar staging of one simulation (uploads, downloads)
  float *D1 = nullptr, *D2 = nullptr;
  double* d_dts = nullptr;
  std::vector<double> h_dts;
  WorkItem* d_items = nullptr;
  int nitems = 0;
  size_t stage_n = 0;
  int block_x = 64, rows = 32, wg = 0;
  int wg_eff = 1;  // persistent workgroups actually launched
  // stages 1 and 2 fused (stage pair), on when every simulation supports it:
  // its own work list of 56-column strips and pair_rows-row tiles
  bool fuse = false;
  int pair_rows = 32, pair_wg = 0, pair_wg_eff = 1;
  WorkItem* d_pitems = nullptr;
  int npitems = 0;
  // pipelined (with the pair): the simulations form two groups of about equal
  // cell counts, group B half a step behind group A, and each step is two
  // mixed launches (stage_mix_batch_kernel): A's stage 3 with B's stage
  // pair, then A's next stage pair with B's stage 3
  bool pipe = false;
  int mix_wg = 0, mix_wg_eff = 1;
  std::vector<int> group;                        // 0 (A) or 1 (B) per simulation
  std::vector<DevParams> hpg[2];                 // the groups' parameters (wall launches)
  DevParams* dpg[2] = {nullptr, nullptr};
  WorkItem *d_k1 = nullptr, *d_k2 = nullptr;     // A pairs + B stage 3; A stage 3 + B pairs
  WorkItem *d_pa = nullptr, *d_sb = nullptr;     // A pairs alone; B stage 3 alone
  int n_k1 = 0, n_k2 = 0, n_pa = 0, n_sb = 0;
  bool any_wall = false;
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  long long steps = 0;
};

// the tiles of every simulation (strips of `ow` output columns, at most
// `rows` rows), largest simulations first so the persistent grid drains
// evenly.  A simulation's tiles share one height, the smallest that keeps
// their number (a 64-row grid in 60-row tiles takes two 32-row tiles, not
// 60 + 4 rows: every tile pays the pipeline's warm-up and halo rows).
static std::vector<WorkItem> batch_items(const dim2d_batch* b, int ow, int rows,
                                         int group = -1, int kind = kItemPair,
                                         int strips_per_item = 1) {
  std::vector<WorkItem> items;
  for (int s = 0; s < b->n; ++s) {
    if (group >= 0 && b->group[s] != group) continue;
    const DevParams& p = b->hp[s];
    const int nj = p.ny - 2 * kBcSize;
    const int nstrips = (p.nx - 2 * kBcSize + ow - 1) / ow;
    const int nbx = (nstrips + strips_per_item - 1) / strips_per_item;
    const int nby = (nj + rows - 1) / rows;
    const int h = (nj + nby - 1) / nby;
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) items.push_back(WorkItem{s, bx, by, h, kind, 0, 0, 0});
  }
  std::stable_sort(items.begin(), items.end(), [&](const WorkItem& x, const WorkItem& y) {
    return b->hp[x.sim].nx * (long long)b->hp[x.sim].ny > b->hp[y.sim].nx * (long long)b->hp[y.sim].ny;
  });
  return items;
}

// two work lists merged so that every stretch of the result holds both in
// their proportion (the launch's resident workgroups then mix the two kinds)
static std::vector<WorkItem> interleave(const std::vector<WorkItem>& x,
                                        const std::vector<WorkItem>& y) {
  std::vector<WorkItem> out;
  out.reserve(x.size() + y.size());
  size_t i = 0, j = 0;
  while (i < x.size() || j < y.size()) {
    // take from x while its share so far is below its share of the total
    if (j >= y.size() || (i < x.size() && i * y.size() <= j * x.size())) out.push_back(x[i++]);
    else out.push_back(y[j++]);
  }
  return out;
}

static int upload_items(const std::vector<WorkItem>& items, WorkItem** d, int* n) {
  (void)hipFree(*d);
  *d = nullptr;
  *n = (int)items.size();
  HIP_TRY(hipMalloc(d, items.size() * sizeof(WorkItem)));
  HIP_TRY(hipMemcpy(*d, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  return DIM2D_OK;
}

// The two groups of a pipelined batch: simulations by decreasing cell count,
// each to the group with fewer cells so far; their parameters for the wall
// launches of each group
static int batch_groups(dim2d_batch* b) {
  std::vector<int> order(b->n);
  for (int i = 0; i < b->n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b->hp[x].ncell > b->hp[y].ncell;
  });
  b->group.assign(b->n, 0);
  long long cells[2] = {0, 0};
  for (int s : order) {
    const int g = cells[1] < cells[0] ? 1 : 0;
    b->group[s] = g;
    cells[g] += b->hp[s].ncell;
  }
  for (int g = 0; g < 2; ++g) {
    b->hpg[g].clear();
    for (int s = 0; s < b->n; ++s)
      if (b->group[s] == g) b->hpg[g].push_back(b->hp[s]);
    (void)hipFree(b->dpg[g]);
    b->dpg[g] = nullptr;
    if (b->hpg[g].empty()) continue;
    HIP_TRY(hipMalloc(&b->dpg[g], b->hpg[g].size() * sizeof(DevParams)));
    HIP_TRY(hipMemcpy(b->dpg[g], b->hpg[g].data(), b->hpg[g].size() * sizeof(DevParams),
                      hipMemcpyHostToDevice));
  }
  return DIM2D_OK;
}

static int batch_rebuild_items(dim2d_batch* b) {
  int st = upload_items(batch_items(b, b->block_x - 2 * kBcSize, b->rows), &b->d_items, &b->nitems);
  if (st) return st;
  b->wg_eff = b->wg > 0 ? std::min(b->wg, b->nitems)
                       : std::min(b->nitems, max_resident_workgroups(b->mixed));
  if (b->wg_eff < 1) b->wg_eff = 1;
  if (b->fuse) {
    st = upload_items(batch_items(b, kPairOutCols, b->pair_rows), &b->d_pitems, &b->npitems);
    if (st) return st;
    b->pair_wg_eff = b->pair_wg > 0 ? std::min(b->pair_wg, b->npitems)
                                    : std::min(b->npitems, max_resident_pair_workgroups(b->mixed));
    if (b->pair_wg_eff < 1) b->pair_wg_eff = 1;
  }
  if (b->pipe) {
    const int ow3 = b->block_x - 2 * kBcSize;
    const auto pa = batch_items(b, kPairOutCols, b->pair_rows, 0, kItemPair);
    const auto pb = batch_items(b, kPairOutCols, b->pair_rows, 1, kItemPair);
    const auto sa = batch_items(b, ow3, b->rows, 0, kItemStage3, 2);
    const auto sb = batch_items(b, ow3, b->rows, 1, kItemStage3, 2);
    if ((st = upload_items(interleave(pa, sb), &b->d_k1, &b->n_k1))) return st;
    if ((st = upload_items(interleave(sa, pb), &b->d_k2, &b->n_k2))) return st;
    if ((st = upload_items(pa, &b->d_pa, &b->n_pa))) return st;
    if ((st = upload_items(sb, &b->d_sb, &b->n_sb))) return st;
    const int most = std::max(std::max(b->n_k1, b->n_k2), 1);
    b->mix_wg_eff = b->mix_wg > 0 ? std::min(b->mix_wg, most)
                                  : std::min(most, max_resident_mix_workgroups(b->mixed));
    if (b->mix_wg_eff < 1) b->mix_wg_eff = 1;
  }
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return DIM2D_OK;
}

// the wall ghost cells of one group after its stage pair (the stage-2 state,
// root guess u^n) or after its stage 3 (u^n+1 in out, root guess u^n)
static int pipe_walls(dim2d_batch* b, int g, bool after_pair, double* un, double* out) {
  if (!b->any_wall || b->hpg[g].empty()) return DIM2D_OK;
  const int n = (int)b->hpg[g].size();
  if (after_pair && b->mixed)
    HIP_TRY(launch_wall_batch(b->hpg[g].data(), b->dpg[g], n, 1, un, nullptr, b->D2, nullptr,
                              b->D2, un, b->stream));
  else if (after_pair)
    HIP_TRY(launch_wall_batch(b->hpg[g].data(), b->dpg[g], n, 0, nullptr, b->B, nullptr, b->B,
                              nullptr, un, b->stream));
  else
    HIP_TRY(launch_wall_batch(b->hpg[g].data(), b->dpg[g], n, 0, nullptr, out, nullptr, out,
                              nullptr, un, b->stream));
  return DIM2D_OK;
}

// One mixed launch of a pipelined batch (and its wall launches): the pairs
// of group gp read u^n from xp, the stage 3 of the other group reads u^n from
// xs and writes u^n+1 to ys; which lists run: 1 (A pairs + B stage 3), 2
// (A stage 3 + B pairs), 3 (A pairs alone), 4 (B stage 3 alone)
static int pipe_launch(dim2d_batch* b, int list, double* xp, double* xs, double* ys) {
  MixArgs m;
  memset(&m, 0, sizeof m);
  m.pU = xp;
  m.pO64 = b->mixed ? nullptr : b->B;
  m.pO32 = b->mixed ? b->D2 : nullptr;
  m.sU = xs;
  m.sS64 = b->mixed ? nullptr : b->B;
  m.sS32 = b->mixed ? b->D2 : nullptr;
  m.sO64 = ys;
  const WorkItem* d = list == 1 ? b->d_k1 : list == 2 ? b->d_k2 : list == 3 ? b->d_pa : b->d_sb;
  const int n = list == 1 ? b->n_k1 : list == 2 ? b->n_k2 : list == 3 ? b->n_pa : b->n_sb;
  if (n > 0)
    HIP_TRY(launch_stage_mix_batch(b->dp, d, n, b->mixed, m, b->d_dts, std::min(b->mix_wg_eff, n),
                                   b->stream));
  int st;
  const bool a_pairs = list == 1 || list == 3, a_stage3 = list == 2;
  const bool b_pairs = list == 2, b_stage3 = list == 1 || list == 4;
  if (a_pairs && (st = pipe_walls(b, 0, true, xp, nullptr))) return st;
  if (a_stage3 && (st = pipe_walls(b, 0, false, xs, ys))) return st;
  if (b_pairs && (st = pipe_walls(b, 1, true, xp, nullptr))) return st;
  if (b_stage3 && (st = pipe_walls(b, 1, false, xs, ys))) return st;
  return DIM2D_OK;
}

// One steady step of a pipelined batch from u^n in x to u^n+1 in y: on
// entry group A has run its step's stage pair (its stage-2 state waits in
// B/D2) and group B has u^n in x; on exit the same holds one step later.
// A's stage 3 goes with B's pair (x -> y), then A's next pair (reading y)
// with B's stage 3 (x -> y).
static int pipe_step(dim2d_batch* b, double* x, double* y) {
  int st = pipe_launch(b, 2, x, x, y);
  if (!st) st = pipe_launch(b, 1, y, x, y);
  return st;
}

// launches of one step: U -> U2 (pointers as given)
static int batch_step_launches(dim2d_batch* b, double* U, double* U2) {
  StageArgs a;
  memset(&a, 0, sizeof a);
  a.dts = b->d_dts;
  const int wg = b->wg_eff;
  hipStream_t s = b->stream;
  if (b->fuse) {
    // stages 1 and 2 in one launch (the wall ghost cells of the stage-1 state
    // inside it), then the stage-2 state's wall ghosts, stage 3 and its walls,
    // as a single context's fused step
    a.U = U;
    if (!b->mixed) a.O64 = b->B;
    else a.O32 = b->D2;
    HIP_TRY(launch_stage_pair_batch(b->dp, b->d_pitems, b->npitems, b->mixed, a, b->pair_wg_eff,
                                    s));
    if (!b->mixed) {
      if (b->any_wall)
        HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, b->B, nullptr, b->B, nullptr, U, s));
      a.S64 = b->B;
      a.O64 = U2;
    } else {
      if (b->any_wall)
        HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 1, U, nullptr, b->D2, nullptr, b->D2, U, s));
      a.S32 = b->D2;
      a.O32 = nullptr;
      a.O64 = U2;
    }
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, b->mixed, 3, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, U2, nullptr, U2, nullptr, U, s));
    return DIM2D_OK;
  }
  if (!b->mixed) {
    a.U = U;
    a.S64 = U;
    a.O64 = b->A;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 0, 1, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, b->A, nullptr, b->A, nullptr, U, s));
    a.S64 = b->A;
    a.O64 = b->B;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 0, 2, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, b->B, nullptr, b->B, nullptr, U, s));
    a.S64 = b->B;
    a.O64 = U2;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 0, 3, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, U2, nullptr, U2, nullptr, U, s));
  } else {
    a.U = U;
    a.O32 = b->D1;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 1, 1, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 1, U, nullptr, b->D1, nullptr, b->D1, U, s));
    a.S32 = b->D1;
    a.O32 = b->D2;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 1, 2, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 1, U, nullptr, b->D2, nullptr, b->D2, U, s));
    a.S32 = b->D2;
    a.O32 = nullptr;
    a.O64 = U2;
    HIP_TRY(launch_stage_batch(b->dp, b->d_items, b->nitems, 1, 3, a, wg, s));
    if (b->any_wall)
      HIP_TRY(launch_wall_batch(b->hp.data(), b->dp, b->n, 0, nullptr, U2, nullptr, U2, nullptr, U, s));
  }
  return DIM2D_OK;
}

extern "C" {

int dim2d_batch_create(const dim2d_params* params, int32_t n, dim2d_batch** out) {
  if (!out || !params || n <= 0) return fail(DIM2D_EINVAL, "invalid batch arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    int st = validate(&params[i]);
    if (st) return fail(st, "simulation %d: %s", i, g_err.c_str());
    if (params[i].precision != params[0].precision)
      return fail(DIM2D_EINVAL, "all simulations of a batch must share the precision mode");
    if (params[i].integrator != DIM2D_INT_RK3TVD)
      return fail(DIM2D_EINVAL, "batches run the RK3-TVD integrator only");
    if (params[i].operators != DIM2D_OPS_MATTSSON)
      return fail(DIM2D_EINVAL, "batches run the Mattsson operators only");
    if (params[i].bc_n == DIM2D_BC_OPEN)
      return fail(DIM2D_EINVAL, "batches do not support open boundaries");
  }
  dim2d_batch* b = new (std::nothrow) dim2d_batch();
  if (!b) return fail(DIM2D_ENOMEM, "host allocation failed");
  b->n = n;
  b->mixed = params[0].precision == DIM2D_PREC_MIXED;
  b->prm.assign(params, params + n);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    DevParams p = make_dev_params(params[i]);
    p.offset = (long long)off;
    off += (size_t)p.ncell * 4;
    off = (off + 63) / 64 * 64;
    b->stage_n = std::max(b->stage_n, (size_t)p.nx * p.ny * 4);
    b->any_wall |= has_wall(p);
    b->hp.push_back(p);
  }
  b->fuse = true;
  for (const DevParams& p : b->hp) b->fuse = b->fuse && pair_supported(p);
  // pipelined by default in FP64: its mixed launches keep the FP64 pair's 2
  // waves per SIMD (2-3.5 % faster a step over 256 simulations), while in mixed
  // precision they would drop the pair from 3 waves to 2 (9 % slower)
  b->pipe = b->fuse && !b->mixed && n >= 2;
  b->total = off;
  b->h_dts.assign(n, 0.0);
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&b->dp, n * sizeof(DevParams));
  if (e == hipSuccess)
    e = hipMemcpy(b->dp, b->hp.data(), n * sizeof(DevParams), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = field_alloc(&b->U, off);
  if (e == hipSuccess) e = field_alloc(&b->U2, off);
  if (e == hipSuccess && !b->mixed && !b->fuse) e = field_alloc(&b->A, off);
  if (e == hipSuccess && !b->mixed) e = field_alloc(&b->B, off);
  if (e == hipSuccess && b->mixed) e = field_alloc(&b->D1, off);
  if (e == hipSuccess && b->mixed) e = field_alloc(&b->D2, off);
  if (e == hipSuccess) e = hipMalloc(&b->d_dts, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->stage, b->stage_n * sizeof(double));
  if (e == hipSuccess) e = hipMemset(b->U, 0, off * sizeof(double));
  if (e == hipSuccess) e = hipMemset(b->U2, 0, off * sizeof(double));
  if (e == hipSuccess && b->A) e = hipMemset(b->A, 0, off * sizeof(double));
  if (e == hipSuccess && b->B) e = hipMemset(b->B, 0, off * sizeof(double));
  if (e == hipSuccess && b->D1) e = hipMemset(b->D1, 0, off * sizeof(float));
  if (e == hipSuccess && b->D2) e = hipMemset(b->D2, 0, off * sizeof(float));
  if (e != hipSuccess) {
    dim2d_batch_destroy(b);
    return fail(e == hipErrorOutOfMemory ? DIM2D_ENOMEM : DIM2D_EDEVICE, "batch allocation: %s",
                hipGetErrorString(e));
  }
  int st = batch_groups(b);
  if (!st) st = batch_rebuild_items(b);
  if (st) {
    dim2d_batch_destroy(b);
    return st;
  }
  *out = b;
  return DIM2D_OK;
}

int dim2d_batch_destroy(dim2d_batch* b) {
  if (!b) return DIM2D_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->dp);
  field_free(b->U);
  field_free(b->U2);
  field_free(b->A);
  field_free(b->B);
  field_free(b->D1);
  field_free(b->D2);
  (void)hipFree(b->d_dts);
  (void)hipFree(b->stage);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_pitems);
  (void)hipFree(b->d_k1);
  (void)hipFree(b->d_k2);
  (void)hipFree(b->d_pa);
  (void)hipFree(b->d_sb);
  (void)hipFree(b->dpg[0]);
  (void)hipFree(b->dpg[1]);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return DIM2D_OK;
}

int dim2d_batch_set_tiles(dim2d_batch* b, int32_t block_x, int32_t rows, int32_t workgroups) {
  if (!b) return fail(DIM2D_EINVAL, "batch is NULL");
  if (block_x != 0 && block_x != 64)
    return fail(DIM2D_EINVAL, "block_x must be 64 (one wave per column strip), got %d", block_x);
  if (rows < 0 || rows > 4096 || workgroups < 0)
    return fail(DIM2D_EINVAL, "invalid tile arguments");
  if (block_x) b->block_x = block_x;
  if (rows) b->rows = rows;
  b->wg = workgroups;
  return batch_rebuild_items(b);
}

int dim2d_batch_set_stage_fusion(dim2d_batch* b, int32_t fuse, int32_t rows_per_block,
                                 int32_t workgroups) {
  if (!b) return fail(DIM2D_EINVAL, "batch is NULL");
  if (rows_per_block < 0 || rows_per_block > 4096 || workgroups < 0 || fuse < 0 || fuse > 2)
    return fail(DIM2D_EINVAL, "invalid stage fusion arguments");
  if (fuse) {
    for (int i = 0; i < b->n; ++i)
      if (!pair_supported(b->hp[i]))
        return fail(DIM2D_EINVAL,
                    "simulation %d: stage fusion needs reflection, periodic or S/N wall edges "
                    "and at least 8 interior points per axis", i);
  } else if (!b->mixed && !b->A) {
    // the separate FP64 stages need the stage-1 buffer, which a fused batch
    // does not allocate
    hipError_t e = field_alloc(&b->A, b->total);
    if (e == hipSuccess) e = hipMemset(b->A, 0, b->total * sizeof(double));
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? DIM2D_ENOMEM : DIM2D_EDEVICE, "batch allocation: %s",
                  hipGetErrorString(e));
  }
  b->fuse = fuse != 0;
  b->pipe = fuse == 2;
  if (rows_per_block) b->pair_rows = rows_per_block;
  if (b->pipe) b->mix_wg = workgroups;
  else b->pair_wg = workgroups;
  return batch_rebuild_items(b);
}

int dim2d_batch_get_tiles(dim2d_batch* b, int32_t out[6]) {
  if (!b || !out) return fail(DIM2D_EINVAL, "NULL argument");
  out[0] = b->pipe ? 2 : b->fuse ? 1 : 0;
  out[1] = b->pair_rows;
  out[2] = b->pipe ? b->mix_wg_eff : b->fuse ? b->pair_wg_eff : 0;
  out[3] = b->block_x;
  out[4] = b->rows;
  out[5] = b->wg_eff;
  return DIM2D_OK;
}

int dim2d_batch_upload(dim2d_batch* b, int32_t sim, const double* nodes) {
  if (!b || !nodes || sim < 0 || sim >= b->n) return fail(DIM2D_EINVAL, "invalid argument");
  const DevParams& p = b->hp[sim];
  HIP_TRY(hipMemcpyAsync(b->stage, nodes, (size_t)p.nx * p.ny * 4 * sizeof(double),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(launch_pack(p, b->stage, b->U + p.offset, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return DIM2D_OK;
}

int dim2d_batch_download(dim2d_batch* b, int32_t sim, double* nodes) {
  if (!b || !nodes || sim < 0 || sim >= b->n) return fail(DIM2D_EINVAL, "invalid argument");
  const DevParams& p = b->hp[sim];
  double* base = b->U + p.offset;
  HIP_TRY(launch_fill_ghosts(p, base, b->stream));
  HIP_TRY(launch_unpack(p, base, b->stage, 0, 0.0, b->stream));
  HIP_TRY(hipMemcpyAsync(nodes, b->stage, (size_t)p.nx * p.ny * 4 * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return DIM2D_OK;
}

// nsteps of a pipelined batch: A's first pair alone, nsteps - 1 steady steps
// (replayed two at a time from a captured graph), then A's last stage 3 with
// B's last pair, and B's last stage 3 alone
static int pipe_integrate(dim2d_batch* b, int nsteps) {
  if (nsteps == 0) return DIM2D_OK;
  double* x = b->U;
  double* y = b->U2;
  int st = pipe_launch(b, 3, x, nullptr, nullptr);
  if (st) return st;
  int s = 0;
  if (nsteps - 1 >= 2 && !b->exec) {
    // the graph of two steady steps (x -> y -> x) is valid while the state
    // starts in U, which every replay keeps
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    st = pipe_step(b, b->U, b->U2);
    if (!st) st = pipe_step(b, b->U2, b->U);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HIP_TRY(e);
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  for (; s + 2 <= nsteps - 1; s += 2) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  if (s < nsteps - 1) {
    if ((st = pipe_step(b, x, y))) return st;
    std::swap(x, y);
    ++s;
  }
  // the last step: A's stage 3 with B's pair, then B's stage 3 alone
  if ((st = pipe_launch(b, 2, x, x, y))) return st;
  if ((st = pipe_launch(b, 4, nullptr, x, y))) return st;
  if (nsteps & 1) {
    std::swap(b->U, b->U2);
    // the captured graph refers to the old buffer order
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph) {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  return DIM2D_OK;
}

int dim2d_batch_integrate(dim2d_batch* b, const double* dts, int32_t nsteps) {
  if (!b || !dts || nsteps < 0) return fail(DIM2D_EINVAL, "invalid argument");
  for (int i = 0; i < b->n; ++i)
    if (!(dts[i] > 0) || !isfinite(dts[i])) return fail(DIM2D_EINVAL, "dt of simulation %d", i);
  HIP_TRY(hipMemcpyAsync(b->d_dts, dts, b->n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->pipe) {
    const int st = pipe_integrate(b, nsteps);
    if (st) return st;
    b->steps += nsteps;
    return DIM2D_OK;
  }
  if (!b->exec) {
    // capture two steps (U -> U2 -> U) so a replay leaves the state in U
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = batch_step_launches(b, b->U, b->U2);
    if (!st) st = batch_step_launches(b, b->U2, b->U);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HIP_TRY(e);
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  for (int s = 0; s + 1 < nsteps; s += 2) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  if (nsteps & 1) {
    int st = batch_step_launches(b, b->U, b->U2);
    if (st) return st;
    std::swap(b->U, b->U2);
    // the captured graph refers to the old buffer order
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph) {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  b->steps += nsteps;
  return DIM2D_OK;
}

int dim2d_batch_autotune(dim2d_batch* b, const double* dts, int32_t nsteps, double out[3]) {
  if (!b || !dts || !out || nsteps < 2) return fail(DIM2D_EINVAL, "invalid argument");
  const int resident = max_resident_workgroups(b->mixed);
  const int rows_c[] = {8, 12, 16, 24, 32};
  const int wg_c[] = {resident, 2 * resident};
  // a pipelined batch tunes the workgroups of its mixed launches in place of
  // the pair's
  const int president = b->pipe ? max_resident_mix_workgroups(b->mixed)
                                : max_resident_pair_workgroups(b->mixed);
  int& pair_wg = b->pipe ? b->mix_wg : b->pair_wg;
  const int prows_c[] = {16, 24, 32, 48, 60, 96, 128};
  const int pwg_c[] = {president, 2 * president};
  const int steps = nsteps & ~1;  // whole graph replays
  // keep the state: candidates run real steps
  double* keep = nullptr;
  HIP_TRY(hipMalloc(&keep, b->total * sizeof(double)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = DIM2D_OK;
  double best = INFINITY;
  bool kept = false;
  // ms per step of the current tiles
  auto time_tiles = [&](double* ms_out) -> int {
    int s2 = batch_rebuild_items(b);
    if (s2) return s2;
    s2 = dim2d_batch_integrate(b, dts, 2);  // warm-up, graph capture
    if (s2) return s2;
    HIP_TRY(hipEventRecord(e0, b->stream));
    s2 = dim2d_batch_integrate(b, dts, steps);
    if (s2) return s2;
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms / steps;
    return DIM2D_OK;
  };
  auto run = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(keep, b->U, b->total * sizeof(double), hipMemcpyDeviceToDevice,
                           b->stream));
    kept = true;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double ms = 0.0;
    if (b->fuse) {
      // the stage pair's tiles first, with the stage-3 tiles as they are
      int best_r = b->pair_rows, best_w = pair_wg;
      double pbest = INFINITY;
      for (int r : prows_c)
        for (int w : pwg_c) {
          b->pair_rows = r;
          pair_wg = w;
          int s2 = time_tiles(&ms);
          if (s2) return s2;
          if (ms < pbest) {
            pbest = ms;
            best_r = r;
            best_w = w;
          }
        }
      b->pair_rows = best_r;
      pair_wg = best_w;
    }
    int best_rows = b->rows, best_wg = b->wg;
    for (int r : rows_c)
      for (int w : wg_c) {
        // the stage-3 workgroup count is the mixed launches' when pipelined
        if (b->pipe && w != wg_c[0]) continue;
        b->rows = r;
        b->wg = w;
        int s2 = time_tiles(&ms);
        if (s2) return s2;
        if (ms < best) {
          best = ms;
          best_rows = r;
          best_wg = w;
        }
      }
    b->rows = best_rows;
    b->wg = best_wg;
    int s2 = batch_rebuild_items(b);
    if (s2) return s2;
    HIP_TRY(hipMemcpyAsync(b->U, keep, b->total * sizeof(double), hipMemcpyDeviceToDevice,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DIM2D_OK;
  };
  const long long steps0 = b->steps;
  st = run();
  // the candidates' steps are not the caller's: the state they started from
  // comes back and the step count with it, also when a candidate failed
  if (st && kept) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipMemcpyAsync(b->U, keep, b->total * sizeof(double), hipMemcpyDeviceToDevice,
                         b->stream);
  }
  b->steps = steps0;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(keep);
  if (st) return st;
  out[0] = b->rows;
  out[1] = b->wg;
  out[2] = best;
  return DIM2D_OK;
}

int dim2d_batch_synchronize(dim2d_batch* b) {
  if (!b) return fail(DIM2D_EINVAL, "batch is NULL");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return DIM2D_OK;
}

}  // extern "C"
