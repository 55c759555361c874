// dim2d_internal.hpp -- structures shared by the host library and the kernels.
//
// Field layout in HBM: cells interleaved, each cell the four conservative
// variables (mass, momentum_x, momentum_y, total energy) in 32 contiguous bytes
// (16 for the FP32 increments of the mixed mode), ny rows of `pitch` cells
// (pitch >= nx, a multiple of 32 cells).  Element (i,j,k) of the reference's
// nodes(nx,ny,ne) (1-based) sits at ((j-1)*pitch + (i-1))*4 + (k-1).  A wave
// reading one row of 64 cells reads 2 KiB contiguous bytes; the upload and
// download calls convert from and to the reference's planar order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dim2d {

constexpr int kBcReflection = 0;
constexpr int kBcPeriodic = 1;
constexpr int kBcWall = 2;
constexpr int kBcOpen = 3;  // hedstrom_xy, on all four edges at once
constexpr int kBcSize = 2;
// largest grid extent per axis (points including ghosts) a context accepts
constexpr int kMaxExtent = 1 << 16;
// the reference's debug_real (NF90_FILL_DOUBLE), written into the ghost cells
// of a time derivative as td_operators_class.f:150-154 does
constexpr double kDebugReal = 9.9692099683868690e+36;

// Per-simulation constants, passed by value (single simulation) or read from
// a device array (batch).  dx, dy are x_map(2)-x_map(1), y_map(2)-y_map(1) as
// in td_operators_class.f:156-157.
struct DevParams {
  int nx, ny, pitch, bc_w;
  int bc_e, bc_s, bc_n, gravity_ac;
  long long ncell;     // pitch * ny cells (4 elements each)
  long long offset;    // element offset of this simulation inside batch buffers
  double dx, dy;
  double idx, idy;         // 1/dx, 1/dy
  double half_idx, half_idy; // 0.5/dx, 0.5/dy  (central gradients)
  double q_idx, q_idy;     // 0.25/dx, 0.25/dy (transverse face gradients)
  double h2x, h2y;         // 0.5/dx^2, 0.5/dy^2
  double re, we, pr, vr, cv;
  double inv_re, inv_we, inv_pr, inv_repr;
  double c15cv, c3cv, c38cv;  // 1.5/cv, 3/cv, 3/(8cv)
  double c3cv_3;           // c3cv / 3 (the quotient, as the sound speed takes it)
  double two_vr;           // 2 + viscous_r
  double two_inv_re, two_inv_we;   // 2/Re, 2/We
  double k_heat_x, k_heat_y;       // 2/(Pr dx), 2/(Pr dy)
  double k_cap_x, k_cap_y;         // 2/(Re Pr dx), 2/(Re Pr dy)

  double gravity;
  double wall_cos_angle;   // cos(contact angle)
  double wall_q;           // constant wall heat flux
  double sigma_pref;       // 2 sqrt(2)/(3 sqrt(We)) (surface tension prefactor)

  // the constants FP32 fluxes take, rounded once (mixed precision): read as
  // scalar operands instead of being converted in every row
  struct F32 {
    float idx, idy, half_idx, half_idy, q_idx, q_idy;
    float vr, two_vr, inv_re, inv_we, two_inv_re, two_inv_we, c15cv;
    float k_heat_x, k_heat_y, k_cap_x, k_cap_y;
    float inv_pr, inv_repr;
    float c3cv_3;          // (float)c3cv / 3 (the mixed sound speed's quotient)
  } f;
};

// Step control block in device memory: dt and time of the running step and
// the reductions (for the CFL time step and the diagnostics) of the state at
// the end of each step, double-buffered by step parity.
struct StepCtl {
  double dt, t;
  double cfl, h;                 // CFL number and min(dx,dy)
  double re, we;
  double sum[2][4];              // sum rho, rho u, rho v, E over interior cells
  unsigned long long amax[2];    // max(|u|+c) as bits of a non-negative double
  unsigned long long rmin[2];    // min rho as bits (non-negative)
  unsigned long long rmax[2];    // max rho as bits
  unsigned int nonfinite[2];
  unsigned int pad[2];
};

// one tile of the batch work list: strip bx, rows 2 + by*rows .. of
// simulation sim (rows: that simulation's tile height); kind: what the tile
// runs in the mixed launch of a pipelined batch (kItemPair: the stage pair on
// a 56-column strip; kItemStage3: stage 3 on 60-column strips 2 bx and
// 2 bx + 1, one per wave)
constexpr int kItemPair = 0, kItemStage3 = 1;
struct WorkItem {
  int sim, bx, by, rows;
  int kind, pad0, pad1, pad2;
};

// kernel launch helpers (dim2d_kernels.hip)
struct StageArgs {
  const double* U;    // u^n (FP64)
  const double* S64;  // stage input state (FP64 mode)
  const float* S32;   // stage increment (mixed mode), state = U + S32
  double* O64;        // FP64 output (FP64 stages, mixed stage 3, time_dev)
  float* O32;         // FP32 output (mixed stages 1, 2)
  double* Q;          // low-storage RK register q (read and written in place)
  float* Q32;         // the same in mixed precision (FP32)
  StepCtl* ctl;
  double* F64;        // stage pair with open edges: s1 of the frame and the band
  float* F32;         // beside it (FP64), or its increment (mixed)
  const double* dts;  // per-simulation dt (batch)
  double dt_fixed;
  int cfl_mode;       // 1: stage 1 takes dt from the reductions
  int parity;         // step parity (reduction slot)
};

hipError_t launch_stage(const DevParams& p, int mixed, int stage, const StageArgs& a, int block_x,
                        int rows, hipStream_t s);
// RK3-TVD stages 1 and 2 in one launch (stage pair): u^n (a.U) -> stage-2
// state in a.O64 (FP64) or its increment in a.O32 (mixed); dt as stage 1.
// pair_supported: reflection, periodic, S/N wall and open edges, grids of at
// least 8 interior points per axis.  With open edges the stage-1 frame kernel
// (output a.F64 / a.F32) runs before it and the stage-2 frame kernel (input
// the same buffer) after it.
bool pair_supported(const DevParams& p);
// lsrk: the low-storage RK3's stages 1 and 2 instead (FP64 without open
// edges): u0 (a.U) -> u2 in a.O64 and q2 in a.Q
hipError_t launch_stage_pair(const DevParams& p, int mixed, int lsrk, const StageArgs& a, int rows,
                             hipStream_t s);
// the stage pair over a batch work list of pair tiles (56-column strips),
// persistent grid of `workgroups` two-wave workgroups, per-simulation dt from
// a.dts; every simulation must satisfy pair_supported
hipError_t launch_stage_pair_batch(const DevParams* d_params, const WorkItem* d_items,
                                   int nitems, int mixed, const StageArgs& a, int workgroups,
                                   hipStream_t s);
constexpr int kPairOutCols = 56;  // output columns per stage-pair strip
// a pipelined batch's mixed launch: the items of kind kItemPair run the stage
// pair (u^n in pU -> the stage-2 state in pO64, or its increment in pO32),
// those of kind kItemStage3 run stage 3 (u^n in sU and the stage-2 state in
// sS64, or its increment in sS32 -> u^n+1 in sO64), in one persistent grid of
// two-wave workgroups; batch base pointers, per-simulation dt from dts
struct MixArgs {
  const double* pU;
  double* pO64;
  float* pO32;
  const double* sU;
  const double* sS64;
  const float* sS32;
  double* sO64;
};
hipError_t launch_stage_mix_batch(const DevParams* d_params, const WorkItem* d_items, int nitems,
                                  int mixed, const MixArgs& m, const double* dts, int workgroups,
                                  hipStream_t s);
int max_resident_mix_workgroups(int mixed);
// one RK stage over a batch work list of 60-column strips, persistent grid
// of `workgroups` one-wave workgroups
hipError_t launch_stage_batch(const DevParams* d_params, const WorkItem* d_items, int nitems,
                              int mixed, int stage, const StageArgs& a, int workgroups,
                              hipStream_t s);
// wall ghost cells of the state `st` (= U + S32 in mixed mode) written into
// O64 or O32; tmp is u^n for the root bracket guess.
hipError_t launch_wall(const DevParams& p, int mixed_in, const double* U, const double* S64,
                       const float* S32, double* O64, float* O32, const double* tmp,
                       hipStream_t s);
hipError_t launch_wall_batch(const DevParams* h_params, const DevParams* d_params, int nsim,
                             int mixed_in, const double* U, const double* S64, const float* S32,
                             double* O64, float* O32, const double* tmp, hipStream_t s);
// store reflection/periodic ghost cells of an FP64 field
hipError_t launch_fill_ghosts(const DevParams& p, double* F, hipStream_t s);
// open boundaries: time derivative and RK update of the ghost frame for a
// stage (0: time derivative into a.O64; 1-3 RK3-TVD; 4-6 low-storage RK3)
hipError_t launch_frame(const DevParams& p, int mixed, int stage, const StageArgs& a,
                        hipStream_t s);
// reductions of an FP64 field into ctl slot `slot` (slot cleared first)
hipError_t launch_reduce(const DevParams& p, const double* F, StepCtl* ctl, int slot,
                         hipStream_t s);
// planar nodes(nx,ny,4) (the reference's order, contiguous) -> interleaved cells
hipError_t launch_pack(const DevParams& p, const double* planar, double* F, hipStream_t s);
// interleaved cells -> planar nodes(nx,ny,4); with fill_ghosts the ghost cells
// receive `ghost_value` instead of the field's values
hipError_t launch_unpack(const DevParams& p, const double* F, double* planar, int fill_ghosts,
                         double ghost_value, hipStream_t s);

// Cockburn and Gau operators, FP64: time derivative (stage 0) or RK3-TVD
// stage 1-3 from a.S64 (and a.U), into a.O64; stage 3 also reduces for the CFL
// step, as launch_stage does
hipError_t launch_cg_stage(const DevParams& p, int mixed, int stage, const StageArgs& a,
                           hipStream_t s);
// max |F| per variable over the interior (and the ghost frame when `frame`)
// into out[0..3] as bits of non-negative doubles (atomicMax; clear out first);
// non-finite values count as +inf
hipError_t launch_absmax(const DevParams& p, int frame, const double* F, unsigned long long* out,
                         hipStream_t s);
// open edges: number of undermined points (imaginary sound speed; with
// md_thr >= 0 also a density inside the two-phase band) in the `dist`
// interior rows/columns nearest each edge, added into counts[0..3] (W, E, S, N)
// rectangles [i0, i1) x [j0, j1) of grid points (0-based)
struct Rects4 {
  int i0[4], i1[4], j0[4], j1[4];
};
// undermined points per rectangle into counts[0..3] (counts cleared by the
// caller); every point's 3x3 neighbourhood must lie in the array
hipError_t launch_undermined_rects(const DevParams& p, const Rects4& r, double md_thr,
                                   const double* U, int* counts, hipStream_t s);
hipError_t launch_undermined(const DevParams& p, int dist, double md_thr, const double* U,
                             int* counts, hipStream_t s);
// state of grid `op` copied into grid `np`, grown by add_w columns west and
// add_s rows south (any number east and north); new points repeat the
// nearest old point
hipError_t launch_extend(const DevParams& np, const DevParams& op, int add_w, int add_s,
                         const double* oldU, double* newU, hipStream_t s);

// open grids grown by one layer (dim2d_newgrdpt.hip): cells of the frame
// snapshot (the two outermost rows and columns) of an nx x ny grid; the
// snapshot of U; the new points of grid `np`, grown from `op` by aw columns
// west and as rows south (0 or 1; east and north implied by the extents),
// whose old points are already in U; x0, y0 are the coordinates of np's
// point (0, 0), ff the conservative far-field state, dt is read from ctl
long long frame_cells(int nx, int ny);
hipError_t launch_save_frame(const DevParams& p, const double* U, double* F0, hipStream_t s);
hipError_t launch_newgrdpt(const DevParams& np, const DevParams& op, int aw, int as, double x0,
                           double y0, const double ff[4], const StepCtl* ctl, const double* F0,
                           double* U, hipStream_t s);

// workgroups of the persistent batch kernels that fit the device at once
int max_resident_workgroups(int mixed);
// the same for the batch stage-pair kernel (two-wave workgroups)
int max_resident_pair_workgroups(int mixed);

}  // namespace dim2d
