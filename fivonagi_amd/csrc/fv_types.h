// fv_types.h -- device-side tables shared by the step kernel and the host
// runtime (fv_runtime.cpp).  Plain C layout, no torch types.
#ifndef FIVONAGI_FV_TYPES_H
#define FIVONAGI_FV_TYPES_H

#include <stdint.h>

namespace fv {

// Output columns per strip: one wave = 64 lanes = 58 output columns plus a
// 3-column halo on each side (dependency radius of y-sweep -> x-sweep -> source).
constexpr int kLanes = 64;
constexpr int kHalo = 3;
constexpr int kStripCols = kLanes - 2 * kHalo;   // 58

enum BoundaryKind : int32_t { kBcZeroGrad = 0, kBcEdgeDrive = 1, kBcTaylor = 2, kBcHifu = 3 };
enum CflMask : int32_t { kCflAll = 0, kCflEven = 1, kCflBitmap = 2 };

// Per-instance control block (device memory): the parameters of one step
// attempt.  Two per instance, alternating with the attempt index: launch a
// reads block a%2 (attempt a-1's parameters) and the reduction slots of
// attempt a-1, forms attempt a's parameters in every wave, and one wave
// stores them into block (a+1)%2.
struct Ctl {
  double dt;            // dt of the next attempt
  double T;             // time at the start of the next attempt
  double last_cfl;      // cfl of the last attempt
  double drive_s, drive_c;   // sin/cos(omega*T) for the next attempt
  int64_t n_acc, n_rej, n_att;
  int64_t target_acc;   // waves skip the instance once n_acc reaches this
  int32_t cur;          // buffer holding the current state (0/1)
  int32_t err;          // 0 ok, 2 imaginary eigenvalue
  float dtx, dty;       // dt/dx, dt/dy of the next attempt
  float ddx2, ddy2;     // delta*dt/dx^2, delta*dt/dy^2
  int32_t mdx, mdy;     // moving-window offset of the current state (MDX, MDY)
  double T_stop;        // ... or once T reaches this (the frame loop's Tprint)
};

// Reduction slots of one attempt, per instance: the waves of an instance
// are split into kGroups groups (local item index mod kGroups), each with its
// own 128-byte line, so the same-address atomics of one step are spread over
// kGroups lines (one line taking every wave's atomics slows the whole step).
// Three sets per instance, by attempt index mod 3: launch a reduces into set
// a%3, reads set (a-1)%3 and clears set (a+1)%3.
constexpr int kGroups = 8;
struct alignas(128) GroupSlot {
  uint32_t cfl_bits;    // max cfl of the group's cells (float bits, >= 0)
  uint32_t pad0;
  uint64_t cfl_bits64;  // the same for FP64 instances (double bits, >= 0)
  uint32_t pad[28];
};
struct alignas(128) SyncSlots {
  GroupSlot grp[kGroups];
  uint32_t err_flag;    // any imaginary eigenvalue in the attempt
  uint32_t ran;         // the attempt ran (the instance was not done or stopped)
  uint32_t pad[30];
};

// Read-only per-instance description (device memory).
struct InstDesc {
  float* q[2][3];       // ping-pong SoA state buffers, offset i + j*nx
  Ctl* ctl;             // 2 control blocks (attempt index mod 2)
  const void* edge_c;   // amp_j * cos(phase_j), length ny (kBcEdgeDrive); float or
  const void* edge_s;   // amp_j * sin(phase_j), length ny         double (fp64)
  const double* dt_seq; // prescribed dt per attempt, or null
  int64_t dt_seq_len;
  Ctl* ctl_log;         // when set: the control block accepted step n_acc - log_base ran with
  int64_t log_base, log_len;
  float* hist;          // when set (FP32): state n_acc - log_base after each accepted
                        // step, slot t = 3 grids, t in [1, log_len] (the extended step kernel)
  double dx, dy, delta, cfl_max, cfl_wish, omega;
  double inv_dx, inv_dy;     // 1/dx, 1/dy
  double dif_x, dif_y;       // delta/dx^2, delta/dy^2
  double omega_turns;        // omega / (2 pi)
  double tay_rhoa, tay_c, tay_cos, tay_sin, tay_yoff;
  int32_t nx, ny;
  int32_t bc, cfl_mask;  // cfl_mask: host-side record of the fold subset kind
  SyncSlots* sync;      // 3 reduction slot sets (attempt index mod 3)
  int32_t n_items;      // work items (waves) of this instance per step
  int32_t pad2;
  int32_t tay_jc;
  int32_t md_on;        // moving window (common/main.cu:109-177)
  float beta2m2;        // 2*(BETA - 1)
  float pad;
  double md_vx, md_vy, md_t;   // MDVX, MDVY, MDT
  // kBcHifu: edge_c/edge_s hold AMPL cos/sin(ISPW phi_j); per row j,
  // hifu_tab[4j..4j+3] = phi_j, nu_x, nu_y, 1 on aperture rows (0 elsewhere)
  const void* hifu_tab; // float or double (fp64)
  float hifu_k;         // FREC/4: envelope argument per unit of T + phi
  int32_t jcut;         // data collection: row (JCUT)
  // data collection (apps/hifu-beam/data_collect.h, apps/taylor-angle-xy/
  // data_collect.h): measure = nx*ny grid (null: off), stage = row jcut of
  // the last attempt before the window shift (hifu); float or double (fp64)
  void* measure;
  void* stage;
  uint32_t* collect_err;  // 4 once a frame holds more than nx steps
  int64_t n0frame;      // accepted steps when the current frame started
  double eta;           // cells per length unit
  int32_t mic[4];
  // taylor-angle-xy cut: i in [icenter - cut_half, icenter + cut_half]
  int32_t cut_half, pad6;
  double tan_theta;
  // FP64 instances (the reference's PRECISION=2): state, tables and the cfl
  // reduction in double
  int32_t fp64, pad5;
  double hbeta_d;       // BETA - 1
  double hifu_k_d;      // FREC/4
};

// One wave's work: a 58-column strip of one instance, rows [y0, y1).
// mrow: index of the item's first row in the CFL lane-mask table; entry
// mrow + (j - y0) has bit l set when lane l's cell (x0 - 3 + l, j) is an
// output cell with i >= 1 that the CFL reduction counts.  local: the item's
// index among its instance's items (0 publishes the instance's control block;
// local % kGroups picks the reduction slot), independent of the launch order.
struct WorkItem {
  int32_t inst;
  int32_t x0;           // first output column of the strip
  int32_t y0, y1;
  int32_t mrow;
  int32_t local;
  int32_t bc;           // the instance's boundary kind (picks the app's march)
  int32_t pad;
};

}  // namespace fv

#endif  // FIVONAGI_FV_TYPES_H
