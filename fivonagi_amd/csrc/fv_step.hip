// fv_step.hip -- the fused FiVoNAGI time step for gfx950 (MI355X).
//
// One launch = one step attempt of every active instance.  It replaces the
// reference's per-step sequence (common/main.cu:310-635): texCpy, boundary,
// getWavesSpeedsCFL_y, step_y, calcLimiters_y, writeLimiters_y, the same four
// in x, 21 reduceCFL launches, reduceCFL2, source, update_T, update_dt and the
// seven blocking device->host copies.
//
// Work decomposition ("march"): one wave64 per work item.  Lane l owns column
// x0 - 3 + l; the wave marches upward through rows [y0-3, y1+3) of its strip:
//   load row r+1 (prefetch; boundary rules applied when the row is consumed)
//   -> y-Riemann at interface r        (registers, same lane)
//   -> y-limiter at interface r-1      (3-interface register window)
//   -> q_y(r-2)                        (first + second order, fused)
//   -> x-Riemann/limiter on row r-2    (neighbours via DPP wave_shr/wave_shl)
//   -> q_xy(r-2)
//   -> source on row r-3               (3-row register window + DPP)
//   -> store row r-3 (lanes 3..60)
// The loop is unrolled three times so the period-3 register windows rotate
// by renaming instead of moves.  No LDS and no barriers: the 3-column halo on
// each side makes the waves independent.
//
// CFL and accept/reject: each wave reduces its cells' cfl to one float and
// folds it into its group's 128-byte line (8 groups per instance) with one
// relaxed agent-scope atomicMax that nothing waits for.  Attempt a's maxima
// are read by every wave of launch a+1, which decides accept/reject, dt, T
// and the ping-pong index of attempt a+1 itself (advance(), identical in
// every wave); one wave per instance stores the result for the launch after.
// The control blocks and reduction slots rotate with the attempt index
// (2 control blocks, 3 slot sets), so no launch reads what it writes and the
// step ends when its last wave ends -- no ticket chain, no fences.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../apps/apps.h"
#include "../apps/hifu-beam/data_collect.h"
#include "../apps/taylor-angle-xy/data_collect.h"
#include "../apps/na/solver.h"
#include "../apps/na/source.h"
#include "fv_types.h"

namespace fv {

typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float cgfloat;
typedef __attribute__((address_space(4))) const float kfloat;   // read-only: scalar loads
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double cgdouble;
typedef __attribute__((address_space(4))) const double kdouble;
// the same three per real type R (float: the FP32 path, double: FP64)
template <typename R> struct Mem;
template <> struct Mem<float> { typedef gfloat g; typedef cgfloat cg; typedef kfloat k; };
template <> struct Mem<double> { typedef gdouble g; typedef cgdouble cg; typedef kdouble k; };

// DPP cross-lane moves over the whole wave64 (gfx9 wave_shr:1 / wave_shl:1).
// shr1: lane l receives lane l-1's value (lane 0 receives 0).
// shl1: lane l receives lane l+1's value (lane 63 receives 0).
// bound_ctrl = 1: the missing source lane reads as 0, so no "old" value has
// to be materialised first.
__device__ __forceinline__ float shr1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float shl1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
// 64-bit values move as two 32-bit halves
__device__ __forceinline__ double dpp64(double v, int ctrl_shr) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)(ctrl_shr ? __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x138, 0xf, 0xf, true)
                                          : __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x130, 0xf, 0xf, true));
  const uint32_t hi = (uint32_t)(ctrl_shr ? __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x138, 0xf, 0xf, true)
                                          : __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x130, 0xf, 0xf, true));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ double shr1(double v) { return dpp64(v, 1); }
__device__ __forceinline__ double shl1(double v) { return dpp64(v, 0); }

typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) const char cgchar;

// Loads/stores at a 32-bit byte offset from a uniform base: the base stays in
// scalar registers (global_load/store with saddr) and each lane carries one
// 32-bit offset instead of a 64-bit address per field.
__device__ __forceinline__ float ldg(cgfloat* base, uint32_t off) {
  return *(cgfloat*)((cgchar*)base + off);
}
__device__ __forceinline__ void stg(gfloat* base, uint32_t off, float v) {
  *(gfloat*)((gchar*)base + off) = v;
}
__device__ __forceinline__ double ldg(cgdouble* base, uint32_t off) {
  return *(cgdouble*)((cgchar*)base + off);
}
__device__ __forceinline__ void stg(gdouble* base, uint32_t off, double v) {
  *(gdouble*)((gchar*)base + off) = v;
}

// s_waitcnt vmcnt(0) with expcnt and lgkmcnt left at their maxima (gfx9 encoding)
constexpr int kWaitVm0 = 0x0F70;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float unif(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double unif(double v) {
  return __longlong_as_double(uni64(__double_as_longlong(v)));
}

// v on the lanes whose bit of the uniform 64-bit mask is set, 0 elsewhere
// (one v_cndmask with the mask as the lane condition).
__device__ __forceinline__ float lane_sel(uint64_t mask, float v) {
  float d;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(d) : "v"(v), "s"(mask));
  return d;
}
__device__ __forceinline__ double lane_sel(uint64_t mask, double v) {
  return ((mask >> __lane_id()) & 1ull) ? v : 0.0;
}

// Pins a value in a register here: its computation cannot sink into a
// branch that follows, so a cross-lane move feeding it stays in its block
// and folds into its user.
__device__ __forceinline__ void pin(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(double& v) { asm volatile("" : "+v"(v)); }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }

__device__ __forceinline__ uint32_t amax_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t aadd_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t axchg_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The per-group cfl maximum of an instance, as float or double bits
// (non-negative values order like their bit patterns).
template <typename R> struct CflSlot;
template <> struct CflSlot<float> {
  typedef uint32_t bits;
  static __device__ __forceinline__ bits load(const GroupSlot& g) { return g.cfl_bits; }
  static __device__ __forceinline__ void clear(GroupSlot& g) { g.cfl_bits = 0u; }
  static __device__ __forceinline__ void fold_in(GroupSlot& g, float v) {
    __hip_atomic_fetch_max(&g.cfl_bits, __float_as_uint(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ float value(bits b) { return __uint_as_float(b); }
  static __device__ __forceinline__ bits lane_max(bits g) {
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) m = max(m, (uint32_t)__builtin_amdgcn_readlane(g, k));
    return m;
  }
};
template <> struct CflSlot<double> {
  typedef uint64_t bits;
  static __device__ __forceinline__ bits load(const GroupSlot& g) { return g.cfl_bits64; }
  static __device__ __forceinline__ void clear(GroupSlot& g) { g.cfl_bits64 = 0ull; }
  static __device__ __forceinline__ void fold_in(GroupSlot& g, double v) {
    __hip_atomic_fetch_max(&g.cfl_bits64, (uint64_t)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ double value(bits b) {
    return __longlong_as_double((long long)b);
  }
  static __device__ __forceinline__ bits lane_max(bits g) {
    uint64_t m = 0ull;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)g, k);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(g >> 32), k);
      m = max(m, ((uint64_t)hi << 32) | lo);
    }
    return m;
  }
};

// What attempt a's reduction slots say about it, with the control block it
// ran with.  fold_load issues every load before anything is decided, so the
// fold costs one memory round trip.
// R: the instance's real type (its cfl reduction is in R).
template <typename R>
struct FoldRaw {
  Ctl p;          // control block attempt a ran with
  uint32_t ran, errf;
  typename CflSlot<R>::bits gbits;
};
template <typename R>
struct Fold {
  Ctl p;
  R cfl;          // max cfl over the counted cells of attempt a
  int state;      // kFoldIdle, kFoldError, kFoldAccept, kFoldReject
};
constexpr int kFoldIdle = 0, kFoldError = 1, kFoldAccept = 2, kFoldReject = 3;

template <typename R>
__device__ __forceinline__ FoldRaw<R> fold_load(const Ctl* __restrict__ c,
                                                const SyncSlots* __restrict__ sl) {
  FoldRaw<R> r;
  Ctl& p = r.p;
  p.dt = c->dt; p.T = c->T; p.last_cfl = c->last_cfl;
  p.drive_s = c->drive_s; p.drive_c = c->drive_c;
  p.n_acc = c->n_acc; p.n_rej = c->n_rej; p.n_att = c->n_att;
  p.target_acc = c->target_acc;
  p.cur = c->cur; p.err = c->err;
  p.dtx = c->dtx; p.dty = c->dty; p.ddx2 = c->ddx2; p.ddy2 = c->ddy2;
  p.mdx = c->mdx; p.mdy = c->mdy; p.T_stop = c->T_stop;
  r.ran = sl->ran;
  r.errf = sl->err_flag;
  // the group maxima with one vector load (lane l reads group l % kGroups)
  // rather than kGroups scalar loads
  r.gbits = CflSlot<R>::load(sl->grp[__lane_id() & (kGroups - 1)]);
  return r;
}

template <typename R>
__device__ __forceinline__ Fold<R> fold_decide(const FoldRaw<R>& r, double cfl_max, bool presc) {
  Fold<R> f;
  f.p = r.p;
  f.p.n_acc = uni64(r.p.n_acc); f.p.n_rej = uni64(r.p.n_rej); f.p.n_att = uni64(r.p.n_att);
  f.p.target_acc = uni64(r.p.target_acc);
  f.p.cur = uni(r.p.cur); f.p.err = uni(r.p.err);
  f.p.mdx = uni(r.p.mdx); f.p.mdy = uni(r.p.mdy);
  const R cfl = CflSlot<R>::value(CflSlot<R>::lane_max(r.gbits));
  f.cfl = cfl;
  const bool acc = presc || ((double)cfl < cfl_max);
  f.state = !uni(r.ran) ? kFoldIdle
                        : (uni(r.errf) ? kFoldError : (acc ? kFoldAccept : kFoldReject));
  return f;
}

// Ping-pong index attempt a+1 reads (known before the rest of advance()).
template <typename R>
__device__ __forceinline__ int next_cur(const Fold<R>& f) {
  return f.state == kFoldAccept ? (f.p.cur ^ 1) : f.p.cur;
}

using fvbc::md_shift;

// Control block of attempt a+1 from attempt a's fold: common/main.cu:199-230
// with a clean retry on rejection.  log: record the control block an accepted
// step ran with (adjoint runs: its recompute and reverse pass replay it); every
// writer records the same values.
template <typename R>
__device__ Ctl advance(const InstDesc& D, const Fold<R>& f, bool log) {
  const Ctl& p = f.p;
  Ctl o = p;
  if (f.state == kFoldIdle) return o;   // attempt a did not run (instance done or stopped)
  if (f.state == kFoldError) {
    o.err = 2;                         // imaginary eigenvalue: step discarded, instance stops
    return o;
  }
  const double cfl = (double)f.cfl;
  const bool presc = D.dt_seq != nullptr;
  const bool acc = f.state == kFoldAccept;
  const int64_t att = p.n_att + 1;
  o.n_att = att;
  o.last_cfl = cfl;
  double T = p.T;
  if (acc) {
    if (log && D.ctl_log && __lane_id() == 0) {
      const int64_t idx = p.n_acc - D.log_base;
      if (idx >= 0 && idx < D.log_len) D.ctl_log[idx] = p;
    }
    T = p.T + p.dt;
    o.T = T;
    o.n_acc = p.n_acc + 1;
    o.cur = p.cur ^ 1;
    if (D.md_on) {
      o.mdx = p.mdx + md_shift(T, D.md_t, D.md_vx, D.dx, p.mdx);
      o.mdy = p.mdy + md_shift(T, D.md_t, D.md_vy, D.dy, p.mdy);
    }
  } else {
    o.n_rej = p.n_rej + 1;
  }
  double ndt;
  if (presc) {
    // attempt att uses dt[att]; past the end the dt is kept, and an attempt
    // that would run with it is refused (prescribed_exhausted)
    ndt = (att < D.dt_seq_len) ? D.dt_seq[att] : p.dt;
  } else {
    ndt = p.dt * D.cfl_wish / cfl;
    if (!acc && ndt > p.dt * 0.9) ndt = p.dt * 0.9;
  }
  o.dt = ndt;
  o.dtx = (float)(ndt * D.inv_dx);
  o.dty = (float)(ndt * D.inv_dy);
  o.ddx2 = (float)(ndt * D.dif_x);
  o.ddy2 = (float)(ndt * D.dif_y);
  // drive phase omega*T in turns, reduced in double to [0, 1); the step uses
  // sin/cos in FP32
  const double turns = D.omega_turns * T;
  if (sizeof(R) == 4) {
    float fs, fc;
    sincospif((float)(2.0 * (turns - floor(turns))), &fs, &fc);
    o.drive_s = fs;
    o.drive_c = fc;
  } else {                      // FP64 instances: the drive phase in double
    double ds, dc;
    sincospi(2.0 * (turns - floor(turns)), &ds, &dc);
    o.drive_s = ds;
    o.drive_c = dc;
  }
  return o;
}

// An attempt that would run (instance not done, not stopped, no error) while
// its prescribed dt sequence has no entry for it: error 9 instead of running.
// The check sits where an attempt is about to run, not where the last entry
// is used, so a sequence of exactly n entries runs n steps.
__device__ __forceinline__ void prescribed_exhausted(const InstDesc& D, Ctl& P) {
  if (D.dt_seq != nullptr && P.err == 0 && P.n_att >= D.dt_seq_len &&
      P.n_acc < P.target_acc && P.T < P.T_stop)
    P.err = 9;
}

// Control block of the attempt a launch runs (phase = a mod 6), read-only.
template <typename R>
__device__ __forceinline__ Ctl current_t(const InstDesc& D, int phase, bool log) {
  Ctl P = advance(D,
                  fold_decide(fold_load<R>(D.ctl + (phase & 1), D.sync + (phase + 2) % 3),
                              D.cfl_max, D.dt_seq != nullptr),
                  log);
  prescribed_exhausted(D, P);
  return P;
}
__device__ __forceinline__ Ctl current(const InstDesc& D, int phase, bool log) {
  return D.fp64 ? current_t<double>(D, phase, log) : current_t<float>(D, phase, log);
}

template <typename R>
struct Row {         // one cell of a row as the y sweep sees it
  R q1, q2, q3;      // rho', rho u, rho v
  R un, ut;          // y normal velocity v = q3/rho, tangential u = q2/rho
};

template <typename R>
struct YI {          // y interface with its limited flux towards the + cell
  fvna::IfaceT<R> f;
  R dl[3];           // W_p . W_p of the interface below (one per wave family)
  R P[3];            // apdq - fadd, in (density, normal, tangential) order
};

template <typename R>
struct XY {          // one cell of q_xy, with its velocities for the source
  R q1, q2, q3, u, v;
};

// One work item of attempt `phase` (mod 6) of an instance whose ghost cells
// follow the boundary policy BC (apps/<app>/boundary.h).
// kOne: the launch covers a single instance, whose control block and
// reduction slots of this attempt are passed directly (ctl_in = its control
// block of attempt a = ctl + a % 2, sync_in = the slots of attempt a-1 =
// sync + (a + 2) % 3), so the fold's loads go out together with the work
// item's instead of after the instance description's.
// kExt: some instance has a moving window or the HIFU boundary.
//  - Moving window: the shift that follows an accepted step (moveDomain,
//    common/main.cu:109-164) is folded into the stores: cell (i, j) of the new
//    state lands at (i - sx, j - sy), and also stays at (i, j) when
//    (i + sx, j + sy) is off the grid (moveDomain's else branch); every cell
//    of the output is written exactly once.  A rejected attempt's output is
//    discarded, as without the window.
//  - the HIFU app's packet and data collection staging.
// R: float (FP32 path) or double (FP64 instances: state, tables, arithmetic
// and the cfl reduction in double; the reference's PRECISION=2).
template <typename R, bool kOne, bool kExt, class BC>
__device__ __forceinline__ void step_impl(const InstDesc* __restrict__ insts, const WorkItem& it,
                                          const uint64_t* __restrict__ rowmask, int phase,
                                          const Ctl* __restrict__ ctl_in,
                                          const SyncSlots* __restrict__ sync_in) {
  const InstDesc& D = insts[kOne ? 0 : it.inst];
  const int local = it.local;
  const int lane = threadIdx.x;

  // ---- the previous attempt's reduction, folded: accept/reject decides
  // which buffer this attempt reads, so its first row can be requested
  // before the rest of the control block (dt, T, drive phase) is formed ----
  // the instance's geometry and both buffers' pointers are requested with
  // the fold's loads; the asm barrier keeps the compiler from re-loading a
  // pointer by the buffer index once accept/reject is known
  const int nx = D.nx, ny = D.ny;
  typedef typename Mem<R>::g gR;
  typedef typename Mem<R>::cg cgR;
  R* qb[2][3];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int f = 0; f < 3; ++f) qb[h][f] = (R*)D.q[h][f];
  double cfl_max = D.cfl_max;
  const double* dt_seq = D.dt_seq;
  fvbc::Tables tb{D.edge_c, D.edge_s, D.hifu_tab};
  const FoldRaw<R> FR = fold_load<R>(kOne ? ctl_in : D.ctl + (phase & 1),
                                     kOne ? sync_in : D.sync + (phase + 2) % 3);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int f = 0; f < 3; ++f) asm volatile("" : "+s"(qb[h][f]));
  asm volatile("" : "+s"(cfl_max), "+s"(dt_seq));
  if (BC::kRowData) asm volatile("" : "+s"(tb.edge_c), "+s"(tb.edge_s));
  const Fold<R> F = fold_decide(FR, cfl_max, dt_seq != nullptr);
  const int cur = next_cur(F);

  cgR* __restrict__ in1 = (cgR*)(cur ? qb[1][0] : qb[0][0]);
  cgR* __restrict__ in2 = (cgR*)(cur ? qb[1][1] : qb[0][1]);
  cgR* __restrict__ in3 = (cgR*)(cur ? qb[1][2] : qb[0][2]);
  gR* __restrict__ o1 = (gR*)(cur ? qb[0][0] : qb[1][0]);
  gR* __restrict__ o2 = (gR*)(cur ? qb[0][1] : qb[1][1]);
  gR* __restrict__ o3 = (gR*)(cur ? qb[0][2] : qb[1][2]);

  const int i = it.x0 - kHalo + lane;
  const bool col_in = (i >= 0) && (i < nx);
  const bool col_x = col_in && (i >= 1);     // lane holds an x interface of the grid
  const int ic = min(max(i, 0), nx - 1);
  const bool col_int = (i >= 2) && (i <= nx - 3);
  const bool out_lane = (lane >= kHalo) && (lane < kLanes - kHalo) && (i < nx);
  const int y0 = it.y0, y1 = it.y1;
  // CFL lane masks of the item's rows: cm[j], j in [y0, y1)
  const uint64_t* __restrict__ cm = rowmask + (it.mrow - y0);

  // the lane's ghost-cell source columns (the app's boundary policy; Taylor
  // ghost cells were written into the input by fv_taylor_ghost_kernel and are
  // plain copies here)
  constexpr uint32_t kSz = sizeof(R);
  fvbc::LaneMap lm;
  BC::lane(ic, nx, lm, kSz);
  const uint32_t row_bytes = kSz * (uint32_t)nx;

  // raw (unpatched) load of row r, with the row's table entries; the
  // boundary patch is applied when the row is consumed one iteration later,
  // so the load latency overlaps a full step
  auto load_raw = [&](int r, R& a, R& b, R& c, fvbc::RowData<R>& h) {
    const int rc = min(max(r, 0), ny - 1);
    const uint32_t o = BC::offset(rc, ny, lm, row_bytes);
    a = ldg(in1, o);
    b = ldg(in2, o);
    c = ldg(in3, o);
    if (BC::kRowData) BC::template load_row<R>(tb, rc, h);
  };
  R n1, n2, n3;
  fvbc::RowData<R> nh{R(0), R(0), R(0), R(0), R(0), R(0)};
  load_raw(it.y0 - kHalo, n1, n2, n3, nh);   // in flight during advance()

  // ---- control block of this attempt ----
  Ctl P = advance<R>(D, F, local == 0);
  prescribed_exhausted(D, P);
  const bool runs = P.err == 0 && P.n_acc < P.target_acc && P.T < P.T_stop;
  if (local == 0 && lane == 0) {
    // one wave per instance: publish the control block for the next launch,
    // clear the slots the next attempt reduces into, mark this attempt
    D.ctl[(phase + 1) & 1] = P;
    SyncSlots* nx_sl = D.sync + (phase + 1) % 3;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) CflSlot<R>::clear(nx_sl->grp[k]);
    nx_sl->err_flag = 0u;
    nx_sl->ran = 0u;
    if (runs) D.sync[phase % 3].ran = 1u;
  }
  if (!runs) return;
  R dtx, dty, ddx2, ddy2, sin_t, cos_t, hbeta;
  if (sizeof(R) == 4) {     // the control block's FP32 forms
    dtx = unif(P.dtx); dty = unif(P.dty);
    ddx2 = unif(P.ddx2); ddy2 = unif(P.ddy2);
    sin_t = unif((float)P.drive_s); cos_t = unif((float)P.drive_c);
    hbeta = R(0.5f * D.beta2m2);            // BETA - 1
  } else {                  // FP64: the same quantities in double
    dtx = unif(R(P.dt * D.inv_dx)); dty = unif(R(P.dt * D.inv_dy));
    ddx2 = unif(R(P.dt * D.dif_x)); ddy2 = unif(R(P.dt * D.dif_y));
    sin_t = unif(R(P.drive_s)); cos_t = unif(R(P.drive_c));
    hbeta = R(D.hbeta_d);
  }
  // shift of the window after this attempt, if it is accepted
  int sx = 0, sy = 0;
  if (kExt && D.md_on) {
    const double Tn = P.T + P.dt;
    sx = uni(md_shift(Tn, D.md_t, D.md_vx, D.dx, P.mdx));
    sy = uni(md_shift(Tn, D.md_t, D.md_vy, D.dy, P.mdy));
  }
  typename Mem<R>::g* __restrict__ stage = (typename Mem<R>::g*)D.stage;
  const bool stage_on = kExt && stage != nullptr;
  // adjoint forward runs that keep every state: this attempt's output also
  // goes to history slot n_acc + 1 - log_base, the slot of the count it
  // makes if accepted.  A rejected attempt's copy is overwritten by the
  // retry, which has the same count, so the slot ends with the accepted
  // state (FP32 instances, extended instantiation only).
  gfloat* __restrict__ hs1 = nullptr;
  gfloat* __restrict__ hs2 = nullptr;
  gfloat* __restrict__ hs3 = nullptr;
  if constexpr (kExt && sizeof(R) == 4) {
    const int64_t t = P.n_acc + 1 - D.log_base;
    if (D.hist != nullptr && t >= 1 && t <= D.log_len) {
      const size_t n = (size_t)nx * (size_t)ny;
      hs1 = (gfloat*)(D.hist + (size_t)t * 3 * n);
      hs2 = hs1 + n;
      hs3 = hs2 + n;
    }
  }
  const bool hist_on = kExt && sizeof(R) == 4 && hs1 != nullptr;
  const int jcut = D.jcut;
  // what the boundary policy sees of this attempt; the moving window's MDX
  // only exists in the extended instantiation (else it is 0)
  fvbc::PatchCtx<R> pc;
  pc.sin_t = sin_t;
  pc.cos_t = cos_t;
  pc.T = P.T;
  pc.hk = sizeof(R) == 4 ? R(D.hifu_k) : R(D.hifu_k_d);
  pc.mdx = kExt ? P.mdx : 0;
  pc.x0 = it.x0;
  pc.nx = nx;
  pc.ny = ny;
  pc.ic = ic;
  pc.left = lm.left;
  pc.any_left = __builtin_amdgcn_ballot_w64(lm.left) != 0;
  auto patch = [&](Row<R>& q, const fvbc::RowData<R>& h, int r) {
    fvbc::Cell<R> c{q.q1, q.q2, q.q3};
    BC::patch(pc, min(max(r, 0), ny - 1), c, h);
    q.q1 = c.q1;
    q.q2 = c.q2;
    q.q3 = c.q3;
    const R ir = fvna::rcp(R(1) + q.q1);
    q.un = q.q3 * ir;
    q.ut = q.q2 * ir;
  };
  // source on row s (apps/na/source.h:86-104) from q_xy rows s-1, s, s+1,
  // then the store of row s (lanes 3..60)
  auto source_store = [&](int s, const XY<R>& lo, const XY<R>& mid, const XY<R>& hi) {
    const bool s_int = col_int && (s >= 2) && (s <= ny - 3);
    // u(i-1) + u(i+1): one cross-lane move, the other folds into the add
    const R uwe = shr1(mid.u) + shl1(mid.u);
    const R vwe = shr1(mid.v) + shl1(mid.v);
    // apps/na/source.h: x and y increments summed, then added to q once
    const R i2 = fvna::source_increment(ddx2, ddy2, mid.u, uwe, lo.u, hi.u);
    const R i3 = fvna::source_increment(ddx2, ddy2, mid.v, vwe, lo.v, hi.v);
    R r2 = fvna::apply_source(mid.q2, i2);
    R r3 = fvna::apply_source(mid.q3, i3);
    r2 = s_int ? r2 : mid.q2;
    r3 = s_int ? r3 : mid.q3;
    pin(r2);
    pin(r3);
    if (!kExt) {
      if (out_lane) {
        const uint32_t o = (uint32_t)s * row_bytes + kSz * (uint32_t)i;
        stg(o1, o, mid.q1);            // q1 passes the source unchanged
        stg(o2, o, r2);
        stg(o3, o, r3);
      }
    } else {
      // one cell of the output, at byte offset o, and of the history slot
      auto put = [&](uint32_t o) {
        stg(o1, o, mid.q1);
        stg(o2, o, r2);
        stg(o3, o, r3);
        if constexpr (sizeof(R) == 4) {
          if (hist_on) {
            stg(hs1, o, mid.q1);
            stg(hs2, o, r2);
            stg(hs3, o, r3);
          }
        }
      };
      if (out_lane && i >= sx && s >= sy)
        put((uint32_t)(s - sy) * row_bytes + kSz * (uint32_t)(i - sx));
      if (out_lane && (i + sx > nx - 1 || s + sy > ny - 1))
        put((uint32_t)s * row_bytes + kSz * (uint32_t)i);
      // data collection reads row jcut as it is before the shift
      if (stage_on && s == jcut && out_lane) stg(stage, kSz * (uint32_t)i, mid.q1);
    }
  };

  // dt/d on the lanes of interior columns, 0 on ghost columns
  const R dty_c = col_int ? dty : R(0), dtx_c = col_int ? dtx : R(0);
  R Sx = R(0), Sy = R(0);
  uint64_t err = 0;                          // lanes that met tmp3 < 0 (or NaN)
  const uint64_t in_mask = __builtin_amdgcn_ballot_w64(col_in);
  const uint64_t x_mask = __builtin_amdgcn_ballot_w64(col_x);

  // One march step, r <= y1 + 2.  bn/bm/bo: q_b rows r, r-1, r-2; fn/fc/fl:
  // y interfaces r, r-1, r-2.  xn/xo/xm hold q_xy rows r-5, r-4, r-3 on
  // entry; the x sweep overwrites xn with row r-2.  n1..n3 (nec, nes) hold the
  // raw prefetched row r on entry and row r+1 on exit.  Order: consume the
  // prefetch, store row r-4, issue the next prefetch -- so each store has a
  // whole step to drain before the next step waits for its prefetch (stores
  // and loads share one in-order counter).
  auto body = [&](int r, Row<R>& bn, Row<R>& bm, Row<R>& bo, YI<R>& fn, YI<R>& fc, YI<R>& fl,
                  XY<R>& xn, XY<R>& xm, XY<R>& xo) {
    // every load of the prefetch (and every store before it) is done: an
    // explicit vmcnt(0) here keeps the compiler's own wait for the prefetched
    // registers from landing after this step's stores
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    bn.q1 = n1; bn.q2 = n2; bn.q3 = n3;
    patch(bn, nh, r);
    if (r >= y0 + 4) source_store(r - 4, xn, xo, xm);
    // the next row, also past the item's last row: the row index is clamped
    // to the grid, and an unconditional load needs no copy of the registers
    // it replaces
    load_raw(r + 1, n1, n2, n3, nh);
    if (r < y0 - 2) return;
    const int m = r - 2;
    // CFL lane masks of rows r (y interfaces) and m (x interfaces); the index
    // is clamped so the loads are unconditional, the row test selects 0
    const uint64_t mYr = cm[min(max(r, y0), y1 - 1)];
    const uint64_t mXr = cm[min(max(m, y0), y1 - 1)];
    const uint64_t mY = (r >= y0 && r < y1 && r >= 1) ? mYr : 0;
    const uint64_t mX = (m >= y0 && m < y1) ? mXr : 0;

    // ---- y interface r (rows r-1 | r): normal = q3, tangential = q2 ----
    fvna::riemann(bm.q1, bm.q3, bm.q2, bm.un, bm.ut, bn.q1, bn.q3, bn.q2, bn.un, bn.ut,
                  hbeta, fn.f);
    // branch-free (a branch here would split the block and stop the
    // cross-lane moves of the x sweep from folding into their users)
    err |= __builtin_amdgcn_ballot_w64(fn.f.bad) & ((r >= 1 && r < ny) ? in_mask : 0);
    Sy = fvna::fmax2(Sy, lane_sel(mY, fn.f.smax));

    // W(r).W(r-1): the left product of interface r, the right one of r-1
    fn.dl[0] = fvna::dot1(fn.f.w, fc.f.w);
    fn.dl[1] = fvna::dot2(fn.f.w, fc.f.w);
    fn.dl[2] = fn.f.w.a3 * fc.f.w.a3;
    if (r < y0) return;

    // ---- y limiter at interface r-1 ----
    R M[3];
    fvna::limited_flux(fc.f, fc.dl[0], fn.dl[0], fc.dl[1], fn.dl[1], fc.dl[2], fn.dl[2], dty,
                       fc.P, M);
    if (r < y0 + 1) return;

    // ---- q_y on row m = r-2 (components: density, normal=q3, tangential=q2)
    const bool row_int = (m >= 2) && (m <= ny - 3);
    // ghost cells keep their value: dt/d is 0 there (the fluxes are finite
    // wherever the attempt is valid, so q - 0*flux == q exactly); the column
    // test is folded in once per item (dty_c, dtx_c), the row test here
    const R dtyu = row_int ? dty_c : R(0), dtxu = row_int ? dtx_c : R(0);
    const R y1q = bo.q1 - dtyu * (fl.P[0] + M[0]);
    const R y3q = bo.q3 - dtyu * (fl.P[1] + M[1]);
    const R y2q = bo.q2 - dtyu * (fl.P[2] + M[2]);

    // ---- x sweep on row m: normal = q2, tangential = q3 ----
    const R ir = fvna::rcp(R(1) + y1q);
    const R un = y2q * ir, ut = y3q * ir;
    const R l1 = shr1(y1q), l2 = shr1(y2q), l3 = shr1(y3q);
    const R lun = shr1(un), lut = shr1(ut);
    fvna::IfaceT<R> fx;
    fvna::riemann(l1, l2, l3, lun, lut, y1q, y2q, y3q, un, ut, hbeta, fx);
    err |= __builtin_amdgcn_ballot_w64(fx.bad) & ((m >= 0 && m < ny) ? x_mask : 0);
    Sx = fvna::fmax2(Sx, lane_sel(mX, fx.smax));
    // waves of the interface to the left (k-1)
    fvna::WavesT<R> wl;
    wl.a1 = shr1(fx.w.a1); wl.w21 = shr1(fx.w.w21); wl.w31 = shr1(fx.w.w31);
    wl.a2 = shr1(fx.w.a2); wl.w22 = shr1(fx.w.w22); wl.w32 = shr1(fx.w.w32);
    wl.a3 = shr1(fx.w.a3);
    const R dl1 = fvna::dot1(fx.w, wl), dl2 = fvna::dot2(fx.w, wl);
    const R dl3 = fx.w.a3 * wl.a3;
    // W(k).W(k+1) is the right neighbour's W(k+1).W(k)
    const R dr1 = shl1(dl1), dr2 = shl1(dl2), dr3 = shl1(dl3);
    R PX[3], MX[3];
    fvna::limited_flux(fx, dl1, dr1, dl2, dr2, dl3, dr3, dtx, PX, MX);
    const R mr0 = shl1(MX[0]), mr1 = shl1(MX[1]), mr2 = shl1(MX[2]);
    xn.q1 = y1q - dtxu * (mr0 + PX[0]);
    xn.q2 = y2q - dtxu * (mr1 + PX[1]);
    xn.q3 = y3q - dtxu * (mr2 + PX[2]);
    const R ir2 = fvna::rcp(R(1) + xn.q1);
    xn.u = xn.q2 * ir2;
    xn.v = xn.q3 * ir2;
  };

  // every window is written before it is read, except the waves of the
  // interface before the first one (their products with the first
  // interface's waves are formed but never used): only those are zeroed, so
  // no unset value is read and the other windows cost no initialisation
  Row<R> B0, B1, B2;
  YI<R> F0, F1, F2;
  XY<R> X0, X1, X2;
  F0.f.w = fvna::WavesT<R>{};
  for (int r = y0 - kHalo; r <= y1 + 2; r += 3) {
    body(r, B0, B2, B1, F0, F2, F1, X0, X2, X1);
    if (r + 1 > y1 + 2) break;
    body(r + 1, B1, B0, B2, F1, F0, F2, X1, X0, X2);
    if (r + 2 > y1 + 2) break;
    body(r + 2, B2, B1, B0, F2, F1, F0, X2, X1, X0);
  }
  // the last row, y1 - 1, with the window of the step r = y1 + 3 would see
  switch ((y1 - y0) % 3) {
    case 0: source_store(y1 - 1, X0, X1, X2); break;
    case 1: source_store(y1 - 1, X1, X2, X0); break;
    default: source_store(y1 - 1, X2, X0, X1); break;
  }

  // ---- CFL / error reduction ----
  // per-cell cfl = |s| * dt/d (fv.h:148-180); max over the wave's cells,
  // folded into the group's slot of this attempt (read by the next launch)
  const R cflw = wave_max(fmax_(Sy * dty, Sx * dtx));
  if (lane != 0) return;
  SyncSlots* sl = D.sync + phase % 3;
  if (err != 0) aadd_relaxed(&sl->err_flag, 1u);
  CflSlot<R>::fold_in(sl->grp[local % kGroups], cflw);
}

// Marker for a step kernel that carries the march of every registered app
// (apps/apps.h) and runs, per work item, that of its instance's boundary kind
// (the item carries it, so the choice is made from the item alone): launches
// over several instances.  A single-instance launch takes its app's policy
// directly, a kernel with one march.
struct AnyApp {};

// The step kernel: one wave per work item.
template <typename R, bool kOne, bool kExt, class BC>
__global__ __launch_bounds__(64) void fv_step_kernel(const InstDesc* __restrict__ insts,
                                                     const WorkItem* __restrict__ items,
                                                     const uint64_t* __restrict__ rowmask,
                                                     int n_items, int phase,
                                                     const Ctl* __restrict__ ctl_in,
                                                     const SyncSlots* __restrict__ sync_in) {
  // XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a contiguous
  // run of work items so vertically adjacent chunks of a strip (which share
  // halo rows) hit the same L2.
  int item;
  {
    const int b = blockIdx.x;
    const int q = n_items >> 3, rem = n_items & 7, x = b & 7, k = b >> 3;
    item = x * q + min(x, rem) + k;
  }
  const WorkItem it = items[item];
  if constexpr (!std::is_same<BC, AnyApp>::value) {
    step_impl<R, kOne, kExt, BC>(insts, it, rowmask, phase, ctl_in, sync_in);
    return;
  }
  switch (uni(it.bc)) {
#define FV_APP_CASE(P)                                                        \
  case P::kId:                                                                \
    if constexpr (kExt || !P::kExtOnly)                                       \
      step_impl<R, kOne, kExt, P>(insts, it, rowmask, phase, ctl_in, sync_in); \
    break;
    FIVONAGI_APPS(FV_APP_CASE)
#undef FV_APP_CASE
    default:
      break;
  }
}

// Control block after the last launched attempt (phase = index of the next
// attempt mod 6), stored where that attempt's launch will store it too; the
// host reads it after a synchronisation.  One block per instance.
__global__ __launch_bounds__(64) void fv_settle_kernel(const InstDesc* __restrict__ insts,
                                                       int phase) {
  const InstDesc& D = insts[blockIdx.x];
  const Ctl P = current(D, phase, true);
  if (threadIdx.x == 0) D.ctl[(phase + 1) & 1] = P;
}

// Data collection of attempt a, run after its step launch: if the attempt
// was accepted, record its new T and its cfl (row 0 and row 1, entry
// n - 1 - n0frame) and the app's measurement (collect_hifu, collect_taylor)
// into the instance's measurement grid (data_collect.h of either app).
// blockIdx.y = instance; threads cover the columns.
template <typename R>
__global__ __launch_bounds__(256) void fv_collect_kernel(const InstDesc* __restrict__ insts,
                                                         int phase) {
  const InstDesc& D = insts[blockIdx.y];
  typename Mem<R>::g* __restrict__ m = (typename Mem<R>::g*)D.measure;
  if (m == nullptr) return;
  // attempt a's control block (its step launch published it) and slots
  const Fold<R> F = fold_decide(fold_load<R>(D.ctl + ((phase + 1) & 1), D.sync + phase % 3),
                                D.cfl_max, D.dt_seq != nullptr);
  if (F.state != kFoldAccept) return;
  const Ctl& P = F.p;
  const double T = P.T + P.dt;                       // update_T
  const int nx = D.nx;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  // While T <= 0 the reference zeroes the whole grid every step and records
  // nothing (data_collect.h:22-30, both apps).  Here the grid is zeroed when
  // collection is switched on, and nothing writes it while T <= 0, so it is
  // already zero: the pass over the grid is skipped.
  if (T <= 0.0) return;
  const int64_t nf = P.n_acc - D.n0frame;          // n - 1 - n0frame, n = n_acc + 1
  if (nf >= nx) {                                   // simError 4
    if (tid == 0) aadd_relaxed(D.collect_err, 4u);
    return;
  }
  if (tid == 0) {                                   // T and cfl of the step
    m[nf] = R(T);
    m[nf + nx] = F.cfl;
  }
  if (D.bc == kBcTaylor)
    fvbc::collect_taylor<R>(D, P, T, (R*)m, tid);
  else
    fvbc::collect_hifu<R>(D, P, T, nf, (R*)m, tid);
}

// Closed-form ghost ring of the Taylor-shock instances (apps/taylor-angle-xy/
// boundary.h:38-55), written into the current input buffer before the step
// reads it.  The values depend on the cell and on T only, never on the state,
// so writing them in place is the reference's boundary kernel.  blockIdx.y =
// instance; the ring (2 cells wide) is enumerated by a flat index.
__global__ __launch_bounds__(256) void fv_taylor_ghost_kernel(const InstDesc* __restrict__ insts,
                                                             int phase) {
  const InstDesc& D = insts[blockIdx.y];
  if (D.bc != kBcTaylor) return;
  const Ctl P = current(D, phase, false);
  if (P.err != 0 || P.n_acc >= P.target_acc || !(P.T < P.T_stop)) return;
  const int nx = D.nx, ny = D.ny;
  const int ring = 4 * nx + 4 * (ny - 4);
  const double T = P.T;
  const int cur = P.cur;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ring; k += gridDim.x * blockDim.x) {
    int i, j;
    if (k < 4 * nx) {            // rows 0, 1, ny-2, ny-1
      const int rr = k / nx;
      i = k - rr * nx;
      j = rr < 2 ? rr : ny - 4 + rr;
    } else {                     // columns 0, 1, nx-2, nx-1 of rows 2..ny-3
      const int kk = k - 4 * nx;
      const int cc = kk / (ny - 4);
      j = 2 + (kk - cc * (ny - 4));
      i = cc < 2 ? cc : nx - 4 + cc;
    }
    // the closed form is taken at the cell's position in the unmoved frame
    // (apps/taylor-angle-xy/boundary.h:45-49: i + MDX, j + MDY)
    const size_t o = (size_t)j * (size_t)nx + (size_t)i;
    if (D.fp64) {
      fvbc::Cell<double> c;
      fvbc::Taylor::prepass(D, T, i + P.mdx, j + P.mdy, c);
      ((double*)D.q[cur][0])[o] = c.q1;
      ((double*)D.q[cur][1])[o] = c.q2;
      ((double*)D.q[cur][2])[o] = c.q3;
    } else {
      fvbc::Cell<float> c;
      fvbc::Taylor::prepass(D, T, i + P.mdx, j + P.mdy, c);
      D.q[cur][0][o] = c.q1;
      D.q[cur][1][o] = c.q2;
      D.q[cur][2][o] = c.q3;
    }
  }
}

}  // namespace fv

extern "C" hipError_t fv_launch_taylor_ghosts(const fv::InstDesc* insts, int n_inst,
                                              int max_ring, int phase, hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  const int blocks = (max_ring + 255) / 256;
  hipLaunchKernelGGL(fv::fv_taylor_ghost_kernel, dim3(blocks, n_inst), dim3(256), 0, stream,
                     insts, phase);
  return hipGetLastError();
}

extern "C" hipError_t fv_launch_collect(const fv::InstDesc* insts, int n_inst, int max_nx,
                                        int phase, int fp64, hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  if (fp64)
    hipLaunchKernelGGL(fv::fv_collect_kernel<double>, dim3((max_nx + 255) / 256, n_inst),
                       dim3(256), 0, stream, insts, phase);
  else
    hipLaunchKernelGGL(fv::fv_collect_kernel<float>, dim3((max_nx + 255) / 256, n_inst),
                       dim3(256), 0, stream, insts, phase);
  return hipGetLastError();
}

extern "C" hipError_t fv_launch_settle(const fv::InstDesc* insts, int n_inst, int phase,
                                       hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  hipLaunchKernelGGL(fv::fv_settle_kernel, dim3(n_inst), dim3(64), 0, stream, insts, phase);
  return hipGetLastError();
}

// Resident step waves per CU (occupancy of the step kernel on this device),
// of the FP64 kernel if fp64.  The lowest over the plain and extended
// kernels of several instances (every app's march: the most registers), so
// the split fits whichever kernel a context launches.
extern "C" hipError_t fv_step_waves_per_cu(int* waves, int fp64) {
  int a = 0, b = 0;
  hipError_t e = fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &a, fv::fv_step_kernel<double, false, false, fv::AnyApp>, fv::kLanes, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &a, fv::fv_step_kernel<float, false, false, fv::AnyApp>, fv::kLanes, 0);
  if (e != hipSuccess) return e;
  e = fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &b, fv::fv_step_kernel<double, false, true, fv::AnyApp>, fv::kLanes, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &b, fv::fv_step_kernel<float, false, true, fv::AnyApp>, fv::kLanes, 0);
  *waves = a < b ? a : b;
  return e;
}

namespace {
template <typename R, bool kOne, bool kExt, class BC>
void launch_step_t(const fv::InstDesc* insts, const fv::WorkItem* items, const uint64_t* rowmask,
                   int n_items, int phase, const fv::Ctl* c_in, const fv::SyncSlots* s_in,
                   hipStream_t stream) {
  hipLaunchKernelGGL((fv::fv_step_kernel<R, kOne, kExt, BC>), dim3(n_items), dim3(fv::kLanes), 0,
                     stream, insts, items, rowmask, n_items, phase, c_in, s_in);
}
// a single instance: its app's own kernel (bc), else the dispatching one
template <typename R, bool kExt>
void launch_one(int bc, const fv::InstDesc* insts, const fv::WorkItem* items,
                const uint64_t* rowmask, int n_items, int phase, const fv::Ctl* c_in,
                const fv::SyncSlots* s_in, hipStream_t stream) {
  switch (bc) {
#define FV_APP_LAUNCH(P)                                                                 \
  case P::kId:                                                                           \
    if constexpr (kExt || !P::kExtOnly) {                                                \
      launch_step_t<R, true, kExt, P>(insts, items, rowmask, n_items, phase, c_in, s_in, \
                                      stream);                                           \
      return;                                                                            \
    }                                                                                    \
    break;
    FIVONAGI_APPS(FV_APP_LAUNCH)
#undef FV_APP_LAUNCH
    default:
      break;
  }
  launch_step_t<R, true, kExt, fv::AnyApp>(insts, items, rowmask, n_items, phase, c_in, s_in,
                                           stream);
}
template <typename R>
void launch_step_r(bool one, bool ext, int bc, const fv::InstDesc* insts,
                   const fv::WorkItem* items, const uint64_t* rowmask, int n_items, int phase,
                   const fv::Ctl* c_in, const fv::SyncSlots* s_in, hipStream_t stream) {
  if (one && !ext)
    launch_one<R, false>(bc, insts, items, rowmask, n_items, phase, c_in, s_in, stream);
  else if (one)
    launch_one<R, true>(bc, insts, items, rowmask, n_items, phase, c_in, s_in, stream);
  else if (!ext)
    launch_step_t<R, false, false, fv::AnyApp>(insts, items, rowmask, n_items, phase, c_in, s_in,
                                               stream);
  else
    launch_step_t<R, false, true, fv::AnyApp>(insts, items, rowmask, n_items, phase, c_in, s_in,
                                              stream);
}
}  // namespace

// One step attempt of every instance.  ctl0/sync0: the control blocks and
// slot sets of the only instance when n_inst == 1 (null otherwise), bc0 its
// boundary kind; ext: some instance has a moving window, the HIFU boundary or
// data collection; fp64: the instances are FP64.
extern "C" hipError_t fv_launch_step(const fv::InstDesc* insts, const fv::WorkItem* items,
                                     const uint64_t* rowmask, int n_items, int phase,
                                     int n_inst, fv::Ctl* ctl0, fv::SyncSlots* sync0, int bc0,
                                     int ext, int fp64, hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  const bool one = n_inst == 1 && ctl0 && sync0;
  const fv::Ctl* c_in = one ? ctl0 + (phase & 1) : nullptr;
  const fv::SyncSlots* s_in = one ? sync0 + (phase + 2) % 3 : nullptr;
  if (fp64)
    launch_step_r<double>(one, ext != 0, bc0, insts, items, rowmask, n_items, phase, c_in, s_in,
                          stream);
  else
    launch_step_r<float>(one, ext != 0, bc0, insts, items, rowmask, n_items, phase, c_in, s_in,
                         stream);
  return hipGetLastError();
}
