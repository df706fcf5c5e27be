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

#include "../apps/common/boundary.h"
#include "../apps/na/solver.h"
#include "fv_types.h"

namespace fv {

typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float cgfloat;
typedef __attribute__((address_space(4))) const float kfloat;   // read-only: scalar loads

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

// v on the lanes whose bit of the uniform 64-bit mask is set, 0 elsewhere
// (one v_cndmask with the mask as the lane condition).
__device__ __forceinline__ float lane_sel(uint64_t mask, float v) {
  float d;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(d) : "v"(v), "s"(mask));
  return d;
}

// Pins a value in a register here: its computation cannot sink into a
// branch that follows, so a cross-lane move feeding it stays in its block
// and folds into its user.
__device__ __forceinline__ void pin(float& v) { asm volatile("" : "+v"(v)); }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ uint32_t amax_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t aadd_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t axchg_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// What attempt a's reduction slots say about it, with the control block it
// ran with.  fold_load issues every load before anything is decided, so the
// fold costs one memory round trip.
struct FoldRaw {
  Ctl p;          // control block attempt a ran with
  uint32_t ran, errf, gbits;
};
struct Fold {
  Ctl p;
  float cfl;      // max cfl over the counted cells of attempt a
  int state;      // kFoldIdle, kFoldError, kFoldAccept, kFoldReject
};
constexpr int kFoldIdle = 0, kFoldError = 1, kFoldAccept = 2, kFoldReject = 3;

__device__ __forceinline__ FoldRaw fold_load(const Ctl* __restrict__ c,
                                             const SyncSlots* __restrict__ sl) {
  FoldRaw r;
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
  r.gbits = sl->grp[__lane_id() & (kGroups - 1)].cfl_bits;
  return r;
}

__device__ __forceinline__ Fold fold_decide(const FoldRaw& r, double cfl_max, bool presc) {
  Fold f;
  f.p = r.p;
  f.p.n_acc = uni64(r.p.n_acc); f.p.n_rej = uni64(r.p.n_rej); f.p.n_att = uni64(r.p.n_att);
  f.p.target_acc = uni64(r.p.target_acc);
  f.p.cur = uni(r.p.cur); f.p.err = uni(r.p.err);
  f.p.mdx = uni(r.p.mdx); f.p.mdy = uni(r.p.mdy);
  // non-negative floats order like their bit patterns
  uint32_t mbits = 0u;
#pragma unroll
  for (int k = 0; k < kGroups; ++k) mbits = max(mbits, (uint32_t)__builtin_amdgcn_readlane(r.gbits, k));
  const float cfl32 = __uint_as_float(mbits);
  f.cfl = cfl32;
  const bool acc = presc || ((double)cfl32 < cfl_max);
  f.state = !uni(r.ran) ? kFoldIdle
                        : (uni(r.errf) ? kFoldError : (acc ? kFoldAccept : kFoldReject));
  return f;
}

// Ping-pong index attempt a+1 reads (known before the rest of advance()).
__device__ __forceinline__ int next_cur(const Fold& f) {
  return f.state == kFoldAccept ? (f.p.cur ^ 1) : f.p.cur;
}

// Moving window (common/main.cu:109-135, :167-176): once T >= MDT the window
// sits int((T - MDT) * MDV / d) cells from its start; after an accepted step
// ending at T it moves by the part not applied yet (md = the current offset).
__device__ __forceinline__ int md_shift(double T, double md_t, double v, double d, int md) {
  return T >= md_t ? (int)((T - md_t) * v / d) - md : 0;
}

// Control block of attempt a+1 from attempt a's fold: common/main.cu:199-230
// with a clean retry on rejection.  log: record dt and T of an accepted step
// (adjoint runs); every writer records the same values.
__device__ Ctl advance(const InstDesc& D, const Fold& f, bool log) {
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
    if (log && D.dt_log) {
      const int64_t idx = p.n_acc - D.log_base;
      if (idx >= 0 && idx < D.log_len) {
        D.dt_log[idx] = p.dt;
        D.T_log[idx] = p.T;
      }
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
    ndt = (att < D.dt_seq_len) ? D.dt_seq[att] : p.dt;
    if (att >= D.dt_seq_len) o.err = 9;   // ran past the prescribed sequence
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
  float fs, fc;
  sincospif((float)(2.0 * (turns - floor(turns))), &fs, &fc);
  o.drive_s = fs;
  o.drive_c = fc;
  return o;
}

// Control block of the attempt a launch runs (phase = a mod 6), read-only.
__device__ __forceinline__ Ctl current(const InstDesc& D, int phase, bool log) {
  return advance(D,
                 fold_decide(fold_load(D.ctl + (phase & 1), D.sync + (phase + 2) % 3),
                             D.cfl_max, D.dt_seq != nullptr),
                 log);
}

struct Row {         // one cell of a row as the y sweep sees it
  float q1, q2, q3;  // rho', rho u, rho v
  float un, ut;      // y normal velocity v = q3/rho, tangential u = q2/rho
};

struct YI {          // y interface with its limited flux towards the + cell
  fvna::Iface f;
  float dl[3];       // W_p . W_p of the interface below (one per wave family)
  float P[3];        // apdq - fadd, in (density, normal, tangential) order
};

struct XY {          // one cell of q_xy, with its velocities for the source
  float q1, q2, q3, u, v;
};

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
//  - HIFU boundary (apps/hifu-beam/boundary.h:20-133): rows 0, 1 read rows
//    4, 3 with rho v negated, the 2x2 corners are 0, and the aperture rows of
//    the left ghost columns take the focused packet while MDX == 0.
template <bool kOne, bool kExt>
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
  const int bc = D.bc;
  float* qb[2][3];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int f = 0; f < 3; ++f) qb[h][f] = D.q[h][f];
  double cfl_max = D.cfl_max;
  const double* dt_seq = D.dt_seq;
  const float* ec_p = D.edge_c;
  const float* es_p = D.edge_s;
  const FoldRaw FR = fold_load(kOne ? ctl_in : D.ctl + (phase & 1),
                               kOne ? sync_in : D.sync + (phase + 2) % 3);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int f = 0; f < 3; ++f) asm volatile("" : "+s"(qb[h][f]));
  asm volatile("" : "+s"(cfl_max), "+s"(dt_seq), "+s"(ec_p), "+s"(es_p));
  const Fold F = fold_decide(FR, cfl_max, dt_seq != nullptr);
  const int cur = next_cur(F);

  cgfloat* __restrict__ in1 = (cgfloat*)(cur ? qb[1][0] : qb[0][0]);
  cgfloat* __restrict__ in2 = (cgfloat*)(cur ? qb[1][1] : qb[0][1]);
  cgfloat* __restrict__ in3 = (cgfloat*)(cur ? qb[1][2] : qb[0][2]);
  gfloat* __restrict__ o1 = (gfloat*)(cur ? qb[0][0] : qb[1][0]);
  gfloat* __restrict__ o2 = (gfloat*)(cur ? qb[0][1] : qb[1][1]);
  gfloat* __restrict__ o3 = (gfloat*)(cur ? qb[0][2] : qb[1][2]);
  kfloat* __restrict__ edge_c = (kfloat*)ec_p;
  kfloat* __restrict__ edge_s = (kfloat*)es_p;

  const int i = it.x0 - kHalo + lane;
  const bool col_in = (i >= 0) && (i < nx);
  const bool col_x = col_in && (i >= 1);     // lane holds an x interface of the grid
  const int ic = min(max(i, 0), nx - 1);
  const bool col_int = (i >= 2) && (i <= nx - 3);
  const bool out_lane = (lane >= kHalo) && (lane < kLanes - kHalo) && (i < nx);
  const int y0 = it.y0, y1 = it.y1;
  // CFL lane masks of the item's rows: cm[j], j in [y0, y1)
  const uint64_t* __restrict__ cm = rowmask + (it.mrow - y0);

  // Ghost-cell source columns (apps/common/boundary.h): colA on rows
  // 2..ny-3, colB on ghost rows (top/bottom rules beat the right rule; the
  // left rule beats everything and keeps its own row).  Taylor ghost cells
  // were written into the input by fv_taylor_ghost_kernel: plain copies here.
  const bool taylor = (bc == kBcTaylor);
  const bool left = !taylor && (ic <= 1);
  const int colA = taylor ? ic : (left ? 2 : (ic >= nx - 2 ? nx - 3 : ic));
  const int colB = taylor ? ic : (left ? 2 : ic);
  const uint32_t bA = 4u * (uint32_t)colA, bB = 4u * (uint32_t)colB;
  const uint32_t row_bytes = 4u * (uint32_t)nx;
  const bool hifu = kExt && bc == kBcHifu;
  kfloat* __restrict__ htab = (kfloat*)D.hifu_tab;
  struct HRow { float ph, nux, nuy, on; };   // aperture row data (hifu)

  // raw (unpatched) load of row r, with the row's drive coefficients; the
  // boundary patch is applied when the row is consumed one iteration later,
  // so the load latency overlaps a full step
  auto load_raw = [&](int r, float& a, float& b, float& c, float& ec, float& es, HRow& h) {
    const int rc = min(max(r, 0), ny - 1);
    const bool grow = (rc <= 1) || (rc >= ny - 2);
    uint32_t o;
    if (hifu) {
      // bottom rows mirror rows 4, 3; top rows copy ny-3; columns as colA
      const int jm = rc <= 1 ? 4 - rc : (rc >= ny - 2 ? ny - 3 : rc);
      o = (uint32_t)jm * row_bytes + bA;
    } else {
      const int jm = taylor ? rc : (rc >= ny - 2 ? ny - 3 : (rc <= 1 ? 2 : rc));
      const uint32_t rowL = (uint32_t)rc * row_bytes, rowM = (uint32_t)jm * row_bytes;
      if (grow) o = (left ? rowL : rowM) + bB;   // ghost rows
      else o = rowM + bA;                        // rows 2..ny-3: rc == jm
    }
    a = ldg(in1, o);
    b = ldg(in2, o);
    c = ldg(in3, o);
    if (bc == kBcEdgeDrive || hifu) {
      ec = edge_c[rc];
      es = edge_s[rc];
    }
    if (hifu) {
      h.ph = htab[4 * rc];
      h.nux = htab[4 * rc + 1];
      h.nuy = htab[4 * rc + 2];
      h.on = htab[4 * rc + 3];
    }
  };
  float n1, n2, n3, nec = 0.0f, nes = 0.0f;
  HRow nh{0.0f, 0.0f, 0.0f, 0.0f};
  load_raw(it.y0 - kHalo, n1, n2, n3, nec, nes, nh);   // in flight during advance()

  // ---- control block of this attempt ----
  const Ctl P = advance(D, F, local == 0);
  const bool runs = P.err == 0 && P.n_acc < P.target_acc && P.T < P.T_stop;
  if (local == 0 && lane == 0) {
    // one wave per instance: publish the control block for the next launch,
    // clear the slots the next attempt reduces into, mark this attempt
    D.ctl[(phase + 1) & 1] = P;
    SyncSlots* nx_sl = D.sync + (phase + 1) % 3;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) nx_sl->grp[k].cfl_bits = 0u;
    nx_sl->err_flag = 0u;
    nx_sl->ran = 0u;
    if (runs) D.sync[phase % 3].ran = 1u;
  }
  if (!runs) return;
  const float dtx = unif(P.dtx), dty = unif(P.dty);
  const float ddx2 = unif(P.ddx2), ddy2 = unif(P.ddy2);
  const float sin_t = unif((float)P.drive_s), cos_t = unif((float)P.drive_c);
  const float hbeta = 0.5f * D.beta2m2;     // BETA - 1
  // the edge source runs while the window has not moved in x
  // (apps/hifu-beam/boundary.h:82); after that the left ghost columns are
  // plain zero-gradient copies
  const bool drive = (bc == kBcEdgeDrive) && (!kExt || P.mdx == 0);
  const bool drive_lane = left && drive;
  // shift of the window after this attempt, if it is accepted
  int sx = 0, sy = 0;
  if (kExt && D.md_on) {
    const double Tn = P.T + P.dt;
    sx = uni(md_shift(Tn, D.md_t, D.md_vx, D.dx, P.mdx));
    sy = uni(md_shift(Tn, D.md_t, D.md_vy, D.dy, P.mdy));
  }
  gfloat* __restrict__ stage = (gfloat*)D.stage;
  const bool stage_on = kExt && stage != nullptr;
  const int jcut = D.jcut;
  // HIFU: the packet runs while the window has not moved (boundary.h:82)
  const bool hdrive = hifu && P.mdx == 0 && it.x0 == 0;
  const double Td = P.T;
  const float hk = D.hifu_k;
  auto patch = [&](Row& q, float ec, float es, const HRow& h, int r) {
    if (drive) {
      float a, b, c;
      fvbc::drive(ec, es, sin_t, cos_t, a, b, c);
      q.q1 = drive_lane ? a : q.q1;
      q.q2 = drive_lane ? b : q.q2;
      q.q3 = drive_lane ? c : q.q3;
    }
    if (hifu) {
      const int rc = min(max(r, 0), ny - 1);
      q.q3 = rc <= 1 ? -q.q3 : q.q3;          // reflective beam axis
      if (hdrive && h.on != 0.0f && rc >= 2 && rc <= ny - 3) {
        float a, b, c;
        // T + phi in FP64: the envelope's 10th power magnifies its rounding
        fvbc::hifu_packet(ec, es, sin_t, cos_t, hk * (float)(Td + (double)h.ph), h.nux, h.nuy,
                          a, b, c);
        q.q1 = left ? a : q.q1;
        q.q2 = left ? b : q.q2;
        q.q3 = left ? c : q.q3;
      }
      const bool corner = (rc <= 1 || rc >= ny - 2) && (ic <= 1 || ic >= nx - 2);
      q.q1 = corner ? 0.0f : q.q1;
      q.q2 = corner ? 0.0f : q.q2;
      q.q3 = corner ? 0.0f : q.q3;
    }
    const float ir = fvna::rcp(1.0f + q.q1);
    q.un = q.q3 * ir;
    q.ut = q.q2 * ir;
  };
  // source on row s (apps/na/source.h:86-104) from q_xy rows s-1, s, s+1,
  // then the store of row s (lanes 3..60)
  auto source_store = [&](int s, const XY& lo, const XY& mid, const XY& hi) {
    const bool s_int = col_int && (s >= 2) && (s <= ny - 3);
    // u(i-1) + u(i+1): one cross-lane move, the other folds into the add
    const float uwe = shr1(mid.u) + shl1(mid.u);
    const float vwe = shr1(mid.v) + shl1(mid.v);
    float r2 = (mid.q2 + ddx2 * fmaf(-2.0f, mid.u, uwe))
               + ddy2 * (fmaf(-2.0f, mid.u, lo.u) + hi.u);
    float r3 = (mid.q3 + ddx2 * fmaf(-2.0f, mid.v, vwe))
               + ddy2 * (fmaf(-2.0f, mid.v, lo.v) + hi.v);
    r2 = s_int ? r2 : mid.q2;
    r3 = s_int ? r3 : mid.q3;
    pin(r2);
    pin(r3);
    if (!kExt) {
      if (out_lane) {
        const uint32_t o = (uint32_t)s * row_bytes + 4u * (uint32_t)i;
        stg(o1, o, mid.q1);            // q1 passes the source unchanged
        stg(o2, o, r2);
        stg(o3, o, r3);
      }
    } else {
      if (out_lane && i >= sx && s >= sy) {
        const uint32_t o = (uint32_t)(s - sy) * row_bytes + 4u * (uint32_t)(i - sx);
        stg(o1, o, mid.q1);
        stg(o2, o, r2);
        stg(o3, o, r3);
      }
      if (out_lane && (i + sx > nx - 1 || s + sy > ny - 1)) {
        const uint32_t o = (uint32_t)s * row_bytes + 4u * (uint32_t)i;
        stg(o1, o, mid.q1);
        stg(o2, o, r2);
        stg(o3, o, r3);
      }
      // data collection reads row jcut as it is before the shift
      if (stage_on && s == jcut && out_lane) stg(stage, 4u * (uint32_t)i, mid.q1);
    }
  };

  float Sx = 0.0f, Sy = 0.0f;
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
  auto body = [&](int r, Row& bn, Row& bm, Row& bo, YI& fn, YI& fc, YI& fl, XY& xn, XY& xm,
                  XY& xo) {
    // every load of the prefetch (and every store before it) is done: an
    // explicit vmcnt(0) here keeps the compiler's own wait for the prefetched
    // registers from landing after this step's stores
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    bn.q1 = n1; bn.q2 = n2; bn.q3 = n3;
    patch(bn, nec, nes, nh, r);
    if (r >= y0 + 4) source_store(r - 4, xn, xo, xm);
    if (r + 1 <= y1 + 2) load_raw(r + 1, n1, n2, n3, nec, nes, nh);
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
    float M[3];
    fvna::limited_flux(fc.f, fc.dl[0], fn.dl[0], fc.dl[1], fn.dl[1], fc.dl[2], fn.dl[2], dty,
                       fc.P, M);
    if (r < y0 + 1) return;

    // ---- q_y on row m = r-2 (components: density, normal=q3, tangential=q2)
    const bool row_int = (m >= 2) && (m <= ny - 3);
    const bool upd = col_int && row_int;
    // ghost cells keep their value: dt/d is 0 there (the fluxes are finite
    // wherever the attempt is valid, so q - 0*flux == q exactly)
    const float dtyu = upd ? dty : 0.0f, dtxu = upd ? dtx : 0.0f;
    const float y1q = bo.q1 - dtyu * (fl.P[0] + M[0]);
    const float y3q = bo.q3 - dtyu * (fl.P[1] + M[1]);
    const float y2q = bo.q2 - dtyu * (fl.P[2] + M[2]);

    // ---- x sweep on row m: normal = q2, tangential = q3 ----
    const float ir = fvna::rcp(1.0f + y1q);
    const float un = y2q * ir, ut = y3q * ir;
    const float l1 = shr1(y1q), l2 = shr1(y2q), l3 = shr1(y3q);
    const float lun = shr1(un), lut = shr1(ut);
    fvna::Iface fx;
    fvna::riemann(l1, l2, l3, lun, lut, y1q, y2q, y3q, un, ut, hbeta, fx);
    err |= __builtin_amdgcn_ballot_w64(fx.bad) & ((m >= 0 && m < ny) ? x_mask : 0);
    Sx = fvna::fmax2(Sx, lane_sel(mX, fx.smax));
    // waves of the interface to the left (k-1)
    fvna::Waves wl;
    wl.a1 = shr1(fx.w.a1); wl.w21 = shr1(fx.w.w21); wl.w31 = shr1(fx.w.w31);
    wl.a2 = shr1(fx.w.a2); wl.w22 = shr1(fx.w.w22); wl.w32 = shr1(fx.w.w32);
    wl.a3 = shr1(fx.w.a3);
    const float dl1 = fvna::dot1(fx.w, wl), dl2 = fvna::dot2(fx.w, wl);
    const float dl3 = fx.w.a3 * wl.a3;
    // W(k).W(k+1) is the right neighbour's W(k+1).W(k)
    const float dr1 = shl1(dl1), dr2 = shl1(dl2), dr3 = shl1(dl3);
    float PX[3], MX[3];
    fvna::limited_flux(fx, dl1, dr1, dl2, dr2, dl3, dr3, dtx, PX, MX);
    const float mr0 = shl1(MX[0]), mr1 = shl1(MX[1]), mr2 = shl1(MX[2]);
    xn.q1 = y1q - dtxu * (mr0 + PX[0]);
    xn.q2 = y2q - dtxu * (mr1 + PX[1]);
    xn.q3 = y3q - dtxu * (mr2 + PX[2]);
    const float ir2 = fvna::rcp(1.0f + xn.q1);
    xn.u = xn.q2 * ir2;
    xn.v = xn.q3 * ir2;
  };

  Row B0{}, B1{}, B2{};
  YI F0{}, F1{}, F2{};
  XY X0{}, X1{}, X2{};
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
  const float cflw = wave_max(fmaxf(Sy * dty, Sx * dtx));
  if (lane != 0) return;
  SyncSlots* sl = D.sync + phase % 3;
  if (err != 0) aadd_relaxed(&sl->err_flag, 1u);
  amax_relaxed(&sl->grp[local % kGroups].cfl_bits, __float_as_uint(cflw));
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

// Data collection of attempt a (apps/hifu-beam/data_collect.h:7-107), run
// after its step launch: if the attempt was accepted, record its new T, its
// cfl and row jcut of its new state (staged by the step before the window
// shift, with the pre-shift MDX) into the instance's measurement grid.
// blockIdx.y = instance; threads cover the columns (and, while T <= 0, the
// whole grid, which the reference zeroes then).
__global__ __launch_bounds__(256) void fv_collect_kernel(const InstDesc* __restrict__ insts,
                                                         int phase) {
  const InstDesc& D = insts[blockIdx.y];
  float* __restrict__ m = D.measure;
  if (m == nullptr) return;
  // attempt a's control block (its step launch published it) and slots
  const Fold F = fold_decide(fold_load(D.ctl + ((phase + 1) & 1), D.sync + phase % 3),
                             D.cfl_max, D.dt_seq != nullptr);
  if (F.state != kFoldAccept) return;
  const Ctl& P = F.p;
  const double T = P.T + P.dt;                       // update_T
  const int nx = D.nx, ny = D.ny;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (T <= 0.0) {                                    // data_collect.h:27-30
    for (int64_t k = tid; k < (int64_t)nx * ny; k += (int64_t)gridDim.x * blockDim.x)
      m[k] = 0.0f;
    return;
  }
  const int i = tid;
  if (i >= nx) return;
  const float u1 = D.stage[i];
  const int mdx = P.mdx;
  const int64_t nf = P.n_acc - D.n0frame;          // n - 1 - n0frame, n = n_acc + 1
  if (nf >= nx) {                                   // data_collect.h:96
    if (i == 0) aadd_relaxed(D.collect_err, 4u);
    return;
  }
  for (int k = 0; k < 4; ++k) {                     // data_collect.h:51-76
    if (i + mdx == D.mic[k]) {
      float* row = m + (size_t)(2 + k) * nx;
      if (nf == 0)
        for (int ii = 0; ii < nx; ++ii) row[ii] = 0.0f;
      row[nf] = u1;
      break;
    }
  }
  if (i == 0) {                                     // data_collect.h:78-94
    m[nf] = (float)T;
    m[nf + nx] = F.cfl;
    if (nf == 0)
      for (int k = 0; k < 4; ++k)
        if (D.mic[k] - mdx < 0 || D.mic[k] - mdx >= nx)
          for (int ii = 0; ii < nx; ++ii) m[(size_t)(2 + k) * nx + ii] = 0.0f;
  }
  // central peak only (data_collect.h:100-106)
  if (fabs((double)(i + mdx) / D.eta - T) > 0.5) return;
  const int64_t o = (int64_t)i + mdx + 6 * (int64_t)nx;
  if (o < (int64_t)nx * ny) m[o] = fmaxf(u1, m[o]);
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
    float a, b, c;
    // the closed form is taken at the cell's position in the unmoved frame
    // (apps/taylor-angle-xy/boundary.h:45-49: i + MDX, j + MDY)
    fvbc::taylor(D.tay_rhoa, D.tay_c, D.tay_cos, D.tay_sin, D.tay_yoff, D.tay_jc, D.dx, D.dy,
                 T, i + P.mdx, j + P.mdy, a, b, c);
    const size_t o = (size_t)j * (size_t)nx + (size_t)i;
    D.q[cur][0][o] = a;
    D.q[cur][1][o] = b;
    D.q[cur][2][o] = c;
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
                                        int phase, hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  hipLaunchKernelGGL(fv::fv_collect_kernel, dim3((max_nx + 255) / 256, n_inst), dim3(256), 0,
                     stream, insts, phase);
  return hipGetLastError();
}

extern "C" hipError_t fv_launch_settle(const fv::InstDesc* insts, int n_inst, int phase,
                                       hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  hipLaunchKernelGGL(fv::fv_settle_kernel, dim3(n_inst), dim3(64), 0, stream, insts, phase);
  return hipGetLastError();
}

// Resident step waves per CU (occupancy of the step kernel on this device).
extern "C" hipError_t fv_step_waves_per_cu(int* waves) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      waves, fv::fv_step_kernel<false, false>, fv::kLanes, 0);
}

// One step attempt of every instance.  ctl0/sync0: the control blocks and
// slot sets of the only instance when n_inst == 1 (null otherwise); md: some
// instance has a moving window.
extern "C" hipError_t fv_launch_step(const fv::InstDesc* insts, const fv::WorkItem* items,
                                     const uint64_t* rowmask, int n_items, int phase,
                                     int n_inst, fv::Ctl* ctl0, fv::SyncSlots* sync0, int md,
                                     hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  const bool one = n_inst == 1 && ctl0 && sync0;
  const fv::Ctl* c_in = one ? ctl0 + (phase & 1) : nullptr;
  const fv::SyncSlots* s_in = one ? sync0 + (phase + 2) % 3 : nullptr;
  if (one && !md)
    hipLaunchKernelGGL((fv::fv_step_kernel<true, false>), dim3(n_items), dim3(fv::kLanes), 0,
                       stream, insts, items, rowmask, n_items, phase, c_in, s_in);
  else if (one)
    hipLaunchKernelGGL((fv::fv_step_kernel<true, true>), dim3(n_items), dim3(fv::kLanes), 0,
                       stream, insts, items, rowmask, n_items, phase, c_in, s_in);
  else if (!md)
    hipLaunchKernelGGL((fv::fv_step_kernel<false, false>), dim3(n_items), dim3(fv::kLanes), 0,
                       stream, insts, items, rowmask, n_items, phase, c_in, s_in);
  else
    hipLaunchKernelGGL((fv::fv_step_kernel<false, true>), dim3(n_items), dim3(fv::kLanes), 0,
                       stream, insts, items, rowmask, n_items, phase, c_in, s_in);
  return hipGetLastError();
}
