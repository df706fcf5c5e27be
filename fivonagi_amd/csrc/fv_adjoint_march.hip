// fv_adjoint_march.hip -- the reverse step of the discrete adjoint as one
// march kernel (gfx950).
//
// The reference has no adjoint.  The step it differentiates (reference:
// common/main.cu:310-635, apps/na/fv.h, apps/na/source.h, common/boundary.h)
// is the one csrc/fv_step.hip runs:
//   q_b = B(q_n), q_y = Y(q_b), q_xy = X(q_y), q_{n+1} = S(q_xy)
// with J = sum_n dt_n sum_box rho'_{n+1}^2 (dt frozen).  One launch of
// adj_reverse_kernel takes lam_{n+1} (the adjoint of q_{n+1}) to
//   lam_n = B^T Y^T X^T S^T (lam_{n+1} + dJ_n/dq_{n+1})
// and adds step n's share to the drive gradient and to J.
//
// Each wave marches up a strip of 56 output columns (4-column halo on each
// side), like the step kernel: per row it recomputes the forward values of
// q_n with the step kernel's own functions (the app's boundary patch, the
// solver.h Riemann solve, limiter and fluxes, and the control block the
// forward step ran with: its dt/dx, dt/dy and drive phase as the step used
// them), so every branch it differentiates is the one the forward step took,
// and applies the transposes in the same pass:
//   row m = r - 2:  q_y(m), the x sweep of m, S^T and X^T of m -> lam_y(m)
//   interface m:    the y limiter's transpose (lam_y of m and m-1)
//   interface m-1:  the Riemann transpose (its wave adjoint is complete once
//                   the limiters of m-2, m-1 and m are done) -> lam_b(m-2)
//   row m-2:        B^T -> lam_n(m-2)
// x neighbours come from DPP lane shifts, y neighbours from register windows.
// A wave runs one or more work items (row ranges of a strip) in turn; the
// host groups them into waves of about equal cost.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../apps/apps.h"
#include "../apps/na/adjoint.h"
#include "../apps/na/source.h"
#include "fv_adjoint.h"
#include "fv_types.h"

namespace fv {
namespace adjm {

constexpr int kLanes = 64;
constexpr int kHalo = 4;                       // x halo of the reverse march, each side
constexpr int kCols = kLanes - 2 * kHalo;      // 56 output columns per strip
static_assert(kCols == 56, "the host cuts the reverse march's strips at 56 columns (fv_tables.cpp)");

typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float cgfloat;
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) const char cgchar;
typedef __attribute__((address_space(4))) const float kfloat;   // read-only: scalar loads

// loads and stores at a 32-bit byte offset from a uniform base
__device__ __forceinline__ float ldg(cgfloat* base, uint32_t off) {
  return *(cgfloat*)((cgchar*)base + off);
}
__device__ __forceinline__ void stg(gfloat* base, uint32_t off, float v) {
  *(gfloat*)((gchar*)base + off) = v;
}

// shr1: lane l receives lane l-1's value (lane 0 receives 0);
// shl1: lane l receives lane l+1's value (lane 63 receives 0).
__device__ __forceinline__ float shr1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float shl1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

struct Vec3 {
  float a, b, c;       // rho', rho u, rho v
};
__device__ __forceinline__ Vec3 v3(float a, float b, float c) { return Vec3{a, b, c}; }
__device__ __forceinline__ Vec3 operator+(const Vec3& x, const Vec3& y) {
  return Vec3{x.a + y.a, x.b + y.b, x.c + y.c};
}
__device__ __forceinline__ Vec3 sel(bool p, const Vec3& x) {
  return Vec3{p ? x.a : 0.0f, p ? x.b : 0.0f, p ? x.c : 0.0f};
}
__device__ __forceinline__ Vec3 shr1(const Vec3& x) { return Vec3{shr1(x.a), shr1(x.b), shr1(x.c)}; }
__device__ __forceinline__ Vec3 shl1(const Vec3& x) { return Vec3{shl1(x.a), shl1(x.b), shl1(x.c)}; }

struct Row {           // one cell of q_b as the y sweep sees it
  float q1, q2, q3;    // rho', rho u, rho v
  float un, ut;        // v = q3/rho (y normal), u = q2/rho (tangential)
};

struct XRow {          // one cell of q_xy (the recomputed-state check)
  float q1, q2, q3, u, v;
};

__device__ __forceinline__ fvad::SumAdj zero_sum() { return fvad::SumAdj{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
// a where keep holds, else 0: one select per field (no branch, so no merge
// of two versions of every field)
__device__ __forceinline__ fvad::SumAdj keep_if(bool keep, const fvad::SumAdj& a) {
  return fvad::SumAdj{keep ? a.s1 : 0.f, keep ? a.s2 : 0.f, keep ? a.s3 : 0.f,
                      keep ? a.dr : 0.f, keep ? a.dn : 0.f, keep ? a.dt : 0.f};
}

// The reverse step of one work item of an instance whose ghost cells follow
// the boundary policy BC.  kCheck: also recompute q_{n+1} on the item's
// cells and count those that differ from the stored one (D.mismatch).
template <class BC, bool kCheck>
__device__ __forceinline__ void reverse_item(const AdjDesc& D, const AdjItem& it, int item,
                                             int k, int slot, int par) {
  const int nx = D.nx, ny = D.ny;
  const uint32_t n = (uint32_t)nx * (uint32_t)ny;
  const Ctl* __restrict__ C = D.ctl_log + k;
  const double dt = C->dt;
  const float dtx = C->dtx, dty = C->dty, ddx2 = C->ddx2, ddy2 = C->ddy2;
  const float sin_t = (float)C->drive_s, cos_t = (float)C->drive_c;
  const float hbeta = D.hbeta;
  const float* qs = D.seg + (size_t)slot * 3 * n;
  cgfloat* __restrict__ in1 = (cgfloat*)qs;
  cgfloat* __restrict__ in2 = (cgfloat*)(qs + n);
  cgfloat* __restrict__ in3 = (cgfloat*)(qs + 2 * (size_t)n);
  cgfloat* __restrict__ nq1 = (cgfloat*)(qs + 3 * (size_t)n);   // q_{n+1}
  cgfloat* __restrict__ nq2 = (cgfloat*)(qs + 4 * (size_t)n);
  cgfloat* __restrict__ nq3 = (cgfloat*)(qs + 5 * (size_t)n);
  cgfloat* __restrict__ g1p = (cgfloat*)D.lam[par];
  cgfloat* __restrict__ g2p = (cgfloat*)(D.lam[par] + n);
  cgfloat* __restrict__ g3p = (cgfloat*)(D.lam[par] + 2 * (size_t)n);
  gfloat* __restrict__ o1 = (gfloat*)D.lam[par ^ 1];
  gfloat* __restrict__ o2 = (gfloat*)(D.lam[par ^ 1] + n);
  gfloat* __restrict__ o3 = (gfloat*)(D.lam[par ^ 1] + 2 * (size_t)n);

  const int lane = threadIdx.x;
  const int i = it.x0 - kHalo + lane;
  const bool col_in = i >= 0 && i < nx;
  const int ic = min(max(i, 0), nx - 1);
  const bool col_int = i >= 2 && i <= nx - 3;
  const bool col_x = col_in && i >= 1;           // lane holds an x interface of the grid
  const bool out_lane = i >= it.xo0 && i < it.xo1;
  // dt/d on the lanes of interior columns, 0 on ghost columns
  const float dty_c = col_int ? dty : 0.0f, dtx_c = col_int ? dtx : 0.0f;
  const bool box_col = i >= D.bx0 && i < D.bx0 + D.bnx;
  const uint32_t row_bytes = 4u * (uint32_t)nx;
  const uint32_t cb = 4u * (uint32_t)ic;
  const float two_dt = (float)(2.0 * dt);
  const int y0 = it.y0, y1 = it.y1;

  // the boundary policy's source columns and patch, as the step kernel
  fvbc::LaneMap lm;
  BC::lane(ic, nx, lm, 4u);
  fvbc::Tables tb{D.edge_c, D.edge_s, nullptr};
  fvbc::PatchCtx<float> pc;
  pc.sin_t = sin_t;
  pc.cos_t = cos_t;
  pc.T = C->T;
  pc.hk = 0.0f;
  pc.mdx = 0;
  pc.x0 = it.x0 - 1;     // the column of lane 3 (the step kernel's first output lane)
  pc.nx = nx;
  pc.ny = ny;
  pc.ic = ic;
  pc.left = lm.left;
  pc.any_left = __builtin_amdgcn_ballot_w64(lm.left) != 0;
  auto load_raw = [&](int r, float& a, float& b, float& c, fvbc::RowData<float>& h) {
    const int rc = min(max(r, 0), ny - 1);
    const uint32_t o = BC::offset(rc, ny, lm, row_bytes);
    a = ldg(in1, o);
    b = ldg(in2, o);
    c = ldg(in3, o);
    if (BC::kRowData) BC::template load_row<float>(tb, rc, h);
  };
  auto patch = [&](Row& q, const fvbc::RowData<float>& h, int r) {
    fvbc::Cell<float> c{q.q1, q.q2, q.q3};
    BC::patch(pc, min(max(r, 0), ny - 1), c, h);
    q.q1 = c.q1;
    q.q2 = c.q2;
    q.q3 = c.q3;
    const float ir = fvna::rcp(1.0f + q.q1);
    q.un = q.q3 * ir;
    q.ut = q.q2 * ir;
  };
  // g(l) = lam_{n+1}(l) + dJ_n/dq_{n+1}(l), where the box term is
  // 2 dt rho'_{n+1}; row l is requested one step before it is used
  // (load_g), then completed (finish_g), which also adds J_n's share of the
  // item's own rows
  double jacc = 0.0;
  const bool in_box_col = col_in && box_col;
  auto box_row = [&](int l) { return l >= 0 && l < ny && l >= D.by0 && l < D.by0 + D.bny; };
  auto load_g = [&](int l, Vec3& g, float& rn) {
    g = v3(0.f, 0.f, 0.f);
    rn = 0.0f;
    if (l < 0 || l >= ny) return;
    const uint32_t o = (uint32_t)l * row_bytes + cb;
    g.a = ldg(g1p, o);
    g.b = ldg(g2p, o);
    g.c = ldg(g3p, o);
    if (box_row(l)) rn = ldg(nq1, o);
  };
  auto finish_g = [&](int l, Vec3& g, float rn) {
    g = sel(col_in, g);
    if (box_row(l)) {
      g.a = in_box_col ? fvna::fma_(two_dt, rn, g.a) : g.a;
      if (l >= y0 && l < y1 && in_box_col && out_lane) jacc += (double)rn * (double)rn;
    }
  };

  // ---- windows (every one zeroed: the first rows of the march only warm
  // them up, and nothing they hold reaches an emitted row) ----
  Row b1{}, b2{}, b3{}, b4{};               // q_b rows r-1 .. r-4
  fvna::Iface f1{}, f2{}, f3{};             // y interfaces r-1, r-2, r-3
  fvad::LimT lim2{};                        // limiter of interface r-2
  float dl1[3] = {0.f, 0.f, 0.f};           // W(r-1) . W(r-2) per family
  float Pp[3] = {0.f, 0.f, 0.f};            // P of interface r-2
  fvad::WaveAdj acc3, acc2;                 // wave adjoints of interfaces r-3, r-2
  fvad::zero(acc3);
  fvad::zero(acc2);
  Vec3 ly1{}, ly2{};                        // lam_y rows r-3, r-4
  fvad::SumAdj gRp = zero_sum();            // transpose of interface r-4 (its cell r-4 part)
  Vec3 gd{}, g0{}, gu{}, gnx{};             // g rows m-1, m, m+1 and the prefetched row
  float rnx = 0.0f;                         // rho' of q_{n+1} on the prefetched row (box rows)
  Vec3 hbot{}, htop{}, hold{};              // lam_b(0) + lam_b(1); lam_b(ny-2) + lam_b(ny-1);
                                            // lam_n(ny-3) waiting for the top rule
  XRow xc1{}, xc2{};                        // check: q_xy rows m-1, m-2
  uint32_t bad = 0;

  const int r0 = y0 - 3, r_end = y1 + 3;
  float n1, n2, n3;
  fvbc::RowData<float> nh{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  load_raw(r0 - 1, n1, n2, n3, nh);
  b1.q1 = n1; b1.q2 = n2; b1.q3 = n3;
  patch(b1, nh, r0 - 1);
  load_raw(r0, n1, n2, n3, nh);
  load_g(r0 - 1, gnx, rnx);

  for (int r = r0; r <= r_end; ++r) {
    Row bn;
    bn.q1 = n1; bn.q2 = n2; bn.q3 = n3;
    patch(bn, nh, r);
    load_raw(r + 1, n1, n2, n3, nh);     // row index clamped (past r_end: unused)
    gd = g0; g0 = gu; gu = gnx;
    finish_g(r - 1, gu, rnx);
    load_g(r, gnx, rnx);

    // ---- y interface r (rows r-1 | r): normal = q3, tangential = q2 ----
    fvna::Iface fn;
    fvna::riemann(b1.q1, b1.q3, b1.q2, b1.un, b1.ut, bn.q1, bn.q3, bn.q2, bn.un, bn.ut, hbeta, fn);
    float dln[3];
    dln[0] = fvna::dot1(fn.w, f1.w);
    dln[1] = fvna::dot2(fn.w, f1.w);
    dln[2] = fn.w.a3 * f1.w.a3;
    // ---- y limiter at interface r-1 ----
    float P1[3], M[3];
    fvad::LimT lim1;
    fvad::limiter_rec(f1, dl1[0], dln[0], dl1[1], dln[1], dl1[2], dln[2], dty, P1, M, lim1);

    // The rows before y0 only warm up the windows of the forward recompute
    // (the y interfaces and limiters the first emitted rows need); everything
    // from q_y on is skipped there, and its window entries stay at zero (no
    // emitted row reads them).  One guard: finer ones split the row into
    // blocks the scheduler cannot interleave, which costs tall items more
    // than the rows they skip.
    Vec3 lyc{};                               // lam_y of row m = r-2
    fvad::WaveAdj acc1;                       // wave adjoint of interface r-1
    fvad::zero(acc1);
    if (r >= y0) {
      // ---- q_y on row m = r-2 (components: density, normal=q3, tangential=q2) ----
      const int m = r - 2;
      const bool row_int = m >= 2 && m <= ny - 3;
      const bool upd = col_int && row_int;
      const float dtyu = row_int ? dty_c : 0.0f, dtxu = row_int ? dtx_c : 0.0f;
      const float y1q = b2.q1 - dtyu * (Pp[0] + M[0]);
      const float y3q = b2.q3 - dtyu * (Pp[1] + M[1]);
      const float y2q = b2.q2 - dtyu * (Pp[2] + M[2]);

      // ---- x sweep on row m: normal = q2, tangential = q3 ----
      const float ir = fvna::rcp(1.0f + y1q);
      const float un = y2q * ir, ut = y3q * ir;
      const float l1 = shr1(y1q), l2 = shr1(y2q), l3 = shr1(y3q);
      const float lun = shr1(un), lut = shr1(ut);
      fvna::Iface fx;
      fvna::riemann(l1, l2, l3, lun, lut, y1q, y2q, y3q, un, ut, hbeta, fx);
      fvna::Waves wl;
      wl.a1 = shr1(fx.w.a1); wl.w21 = shr1(fx.w.w21); wl.w31 = shr1(fx.w.w31);
      wl.a2 = shr1(fx.w.a2); wl.w22 = shr1(fx.w.w22); wl.w32 = shr1(fx.w.w32);
      wl.a3 = shr1(fx.w.a3);
      const float dlx1 = fvna::dot1(fx.w, wl), dlx2 = fvna::dot2(fx.w, wl);
      const float dlx3 = fx.w.a3 * wl.a3;
      const float drx1 = shl1(dlx1), drx2 = shl1(dlx2), drx3 = shl1(dlx3);
      float PX[3], MX[3];
      fvad::LimT limx;
      fvad::limiter_rec(fx, dlx1, drx1, dlx2, drx2, dlx3, drx3, dtx, PX, MX, limx);
      const float mr0 = shl1(MX[0]), mr1 = shl1(MX[1]), mr2 = shl1(MX[2]);
      XRow xc;
      xc.q1 = y1q - dtxu * (mr0 + PX[0]);
      xc.q2 = y2q - dtxu * (mr1 + PX[1]);
      xc.q3 = y3q - dtxu * (mr2 + PX[2]);
      const float ir2 = fvna::rcp(1.0f + xc.q1);
      xc.u = xc.q2 * ir2;
      xc.v = xc.q3 * ir2;

      if (kCheck) {
        // the source of row m-1 from q_xy rows m-2, m-1, m, against q_{n+1}
        const int s = m - 1;
        const float uwe = shr1(xc1.u) + shl1(xc1.u);
        const float vwe = shr1(xc1.v) + shl1(xc1.v);
        const float i2 = fvna::source_increment(ddx2, ddy2, xc1.u, uwe, xc2.u, xc.u);
        const float i3 = fvna::source_increment(ddx2, ddy2, xc1.v, vwe, xc2.v, xc.v);
        const bool s_int = col_int && s >= 2 && s <= ny - 3;
        const float r2 = s_int ? fvna::apply_source(xc1.q2, i2) : xc1.q2;
        const float r3 = s_int ? fvna::apply_source(xc1.q3, i3) : xc1.q3;
        // a Taylor state's ghost ring is rewritten in place by the next step's
        // prepass (its closed form at the next T): only the cells the step
        // computes are compared there
        const bool cmp = BC::kId != kBcTaylor || s_int;
        if (s >= y0 && s < y1 && out_lane && cmp) {
          const uint32_t o = (uint32_t)s * row_bytes + cb;
          const bool same = __float_as_uint(ldg(nq1, o)) == __float_as_uint(xc1.q1) &&
                            __float_as_uint(ldg(nq2, o)) == __float_as_uint(r2) &&
                            __float_as_uint(ldg(nq3, o)) == __float_as_uint(r3);
          bad += same ? 0u : 1u;
        }
        xc2 = xc1;
        xc1 = xc;
      }

      // ---- S^T on row m: the source adds ddx2 (u_W - 2u + u_E) +
      // ddy2 (u_S - 2u + u_N) to rho u, the same with v to rho v, on interior cells ----
      const bool int_d = col_int && m - 1 >= 2 && m - 1 <= ny - 3;
      const bool int_u = col_int && m + 1 >= 2 && m + 1 <= ny - 3;
      const float G2c = upd ? g0.b : 0.0f, G3c = upd ? g0.c : 0.0f;
      const float G2y = (int_d ? gd.b : 0.0f) + (int_u ? gu.b : 0.0f);
      const float G3y = (int_d ? gd.c : 0.0f) + (int_u ? gu.c : 0.0f);
      const float sU = fvna::fma_(ddx2, fvna::fma_(-2.0f, G2c, shr1(G2c) + shl1(G2c)),
                                  ddy2 * fvna::fma_(-2.0f, G2c, G2y));
      const float sV = fvna::fma_(ddx2, fvna::fma_(-2.0f, G3c, shr1(G3c) + shl1(G3c)),
                                  ddy2 * fvna::fma_(-2.0f, G3c, G3y));
      Vec3 lxy;
      lxy.a = g0.a - fvna::fma_(sU, xc.u, sV * xc.v) * ir2;
      lxy.b = fvna::fma_(sU, ir2, g0.b);
      lxy.c = fvna::fma_(sV, ir2, g0.c);

      // ---- X^T on row m ----
      // q_xy(k) = q_y(k) - dtxu(k) (PX(k) + MX(k+1))
      float gPX[3] = {-dtxu * lxy.a, -dtxu * lxy.b, -dtxu * lxy.c};
      float gMX[3] = {shr1(gPX[0]), shr1(gPX[1]), shr1(gPX[2])};
      fvad::WaveAdj gx;
      fvad::zero(gx);
      float gdx[3];
      fvad::limiter_bw(fx, limx, dtx, gPX, gMX, gx, gdx);
      {
        // family p of interface k dotted with k-1 (L_p > 0) or with k+1
        fvna::Waves wr;
        wr.a1 = shl1(fx.w.a1); wr.w21 = shl1(fx.w.w21); wr.w31 = shl1(fx.w.w31);
        wr.a2 = shl1(fx.w.a2); wr.w22 = shl1(fx.w.w22); wr.w32 = shl1(fx.w.w32);
        wr.a3 = shl1(fx.w.a3);
        const float Ls[3] = {fx.L1, fx.L2, fx.L3};
        float kl[3], kr[3];
  #pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float gdl = Ls[p] > 0.0f ? gdx[p] : 0.0f;
          const float gdr = gdx[p] - gdl;
          kl[p] = gdl + shr1(gdr);      // times W(k-1): own left dot, and k-1's right dot
          kr[p] = gdr + shl1(gdl);      // times W(k+1): own right dot, and k+1's left dot
        }
        fvad::add_scaled(gx, kl, wl);
        fvad::add_scaled(gx, kr, wr);
      }
      fvad::SumAdj sx = fvad::riemann_bw(fvna::add_nc(l1, y1q), fvna::add_nc(lun, un),
                                         fvna::add_nc(lut, ut), y1q - l1, y2q - l2, hbeta, fx, gx);
      sx = keep_if(col_x, sx);
      {
        // cell k: interface k's R part and interface k+1's L part (lane k+1)
        const float ar1 = (sx.s1 + sx.dr) + shl1(sx.s1 - sx.dr);
        const float an = sx.dn - shl1(sx.dn);
        const float at = sx.dt - shl1(sx.dt);
        const float aun = sx.s2 + shl1(sx.s2);
        const float aut = sx.s3 + shl1(sx.s3);
        lyc.a = lxy.a + ar1 - fvna::fma_(aun, un, aut * ut) * ir;
        lyc.b = lxy.b + fvna::fma_(aun, ir, an);
        lyc.c = lxy.c + fvna::fma_(aut, ir, at);
      }

      // ---- Y^T: limiter of interface m (P -> cell m, M -> cell m-1) ----
      {
        const float dtyu1 = (m - 1 >= 2 && m - 1 <= ny - 3) ? dty_c : 0.0f;
        // y component order (rho', normal = rho v, tangential = rho u)
        float gPy[3] = {-dtyu * lyc.a, -dtyu * lyc.c, -dtyu * lyc.b};
        float gMy[3] = {-dtyu1 * ly1.a, -dtyu1 * ly1.c, -dtyu1 * ly1.b};
        float gdy[3];
        fvad::limiter_bw(f2, lim2, dty, gPy, gMy, acc2, gdy);
        const float Ls[3] = {f2.L1, f2.L2, f2.L3};
        float kdn[3], kup[3];
  #pragma unroll
        for (int p = 0; p < 3; ++p) {
          kdn[p] = Ls[p] > 0.0f ? gdy[p] : 0.0f;   // dotted with interface m-1
          kup[p] = gdy[p] - kdn[p];                // ... with m+1
        }
        fvad::add_scaled(acc2, kdn, f3.w);
        fvad::add_scaled(acc2, kup, f1.w);
        fvad::add_scaled(acc3, kdn, f2.w);
        fvad::add_scaled(acc1, kup, f2.w);
      }
      // ---- Riemann transpose of interface m-1 = r-3 (rows r-4 | r-3) ----
      fvad::SumAdj sy = fvad::riemann_bw(fvna::add_nc(b4.q1, b3.q1), fvna::add_nc(b4.un, b3.un),
                                         fvna::add_nc(b4.ut, b3.ut), b3.q1 - b4.q1, b3.q3 - b4.q3,
                                         hbeta, f3, acc3);
      sy = keep_if(col_in && r - 3 >= 1 && r - 3 <= ny - 1, sy);
      // ---- lam_b of row r-4: lam_y, interface r-4's R part, interface r-3's L part ----
      Vec3 lb;
      {
        const float br1 = (gRp.s1 + gRp.dr) + (sy.s1 - sy.dr);
        const float bn_ = gRp.dn - sy.dn;
        const float bt = gRp.dt - sy.dt;
        const float bun = gRp.s2 + sy.s2;
        const float but = gRp.s3 + sy.s3;
        const float ir4 = fvna::rcp(1.0f + b4.q1);
        lb.a = ly2.a + br1 - fvna::fma_(bun, b4.un, but * b4.ut) * ir4;
        lb.c = ly2.c + fvna::fma_(bun, ir4, bn_);     // normal = rho v
        lb.b = ly2.b + fvna::fma_(but, ir4, bt);      // tangential = rho u
      }
      gRp = sy;

      // ---- B^T and the drive gradient of row j = r-4 ----
      const int j = r - 4;
      if (j >= y0 && j < y1) {
        const bool interior = col_int && j >= 2 && j <= ny - 3;
        Vec3 a = sel(interior, lb);
        bool defer = false;
        if (BC::kId != kBcTaylor) {
          // zero-gradient copies (apps/common/boundary.h): left columns <- 2
          // (zero-gradient app only; the edge drive sets them), right columns
          // <- nx-3 on rows 2..ny-3, rows 0, 1 <- 2 and ny-2, ny-1 <- ny-3 off
          // the left columns
          if (BC::kId == kBcZeroGrad) {
            const Vec3 w1 = shr1(lb), w2 = shr1(w1);
            if (i == 2) a = a + w1 + w2;
          }
          {
            const Vec3 e1 = shl1(lb), e2 = shl1(e1);
            if (i == nx - 3 && j >= 2 && j <= ny - 3) a = a + e1 + e2;
          }
          if (j == 0) hbot = lb;
          if (j == 1) hbot = hbot + lb;
          if (j == 2 && i >= 2) a = a + hbot;
          if (j >= ny - 2) htop = htop + lb;
          if (j == ny - 3) {
            hold = a;
            defer = true;
          }
        }
        const uint32_t o = (uint32_t)j * row_bytes + cb;
        if (out_lane && !defer) {
          stg(o1, o, a.a);
          stg(o2, o, a.b);
          stg(o3, o, a.c);
        }
        if (BC::kId != kBcTaylor && j == ny - 1) {
          const Vec3 t = i >= 2 ? hold + htop : hold;
          const uint32_t ot = (uint32_t)(ny - 3) * row_bytes + cb;
          if (out_lane) {
            stg(o1, ot, t.a);
            stg(o2, ot, t.b);
            stg(o3, ot, t.c);
          }
        }
        if (BC::kId == kBcEdgeDrive && it.xo0 == 0) {
          // left columns hold p = ampc sin(theta) + amps cos(theta), rho u = p (p + 1)
          const float ec = ((kfloat*)D.edge_c)[j], es = ((kfloat*)D.edge_s)[j];
          fvbc::Cell<float> dc;
          fvbc::drive(ec, es, sin_t, cos_t, dc);
          const float p = dc.q1;
          const float tp = (i == 0 || i == 1) ? fvna::fma_(lb.b, 2.0f * p + 1.0f, lb.a) : 0.0f;
          // columns 0 and 1 are lanes kHalo and kHalo + 1 of the first strip
          const int tb0 = __builtin_amdgcn_readlane(__float_as_int(tp), kHalo);
          const int tb1 = __builtin_amdgcn_readlane(__float_as_int(tp), kHalo + 1);
          const double g = (double)__int_as_float(tb0) + (double)__int_as_float(tb1);
          if (lane == 0) {
            D.grad_ec[j] += g * (double)sin_t;
            D.grad_es[j] += g * (double)cos_t;
          }
        }
      }

    }

    // ---- rotate the windows ----
    b4 = b3; b3 = b2; b2 = b1; b1 = bn;
    f3 = f2; f2 = f1; f1 = fn;
    lim2 = lim1;
    dl1[0] = dln[0]; dl1[1] = dln[1]; dl1[2] = dln[2];
    Pp[0] = P1[0]; Pp[1] = P1[1]; Pp[2] = P1[2];
    acc3 = acc2; acc2 = acc1;
    ly2 = ly1; ly1 = lyc;
  }

  // J of this step over the item's box cells, one double per item
  double v = jacc;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0 && v != 0.0) D.jpart[item] += dt * v;
  if (kCheck) {
    if (__builtin_amdgcn_ballot_w64(bad != 0) != 0) {
      uint32_t s = bad;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) atomicAdd(D.mismatch, s);
    }
  }
}

}  // namespace adjm

using namespace adjm;

// One reverse step of every instance: step index k of the run (its logged
// control block), segment slot `slot` (q_n; slot+1 holds q_{n+1}), adjoint
// ping-pong index par (lam[par] -> lam[par^1]).  Wave w runs the items
// wave_first[w] .. wave_first[w+1] in turn (the end of one strip and the
// start of the next, when the host balances the waves' rows).
template <bool kCheck>
__global__ __launch_bounds__(64) void adj_reverse_kernel(const AdjDesc* __restrict__ ds,
                                                         const AdjItem* __restrict__ items,
                                                         const int32_t* __restrict__ wave_first,
                                                         int n_waves, int k, int slot, int par) {
  // XCD-aware order, as the step kernel: consecutive waves of a strip on one XCD
  int w;
  {
    const int b = blockIdx.x;
    const int q = n_waves >> 3, rem = n_waves & 7, x = b & 7, kk = b >> 3;
    w = x * q + min(x, rem) + kk;
  }
  const int i0 = wave_first[w], i1 = wave_first[w + 1];
  for (int item = i0; item < i1; ++item) {
    const AdjItem it = items[item];
    const AdjDesc& D = ds[it.inst];
    switch (__builtin_amdgcn_readfirstlane(D.bc)) {
      case kBcZeroGrad:
        reverse_item<fvbc::ZeroGradient, kCheck>(D, it, item, k, slot, par);
        break;
      case kBcEdgeDrive:
        reverse_item<fvbc::EdgeDrive, kCheck>(D, it, item, k, slot, par);
        break;
      case kBcTaylor:
        reverse_item<fvbc::Taylor, kCheck>(D, it, item, k, slot, par);
        break;
      default:
        break;
    }
  }
}

}  // namespace fv

// ---------------------------------------------------------------- launches --

extern "C" hipError_t fv_adj_reverse(const fv::AdjDesc* ds, const fv::AdjItem* items,
                                     const int32_t* wave_first, int n_waves, int k, int slot,
                                     int par, int check, hipStream_t s) {
  if (n_waves <= 0) return hipSuccess;
  if (check)
    hipLaunchKernelGGL(fv::adj_reverse_kernel<true>, dim3(n_waves), dim3(64), 0, s, ds, items,
                       wave_first, n_waves, k, slot, par);
  else
    hipLaunchKernelGGL(fv::adj_reverse_kernel<false>, dim3(n_waves), dim3(64), 0, s, ds, items,
                       wave_first, n_waves, k, slot, par);
  return hipGetLastError();
}

// Resident reverse waves per CU.
extern "C" hipError_t fv_adj_occupancy(int* waves_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(waves_per_cu, fv::adj_reverse_kernel<false>,
                                                      64, 0);
}
