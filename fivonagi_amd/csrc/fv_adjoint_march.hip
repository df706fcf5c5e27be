// fv_adjoint_march.hip -- the adjoint step as three march kernels (gfx950).
//
// The reference has no adjoint.  The step (reference: common/main.cu:310-635,
// apps/na/fv.h, apps/na/source.h, common/boundary.h) is
//   q_b = B(q_n, T_n), q_y = Y(q_b), q_xy = X(q_y), q_{n+1} = S(q_xy)
// and the objective J = sum_n dt_n sum_box rho'_{n+1}^2 (dt frozen, the
// executed branch path differentiated).  Per step:
//   adj_march_x<false>  q_n -> q_{n+1}                      (segment recompute)
//   adj_march_x<true>   q_n, lam_{n+1} -> lam_y = X^T S^T (lam_{n+1} + dJ/dq_{n+1})
//   adj_march_y         q_n, lam_y -> lam_n = B^T Y^T lam_y, drive gradient
// Each is a wave64 marching up a strip of columns, like the forward step
// kernel: the forward quantities the transposes need (waves, speeds,
// limiter branches) are recomputed in registers from q_n instead of being
// stored; x neighbours come from DPP lane shifts, y neighbours from
// register windows.  adj_march_x uses a 4-column halo (56 output columns),
// adj_march_y a 2-column halo (60).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../apps/na/adjoint.h"
#include "fv_adjoint.h"

namespace fv {
namespace adjm {

constexpr int kLanes = 64;

__device__ __forceinline__ float shr1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float shl1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ int shr1i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int shl1i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}

struct Cell { float r, u, v; };   // rho', rho u, rho v

// Cell (i, r) of q_b = B(q_n, T) for the lane's column i (rows and columns
// clamped into the grid; lanes and rows outside it only feed halo work).
// Ghost rules as apps/common/boundary.h: left, then top, then bottom, then
// right; the drive value p and the Taylor closed form are evaluated in double.
struct Loader {
  const float* q;
  int nx, ny, n, bc, i, ic;
  double T, omega;
  const float* amp;
  const float* phase;
  double tay_rhoa, tay_c, tay_cos, tay_sin, tay_yoff, dx, dy;
  int tay_jc;

  __device__ Cell operator()(int r) const {
    const int rc = min(max(r, 0), ny - 1);
    Cell c;
    if (bc == 2) {
      if (ic <= 1 || rc <= 1 || ic >= nx - 2 || rc >= ny - 2) {
        const double arg = tay_c * ((double)ic * dx * tay_cos +
                                    ((double)(rc - tay_jc) * dy + tay_yoff) * tay_sin - T);
        const double rr = tay_rhoa * tanh(arg);
        c.r = (float)rr; c.u = (float)(rr * tay_cos); c.v = (float)(rr * tay_sin);
        return c;
      }
      const int o = ic + rc * nx;
      c.r = q[o]; c.u = q[n + o]; c.v = q[2 * n + o];
      return c;
    }
    int si = ic, sj = rc;
    if (ic <= 1) {
      if (bc == 1) {
        const double th = omega * T + (double)phase[rc];
        const float p = (float)((double)amp[rc] * sin(th));
        c.r = p; c.u = p * (p + 1.0f); c.v = 0.0f;
        return c;
      }
      si = 2;
    } else if (rc >= ny - 2) {
      sj = ny - 3;
    } else if (rc <= 1) {
      sj = 2;
    } else if (ic >= nx - 2) {
      si = nx - 3;
    }
    const int o = si + sj * nx;
    c.r = q[o]; c.u = q[n + o]; c.v = q[2 * n + o];
    return c;
  }
};

__device__ __forceinline__ Loader make_loader(const AdjDesc& d, const float* q, double T, int i) {
  Loader L;
  L.q = q; L.nx = d.nx; L.ny = d.ny; L.n = d.nx * d.ny; L.bc = d.bc; L.i = i;
  L.ic = min(max(i, 0), d.nx - 1);
  L.T = T; L.omega = d.omega; L.amp = d.edge_amp; L.phase = d.edge_phase;
  L.tay_rhoa = d.tay_rhoa; L.tay_c = d.tay_c; L.tay_cos = d.tay_cos; L.tay_sin = d.tay_sin;
  L.tay_yoff = d.tay_yoff; L.dx = d.dx; L.dy = d.dy; L.tay_jc = d.tay_jc;
  return L;
}

struct YIf { float W[7]; float s[3]; };   // y interface: waves (rho', rho v, rho u) and speeds

__device__ __forceinline__ bool interior(int i, int j, int nx, int ny) {
  return i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
}

// y interface between rows a (below) and b (above), y component order
__device__ __forceinline__ void y_waves(const Cell& a, const Cell& b, float beta, YIf& f) {
  fvad::RW w;
  fvad::riemann_fw(a.r, a.v, a.u, b.r, b.v, b.u, beta, w);
#pragma unroll
  for (int c = 0; c < 7; ++c) f.W[c] = w.W[c];
  f.s[0] = w.s[0]; f.s[1] = w.s[1]; f.s[2] = w.s[2];
}

}  // namespace adjm

using namespace adjm;

// ------------------------------------------------------------ x march --
// kRev = false: q_{n+1} = S(X(Y(B(q_n)))) into state slot+1.
// kRev = true:  lam_y = X^T S^T (lam + dJ/dq_{n+1}) into d.lamx, and
//               jstep[k] = dt_k sum_box rho'_{n+1}^2.
template <bool kRev>
__global__ __launch_bounds__(64) void adj_march_x(const AdjDesc* __restrict__ ds,
                                                  const AdjItem* __restrict__ items, int k,
                                                  int slot) {
  constexpr int kHalo = 4;
  const AdjItem it = items[blockIdx.x];
  const AdjDesc& d = ds[it.inst];
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const float* q = d.seg + (size_t)slot * 3 * n;
  const double dt = d.dt_log[k];
  const double T = d.T_log[k];
  const float dtx = (float)(dt / d.dx), dty = (float)(dt / d.dy);
  const float ddx2 = (float)(d.delta * dt / (d.dx * d.dx));
  const float ddy2 = (float)(d.delta * dt / (d.dy * d.dy));
  const float beta = d.beta;
  const int lane = threadIdx.x;
  const int i = it.x0 - kHalo + lane;
  const bool col_int = i >= 2 && i <= nx - 3;
  const bool out_lane = lane >= kHalo && lane < kLanes - kHalo && i < nx;
  const bool iface_x = i >= 1 && i <= nx - 1;   // x interface (i-1 | i) exists
  const int y0 = it.y0, y1 = it.y1;
  const Loader load = make_loader(d, q, T, i);
  const bool in_box_col = i >= d.bx0 && i < d.bx0 + d.bnx;

  // windows: q_b rows r, r-1, r-2; y interfaces r, r-1, r-2; P of interface r-2
  Cell b0{}, b1{}, b2{};
  YIf f0{}, f1{}, f2{};
  float Pp[3] = {0.f, 0.f, 0.f};
  // forward-only: q_xy rows m-1, m (the source at m-1 needs m-2..m)
  Cell xm2{}, xm1{};
  float um2 = 0.f, vm2 = 0.f, um1 = 0.f, vm1 = 0.f;
  // reverse-only: lam rows r-1 (next), r-2, r-3
  float la[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  double jacc = 0.0;

  auto load_lam = [&](int row, float out[3]) {
    if (row < 0 || row >= ny || i < 0 || i >= nx) {
      out[0] = out[1] = out[2] = 0.0f;
      return;
    }
    const int o = i + row * nx;
    out[0] = d.lam[o]; out[1] = d.lam[n + o]; out[2] = d.lam[2 * n + o];
  };

  const int r_end = kRev ? y1 + 1 : y1 + 2;
  if (kRev) {
    load_lam(y0 - 1, la[1]);   // becomes row r-3 at r = y0 + 2
    load_lam(y0, la[0]);
  }
  Cell nx0 = load(y0 - 3);     // q_b rows are fetched one iteration ahead
  for (int r = y0 - 3; r <= r_end; ++r) {
    // ---- q_b row r, y interface r, y limiter at r-1 ----
    b2 = b1; b1 = b0;
    b0 = nx0;
    nx0 = load(r + 1);
    f2 = f1; f1 = f0;
    if (r >= y0 - 2) y_waves(b1, b0, beta, f0);
    if (r < y0) continue;
    float Py[3], My[3];
    {
      fvad::Lim l;
      fvad::limiter_fw(f2.W, f1.W, f1.s, f0.W, dty, l, Py, My);
    }
    // q_y(m = r-2) = q_b(m) - dty (P(m) + M(m+1)), y component order (rho', rho v, rho u)
    const int m = r - 2;
    float yq[3];
    {
      const bool upd = col_int && m >= 2 && m <= ny - 3;
      const float du = upd ? dty : 0.0f;
      yq[0] = b2.r - du * (Pp[0] + My[0]);
      yq[2] = b2.v - du * (Pp[1] + My[1]);
      yq[1] = b2.u - du * (Pp[2] + My[2]);
    }
    Pp[0] = Py[0]; Pp[1] = Py[1]; Pp[2] = Py[2];
    if (r < y0 + 1) continue;
    if (kRev && r < y0 + 2) {
      float nl[3];
      load_lam(r, nl);   // keep the lam window aligned (row r - 1 next)
      la[2][0] = la[1][0]; la[2][1] = la[1][1]; la[2][2] = la[1][2];
      la[1][0] = la[0][0]; la[1][1] = la[0][1]; la[1][2] = la[0][2];
      la[0][0] = nl[0]; la[0][1] = nl[1]; la[0][2] = nl[2];
      continue;
    }

    // ---- x sweep on row m (x component order: rho', rho u, rho v) ----
    fvad::RW w;
    fvad::riemann_fw(shr1(yq[0]), shr1(yq[1]), shr1(yq[2]), yq[0], yq[1], yq[2], beta, w);
    float Wl[7], Wr[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) { Wl[c] = shr1(w.W[c]); Wr[c] = shl1(w.W[c]); }
    fvad::Lim lx;
    float Px[3], Mx[3];
    fvad::limiter_fw(Wl, w.W, w.s, Wr, dtx, lx, Px, Mx);
    float xq[3];
    {
      const bool upd = col_int && m >= 2 && m <= ny - 3;
      const float du = upd ? dtx : 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) xq[c] = yq[c] - du * (Px[c] + shl1(Mx[c]));
    }

    if (!kRev) {
      // ---- source on row m-1 (apps/na/source.h:84-104), store q_{n+1} ----
      const float irho = fvad::rcp(1.0f + xq[0]);
      const float u0 = xq[1] * irho, v0 = xq[2] * irho;
      // cross-lane moves with every lane active (a disabled source lane reads 0)
      const float uw = shr1(um1), ue = shl1(um1), vw = shr1(vm1), ve = shl1(vm1);
      if (r >= y0 + 3 && out_lane) {
        const int s = m - 1;
        float o1 = xm1.u, o2 = xm1.v;
        if (col_int && s >= 2 && s <= ny - 3) {
          o1 = xm1.u + ddx2 * (uw - 2.0f * um1 + ue) + ddy2 * (um2 - 2.0f * um1 + u0);
          o2 = xm1.v + ddx2 * (vw - 2.0f * vm1 + ve) + ddy2 * (vm2 - 2.0f * vm1 + v0);
        }
        float* out = d.seg + (size_t)(slot + 1) * 3 * n;
        const int o = i + s * nx;
        out[o] = xm1.r;
        out[n + o] = o1;
        out[2 * n + o] = o2;
      }
      xm2 = xm1; um2 = um1; vm2 = vm1;
      xm1.r = xq[0]; xm1.u = xq[1]; xm1.v = xq[2];
      um1 = u0; vm1 = v0;
      continue;
    } else {
      // ---- lam window: rows m-1 (la[2]), m (la[1]), m+1 (la[0]); prefetch m+2 ----
      float nl[3];
      load_lam(r, nl);
      // ---- S^T at row m ----
      const bool in_box = in_box_col && m >= d.by0 && m < d.by0 + d.bny;
      float g0 = la[1][0];
      if (in_box) {
        g0 += (float)(2.0 * dt) * xq[0];
        jacc += (double)xq[0] * (double)xq[0];
      }
      const bool cint = col_int && m >= 2 && m <= ny - 3;
      const bool ilo = col_int && m - 1 >= 2 && m - 1 <= ny - 3;
      const bool ihi = col_int && m + 1 >= 2 && m + 1 <= ny - 3;
      const float mu1 = cint ? la[1][1] : 0.0f, mu2 = cint ? la[1][2] : 0.0f;
      const float cc = -2.0f * (ddx2 + ddy2);
      const float gU = cc * mu1 + ddx2 * (shr1(mu1) + shl1(mu1)) +
                       ddy2 * ((ilo ? la[2][1] : 0.0f) + (ihi ? la[0][1] : 0.0f));
      const float gV = cc * mu2 + ddx2 * (shr1(mu2) + shl1(mu2)) +
                       ddy2 * ((ilo ? la[2][2] : 0.0f) + (ihi ? la[0][2] : 0.0f));
      const float irho = fvad::rcp(1.0f + xq[0]);
      const float u = xq[1] * irho, v = xq[2] * irho;
      float lxy[3];
      lxy[0] = g0 - (gU * u + gV * v) * irho;
      lxy[1] = la[1][1] + gU * irho;
      lxy[2] = la[1][2] + gV * irho;

      // ---- X^T at row m ----
      // gP(k) = -dtx lam_xy(k) on interior k; gM(k) = -dtx lam_xy(k-1) on interior k-1
      const bool rint = m >= 2 && m <= ny - 3;
      const bool own = rint && col_int;
      const bool lft = rint && (i - 1 >= 2) && (i - 1 <= nx - 3);
      float gP[3], gM[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gP[c] = own ? -dtx * lxy[c] : 0.0f;
        const float ll = shr1(lxy[c]);
        gM[c] = lft ? -dtx * ll : 0.0f;
      }
      float gW[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gs[3] = {0.f, 0.f, 0.f}, gd[3];
      fvad::limiter_bw(w.W, w.s, lx, dtx, gP, gM, gW, gs, gd);
      // the dot products W_p(k) . W_p(k +- 1) of the limiters of k-1, k, k+1
      const int nb = lx.nb, nbl = shr1i(nb), nbr = shl1i(nb);
      float gdl[3], gdr[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) { gdl[p] = shr1(gd[p]); gdr[p] = shl1(gd[p]); }
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const bool right = (nb >> p) & 1;
        // own limiter: its upwind neighbour
        const float go = gd[p];
        // limiter of k-1 used W(k) when its upwind side was +1
        const float gl = ((nbl >> p) & 1) ? gdl[p] : 0.0f;
        // limiter of k+1 used W(k) when its upwind side was -1
        const float gr = ((nbr >> p) & 1) ? 0.0f : gdr[p];
        if (p < 2) {
          const int b = 3 * p;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            gW[b + c] += go * (right ? Wr[b + c] : Wl[b + c]) + gl * Wl[b + c] + gr * Wr[b + c];
        } else {
          gW[6] += go * (right ? Wr[6] : Wl[6]) + gl * Wl[6] + gr * Wr[6];
        }
      }
      float gL[3] = {0.f, 0.f, 0.f}, gR[3] = {0.f, 0.f, 0.f};
      fvad::riemann_bw(w, beta, gW, gs, gL, gR);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gL[c] = iface_x ? gL[c] : 0.0f;
        gR[c] = iface_x ? gR[c] : 0.0f;
      }
      float ly[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) ly[c] = lxy[c] + gR[c] + shl1(gL[c]);
      if (out_lane) {
        const int o = i + m * nx;
        d.lamx[o] = ly[0];
        d.lamx[n + o] = ly[1];
        d.lamx[2 * n + o] = ly[2];
      }
      la[2][0] = la[1][0]; la[2][1] = la[1][1]; la[2][2] = la[1][2];
      la[1][0] = la[0][0]; la[1][1] = la[0][1]; la[1][2] = la[0][2];
      la[0][0] = nl[0]; la[0][1] = nl[1]; la[0][2] = nl[2];
    }
  }
  if (kRev) {
    // J of this step: the wave's box cells, one double atomic per wave
    const bool any = __any(jacc != 0.0);
    if (any) {
      double v = out_lane ? jacc : 0.0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) atomicAdd(&d.jstep[k], dt * v);
    }
  }
}

// ------------------------------------------------------------ y march --
// lam = B^T Y^T lam_y (d.lamx -> d.lam), drive gradient of the step.
__global__ __launch_bounds__(64) void adj_march_y(const AdjDesc* __restrict__ ds,
                                                  const AdjItem* __restrict__ items, int k,
                                                  int slot) {
  constexpr int kHalo = 2;
  const AdjItem it = items[blockIdx.x];
  const AdjDesc& d = ds[it.inst];
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const float* q = d.seg + (size_t)slot * 3 * n;
  const double dt = d.dt_log[k];
  const double T = d.T_log[k];
  const float dty = (float)(dt / d.dy);
  const float beta = d.beta;
  const int lane = threadIdx.x;
  const int i = it.x0 - kHalo + lane;
  const bool col_int = i >= 2 && i <= nx - 3;
  const bool out_lane = lane >= kHalo && lane < kLanes - kHalo && i < nx;
  const int y0 = it.y0, y1 = it.y1;
  const Loader load = make_loader(d, q, T, i);
  const int bc = d.bc;

  // windows: q_b rows r..r-3; y interfaces r..r-3; limiter adjoint of r-2, r-3
  Cell b0{}, b1{}, b2{}, b3{};
  YIf f0{}, f1{}, f2{}, f3{};
  float gWo2[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gs2[3] = {0.f, 0.f, 0.f};
  float gd2[3] = {0.f, 0.f, 0.f}, gd3[3] = {0.f, 0.f, 0.f};
  int nb2 = 0, nb3 = 0;
  float gRp[3] = {0.f, 0.f, 0.f};   // gR of interface r-3 (cell r-3)
  // lam_y rows r (prefetched), r-1, r-2, r-3
  float ly[4][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  // held lam_b rows: 0 and 1 (bottom rule, for row 2), ny-3 (waits for the
  // top rule) and the sum of rows ny-2 and ny-1 (top rule, for row ny-3);
  // only the bottom and top items of a strip ever fill them
  float lb0[3] = {0.f, 0.f, 0.f}, lb1[3] = {0.f, 0.f, 0.f};
  float lbt[3] = {0.f, 0.f, 0.f}, ltop[3] = {0.f, 0.f, 0.f};

  auto load_ly = [&](int row, float out[3]) {
    if (row < 0 || row >= ny || i < 0 || i >= nx) {
      out[0] = out[1] = out[2] = 0.0f;
      return;
    }
    const int o = i + row * nx;
    out[0] = d.lamx[o]; out[1] = d.lamx[n + o]; out[2] = d.lamx[2 * n + o];
  };

  // B^T of row j from lam_b(j) (fields rho', rho u, rho v) and the held rows
  auto emit = [&](int j, const float lbj[3]) {
    const bool ghost = i <= 1 || j <= 1 || i >= nx - 2 || j >= ny - 2;
    float v[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const float l1 = shr1(lbj[f]);                // lane i-1
      const float l2 = shr1(l1);                    // lane i-2
      const float r1 = shl1(lbj[f]);                // lane i+1
      const float r2 = shl1(r1);                    // lane i+2
      float a = ghost ? 0.0f : lbj[f];
      if (bc != 2) {
        if (i == 2 && bc == 0) a += l2 + l1;                               // left rule
        if (j == 2 && i >= 2) a += lb0[f] + lb1[f];                        // bottom rule
        if (j == ny - 3 && i >= 2) a += ltop[f];                           // top rule
        if (i == nx - 3 && j >= 2 && j <= ny - 3) a += r1 + r2;            // right rule
      }
      v[f] = a;
    }
    if (out_lane) {
      const int o = i + j * nx;
      d.lam[o] = v[0];
      d.lam[n + o] = v[1];
      d.lam[2 * n + o] = v[2];
    }
    // drive parameters: the left ghost cells hold p, p(p+1), 0 with
    // p = amp sin(omega T + phase)
    if (bc == 1 && it.x0 == 0) {
      const double th = d.omega * T + (double)d.edge_phase[j];
      const double sn = sin(th), cs = cos(th);
      const double p = (double)d.edge_amp[j] * sn;
      const bool gl = i >= 0 && i <= 1;
      const double t = gl ? (double)lbj[0] + (double)lbj[1] * (2.0 * p + 1.0) : 0.0;
      const double g = __shfl(t, kHalo, 64) + __shfl(t, kHalo + 1, 64);
      if (lane == 0) {
        d.grad_amp[j] += g * sn;
        d.grad_phase[j] += g * (double)d.edge_amp[j] * cs;
      }
    }
  };

  Cell nx0 = load(y0 - 3);     // q_b rows are fetched one iteration ahead
  for (int r = y0 - 3; r <= y1 + 2; ++r) {
    b3 = b2; b2 = b1; b1 = b0;
    b0 = nx0;
    nx0 = load(r + 1);
    f3 = f2; f2 = f1; f1 = f0;
    if (r >= y0 - 2) y_waves(b1, b0, beta, f0);
    // lam_y window: ly[3] = row r-3, ly[2] = r-2, ly[1] = r-1, ly[0] = r
    // (row r is first used one iteration later)
    {
      float nl[3];
      load_ly(r, nl);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ly[3][c] = ly[2][c]; ly[2][c] = ly[1][c]; ly[1][c] = ly[0][c]; ly[0][c] = nl[c];
      }
    }
    if (r < y0) continue;

    // ---- limiter and its adjoint at interface r-1 ----
    float gWo1[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gs1[3] = {0.f, 0.f, 0.f}, gd1[3];
    int nb1;
    {
      fvad::Lim l;
      float P[3], M[3];
      fvad::limiter_fw(f2.W, f1.W, f1.s, f0.W, dty, l, P, M);
      const int j = r - 1;
      const bool own = col_int && j >= 2 && j <= ny - 3;           // P(j) updates cell j
      const bool blw = col_int && j - 1 >= 2 && j - 1 <= ny - 3;   // M(j) updates cell j-1
      // y component order (rho', rho v, rho u) -> lam_y fields (0, 2, 1)
      float gP[3], gM[3];
      gP[0] = own ? -dty * ly[1][0] : 0.0f;
      gP[1] = own ? -dty * ly[1][2] : 0.0f;
      gP[2] = own ? -dty * ly[1][1] : 0.0f;
      gM[0] = blw ? -dty * ly[2][0] : 0.0f;
      gM[1] = blw ? -dty * ly[2][2] : 0.0f;
      gM[2] = blw ? -dty * ly[2][1] : 0.0f;
      fvad::limiter_bw(f1.W, f1.s, l, dty, gP, gM, gWo1, gs1, gd1);
      nb1 = l.nb;
    }
    if (r >= y0 + 2) {
      // ---- total adjoint of the waves of interface r-2, then its Riemann^T ----
      float gW[7], gs[3];
#pragma unroll
      for (int c = 0; c < 7; ++c) gW[c] = gWo2[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) gs[c] = gs2[c];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const bool right = (nb2 >> p) & 1;
        const float go = gd2[p];
        const float gl = ((nb3 >> p) & 1) ? gd3[p] : 0.0f;   // limiter of r-3 looked up (+1)
        const float gr = ((nb1 >> p) & 1) ? 0.0f : gd1[p];   // limiter of r-1 looked down (-1)
        if (p < 2) {
          const int b = 3 * p;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            gW[b + c] += go * (right ? f1.W[b + c] : f3.W[b + c]) + gl * f3.W[b + c] +
                         gr * f1.W[b + c];
        } else {
          gW[6] += go * (right ? f1.W[6] : f3.W[6]) + gl * f3.W[6] + gr * f1.W[6];
        }
      }
      fvad::RW w;
      fvad::riemann_fw(b3.r, b3.v, b3.u, b2.r, b2.v, b2.u, beta, w);
      float gL[3] = {0.f, 0.f, 0.f}, gR[3] = {0.f, 0.f, 0.f};
      const int j = r - 2;
      if (j >= 1 && j <= ny - 1) fvad::riemann_bw(w, beta, gW, gs, gL, gR);
      if (r >= y0 + 3) {
        // lam_b(r-3) = lam_y(r-3) + gR(interface r-3) + gL(interface r-2)
        const int jj = r - 3;
        float lb[3];
        lb[0] = ly[3][0] + gRp[0] + gL[0];
        lb[2] = ly[3][2] + gRp[1] + gL[1];
        lb[1] = ly[3][1] + gRp[2] + gL[2];
        if (jj == 0) { lb0[0] = lb[0]; lb0[1] = lb[1]; lb0[2] = lb[2]; }
        if (jj == 1) { lb1[0] = lb[0]; lb1[1] = lb[1]; lb1[2] = lb[2]; }
        if (jj == ny - 3) { lbt[0] = lb[0]; lbt[1] = lb[1]; lbt[2] = lb[2]; }
        if (jj >= ny - 2) { ltop[0] += lb[0]; ltop[1] += lb[1]; ltop[2] += lb[2]; }
        if (jj != ny - 3) emit(jj, lb);
        // row ny-3 waited for the top ghost rows ny-2 and ny-1
        if (jj == ny - 1) emit(ny - 3, lbt);
      }
      gRp[0] = gR[0]; gRp[1] = gR[1]; gRp[2] = gR[2];
    }
    // shift the limiter-adjoint window
#pragma unroll
    for (int c = 0; c < 3; ++c) { gd3[c] = gd2[c]; gd2[c] = gd1[c]; }
    nb3 = nb2; nb2 = nb1;
#pragma unroll
    for (int c = 0; c < 7; ++c) gWo2[c] = gWo1[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) gs2[c] = gs1[c];
  }
}

}  // namespace fv

// ---------------------------------------------------------------- launches --

extern "C" hipError_t fv_adj_march_forward(const fv::AdjDesc* ds, const fv::AdjItem* items,
                                           int n_items, int k, int slot, hipStream_t s) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(fv::adj_march_x<false>, dim3(n_items), dim3(64), 0, s, ds, items, k, slot);
  return hipGetLastError();
}

extern "C" hipError_t fv_adj_march_reverse(const fv::AdjDesc* ds, const fv::AdjItem* items_x,
                                           int n_items_x, const fv::AdjItem* items_y,
                                           int n_items_y, int k, int slot, hipStream_t s) {
  if (n_items_x <= 0 || n_items_y <= 0) return hipSuccess;
  hipLaunchKernelGGL(fv::adj_march_x<true>, dim3(n_items_x), dim3(64), 0, s, ds, items_x, k,
                     slot);
  hipLaunchKernelGGL(fv::adj_march_y, dim3(n_items_y), dim3(64), 0, s, ds, items_y, k, slot);
  return hipGetLastError();
}

extern "C" hipError_t fv_adj_march_occupancy(int* x_fwd, int* x_rev, int* y_rev) {
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(x_fwd, fv::adj_march_x<false>, 64, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(x_rev, fv::adj_march_x<true>, 64, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(y_rev, fv::adj_march_y, 64, 0);
  return e;
}
