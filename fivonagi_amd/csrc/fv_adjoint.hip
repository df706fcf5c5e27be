// fv_adjoint.hip -- discrete adjoint of the FiVoNAGI step (gfx950).
//
// The reference has no adjoint.  This differentiates the step of
// fv_step.hip (boundary B, y sweep Y, x sweep X, source S; reference:
// common/main.cu:310-635, apps/na/fv.h, apps/na/source.h) along the executed
// branch path with the dt sequence frozen (the dt of every accepted step of
// the forward run is an input).  Objective J = sum_n dt_n sum_box rho'_{n+1}^2;
// parameters: the left-edge drive amplitude and phase per row, and the
// initial state (the final adjoint).
//
// One adjoint step is a chain of grid-wide passes, each launched over every
// instance at once (blockIdx.y = instance):
//   recompute: B -> (Riemann, limiter, update) in y -> the same in x
//   reverse:   S^T -> X^T (limiter^T, Riemann^T, gather) -> Y^T -> B^T
// The sweeps are per-interface passes with the interface data (waves, speeds,
// adjoint of the waves) in SoA scratch grids indexed like the reference's
// s/W/amdq grids (offset of the + cell).  The states of a segment of K steps
// are recomputed from a checkpoint before the segment is reversed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../apps/na/adjoint.h"
#include "fv_adjoint.h"

namespace fv {

__device__ __forceinline__ bool interior(int i, int j, int nx, int ny) {
  return i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
}

// thread -> cell of instance blockIdx.y; false when outside the grid
__device__ __forceinline__ bool cell_of(const AdjDesc& d, int& o, int& i, int& j) {
  o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= d.nx * d.ny) return false;
  j = o / d.nx;
  i = o - j * d.nx;
  return true;
}

// field order (rho', normal, tangential) -> state field index
__device__ __forceinline__ int fld(int axis, int c) {
  return c == 0 ? 0 : (axis == 1 ? (c == 1 ? 2 : 1) : (c == 1 ? 1 : 2));
}

// ---------------------------------------------------------------- forward --

// B: ghost rules of common/boundary.h:27-81 / apps/hifu-beam/boundary.h:77-93 /
// apps/taylor-angle-xy/boundary.h:38-55 applied to state `slot`
__global__ __launch_bounds__(256) void adj_bnd_fw(const AdjDesc* __restrict__ ds, int k,
                                                  int slot) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const float* q = d.seg + (size_t)slot * 3 * n;
  const double T = d.T_log[k];
  float* b = d.qb;
  const bool ghost = i <= 1 || j <= 1 || i >= nx - 2 || j >= ny - 2;
  if (d.bc == 2) {
    if (ghost) {
      const double arg = d.tay_c * ((double)i * d.dx * d.tay_cos +
                                    ((double)(j - d.tay_jc) * d.dy + d.tay_yoff) * d.tay_sin - T);
      const double r = d.tay_rhoa * tanh(arg);
      b[o] = (float)r; b[n + o] = (float)(r * d.tay_cos); b[2 * n + o] = (float)(r * d.tay_sin);
    } else {
      b[o] = q[o]; b[n + o] = q[n + o]; b[2 * n + o] = q[2 * n + o];
    }
    return;
  }
  int si = i, sj = j;
  if (i <= 1) {
    if (d.bc == 1) {
      const double th = d.omega * T + (double)d.edge_phase[j];
      const float p = (float)((double)d.edge_amp[j] * sin(th));
      b[o] = p; b[n + o] = p * (p + 1.0f); b[2 * n + o] = 0.0f;
      return;
    }
    si = 2;
  } else if (j >= ny - 2) {
    sj = ny - 3;
  } else if (j <= 1) {
    sj = 2;
  } else if (i >= nx - 2) {
    si = nx - 3;
  }
  const int os = si + sj * nx;
  b[o] = q[os]; b[n + o] = q[n + os]; b[2 * n + o] = q[2 * n + os];
}

// Riemann solve at every interface m >= 1 along `axis` (0 = x, 1 = y) -> iw
__global__ __launch_bounds__(256) void adj_waves(const AdjDesc* __restrict__ ds, int axis,
                                                 int src_sel) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int n = d.nx * d.ny, m = axis ? j : i, st = axis ? d.nx : 1;
  const float* q = src_sel == 0 ? d.qb : d.qy;
  float* iw = d.iw;
  if (m < 1) {
    for (int c = 0; c < 10; ++c) iw[(size_t)c * n + o] = 0.0f;
    return;
  }
  const int ol = o - st;
  fvad::RW w;
  fvad::riemann_fw(q[ol], q[(size_t)fld(axis, 1) * n + ol], q[(size_t)fld(axis, 2) * n + ol],
                   q[o], q[(size_t)fld(axis, 1) * n + o], q[(size_t)fld(axis, 2) * n + o], d.beta,
                   w);
  for (int c = 0; c < 7; ++c) iw[(size_t)c * n + o] = w.W[c];
  for (int c = 0; c < 3; ++c) iw[(size_t)(7 + c) * n + o] = w.s[c];
}

__device__ __forceinline__ void load_w(const float* iw, int n, int o, float W[7], float s[3]) {
  for (int c = 0; c < 7; ++c) W[c] = iw[(size_t)c * n + o];
  for (int c = 0; c < 3; ++c) s[c] = iw[(size_t)(7 + c) * n + o];
}

// limiter + folded flux P, M at interfaces 2 <= m <= L-2 -> pm
__global__ __launch_bounds__(256) void adj_flux(const AdjDesc* __restrict__ ds, int axis, int k) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int n = d.nx * d.ny, m = axis ? j : i, L = axis ? d.ny : d.nx, st = axis ? d.nx : 1;
  if (m < 2 || m > L - 2) return;
  const double dt = d.dt_log[k];
  const float dtd = (float)(dt / (axis ? d.dy : d.dx));
  float Wm[7], W[7], Wp[7], s[3], sx[3];
  load_w(d.iw, n, o - st, Wm, sx);
  load_w(d.iw, n, o, W, s);
  load_w(d.iw, n, o + st, Wp, sx);
  fvad::Lim l;
  float P[3], M[3];
  fvad::limiter_fw(Wm, W, s, Wp, dtd, l, P, M);
  for (int c = 0; c < 3; ++c) {
    d.pm[(size_t)c * n + o] = P[c];
    d.pm[(size_t)(3 + c) * n + o] = M[c];
  }
}

// q' = q - dtd (P(k) + M(k+1)) on interior cells (src_sel: 0 qb -> qy, 1 qy -> qxy)
__global__ __launch_bounds__(256) void adj_update(const AdjDesc* __restrict__ ds, int axis, int k,
                                                  int src_sel) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int n = d.nx * d.ny, st = axis ? d.nx : 1;
  const float* q = src_sel == 0 ? d.qb : d.qy;
  float* out = src_sel == 0 ? d.qy : d.qxy;
  const float dtd = (float)(d.dt_log[k] / (axis ? d.dy : d.dx));
  const bool upd = interior(i, j, d.nx, d.ny);
  for (int c = 0; c < 3; ++c) {
    const int f = fld(axis, c);
    float v = q[(size_t)f * n + o];
    if (upd) v = v - dtd * (d.pm[(size_t)c * n + o] + d.pm[(size_t)(3 + c) * n + o + st]);
    out[(size_t)f * n + o] = v;
  }
}

// S (apps/na/source.h:84-104): qxy -> state slot+1
__global__ __launch_bounds__(256) void adj_src_fw(const AdjDesc* __restrict__ ds, int k,
                                                  int slot) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, n = nx * d.ny;
  const float* q = d.qxy;
  float* out = d.seg + (size_t)(slot + 1) * 3 * n;
  out[o] = q[o];
  if (!interior(i, j, nx, d.ny)) {
    out[n + o] = q[n + o];
    out[2 * n + o] = q[2 * n + o];
    return;
  }
  const double dt = d.dt_log[k];
  const float ddx2 = (float)(d.delta * dt / (d.dx * d.dx));
  const float ddy2 = (float)(d.delta * dt / (d.dy * d.dy));
  for (int f = 1; f <= 2; ++f) {
    const float* qf = q + (size_t)f * n;
    const float u0 = qf[o] / (1.0f + q[o]);
    const float uW = qf[o - 1] / (1.0f + q[o - 1]), uE = qf[o + 1] / (1.0f + q[o + 1]);
    const float uS = qf[o - nx] / (1.0f + q[o - nx]), uN = qf[o + nx] / (1.0f + q[o + nx]);
    out[(size_t)f * n + o] = qf[o] + ddx2 * (uW - 2.0f * u0 + uE) + ddy2 * (uS - 2.0f * u0 + uN);
  }
}

// J contribution of step k: dt_k * sum_box rho'_{k+1}^2 (one block per instance)
__global__ __launch_bounds__(256) void adj_jbox(const AdjDesc* __restrict__ ds, int k, int slot) {
  const AdjDesc& d = ds[blockIdx.y];
  const int n = d.nx * d.ny;
  const float* q1 = d.seg + (size_t)(slot + 1) * 3 * n;
  double acc = 0.0;
  const int cnt = d.bnx * d.bny;
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int bj = t / d.bnx, bi = t - bj * d.bnx;
    const float v = q1[(d.by0 + bj) * d.nx + d.bx0 + bi];
    acc += (double)v * (double)v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.jstep[k] = d.dt_log[k] * red[0];
}

// ---------------------------------------------------------------- reverse --

// S^T with the objective term: lam2 = S^T (lam + dJ/dq_{k+1})
__global__ __launch_bounds__(256) void adj_src_bw(const AdjDesc* __restrict__ ds, int k,
                                                  int slot) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const double dt = d.dt_log[k];
  const float* lam = d.lam;
  float g1 = lam[o];
  if (i >= d.bx0 && i < d.bx0 + d.bnx && j >= d.by0 && j < d.by0 + d.bny) {
    const float* q1 = d.seg + (size_t)(slot + 1) * 3 * n;
    g1 += (float)(2.0 * dt) * q1[o];
  }
  const float ddx2 = (float)(d.delta * dt / (d.dx * d.dx));
  const float ddy2 = (float)(d.delta * dt / (d.dy * d.dy));
  const float rho = 1.0f + d.qxy[o];
  const float u = d.qxy[n + o] / rho, v = d.qxy[2 * n + o] / rho;
  float gU = 0.0f, gV = 0.0f;
  // transposed 5-point stencil over interior centres c
  const int cs[5] = {o, o - 1, o + 1, o - nx, o + nx};
  const int ci[5] = {i, i - 1, i + 1, i, i};
  const int cj[5] = {j, j, j, j - 1, j + 1};
  const float cf[5] = {-2.0f * (ddx2 + ddy2), ddx2, ddx2, ddy2, ddy2};
  for (int t = 0; t < 5; ++t) {
    if (ci[t] < 0 || ci[t] >= nx || cj[t] < 0 || cj[t] >= ny) continue;
    if (!interior(ci[t], cj[t], nx, ny)) continue;
    gU += cf[t] * lam[n + cs[t]];
    gV += cf[t] * lam[2 * n + cs[t]];
  }
  d.lam2[o] = g1 - (gU * u + gV * v) / rho;
  d.lam2[n + o] = lam[n + o] + gU / rho;
  d.lam2[2 * n + o] = lam[2 * n + o] + gV / rho;
}

// sweep^T part 1: gP, gM -> limiter^T at interfaces 2..L-2 -> ig, inb
// lam_sel: 0 = lam2 (x sweep output adjoint), 1 = lamx (y sweep output adjoint)
__global__ __launch_bounds__(256) void adj_flux_bw(const AdjDesc* __restrict__ ds, int axis, int k,
                                                   int lam_sel) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const int m = axis ? j : i, L = axis ? ny : nx, st = axis ? nx : 1;
  float* ig = d.ig;
  if (m < 2 || m > L - 2) {
    for (int c = 0; c < 13; ++c) ig[(size_t)c * n + o] = 0.0f;
    d.inb[o] = 0;
    return;
  }
  const float* lo = lam_sel == 0 ? d.lam2 : d.lamx;
  const float dtd = (float)(d.dt_log[k] / (axis ? d.dy : d.dx));
  float gP[3] = {0.f, 0.f, 0.f}, gM[3] = {0.f, 0.f, 0.f};
  // P(m) updates cell m; M(m) updates cell m-1
  const int im = axis ? i : i - 1, jm = axis ? j - 1 : j;
  const bool own = interior(i, j, nx, ny), left = interior(im, jm, nx, ny);
  for (int c = 0; c < 3; ++c) {
    const int f = fld(axis, c);
    if (own) gP[c] = -dtd * lo[(size_t)f * n + o];
    if (left) gM[c] = -dtd * lo[(size_t)f * n + o - st];
  }
  float Wm[7], W[7], Wp[7], s[3], sx[3];
  load_w(d.iw, n, o - st, Wm, sx);
  load_w(d.iw, n, o, W, s);
  load_w(d.iw, n, o + st, Wp, sx);
  fvad::Lim l;
  float P[3], M[3];
  fvad::limiter_fw(Wm, W, s, Wp, dtd, l, P, M);
  float gW[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gs[3] = {0.f, 0.f, 0.f}, gd[3];
  fvad::limiter_bw(W, s, l, dtd, gP, gM, gW, gs, gd);
  for (int c = 0; c < 7; ++c) ig[(size_t)c * n + o] = gW[c];
  for (int c = 0; c < 3; ++c) ig[(size_t)(7 + c) * n + o] = gs[c];
  for (int c = 0; c < 3; ++c) ig[(size_t)(10 + c) * n + o] = gd[c];
  d.inb[o] = l.nb;
}

__device__ __forceinline__ void add_wave(float gW[7], const float* iw, int n, int o, int p, float g) {
  if (p == 0) { for (int c = 0; c < 3; ++c) gW[c] += g * iw[(size_t)c * n + o]; }
  else if (p == 1) { for (int c = 0; c < 3; ++c) gW[3 + c] += g * iw[(size_t)(3 + c) * n + o]; }
  else { gW[6] += g * iw[(size_t)6 * n + o]; }
}

// sweep^T part 2: total wave adjoint (own + the dot products shared with the
// neighbours), then Riemann^T -> iq (gL, gR); src_sel: 0 qb, 1 qy
__global__ __launch_bounds__(256) void adj_riemann_bw(const AdjDesc* __restrict__ ds, int axis,
                                                      int src_sel) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const int m = axis ? j : i, L = axis ? ny : nx, st = axis ? nx : 1;
  float* iq = d.iq;
  if (m < 1) {
    for (int c = 0; c < 6; ++c) iq[(size_t)c * n + o] = 0.0f;
    return;
  }
  const float* ig = d.ig;
  float gW[7], gs[3];
  for (int c = 0; c < 7; ++c) gW[c] = ig[(size_t)c * n + o];
  for (int c = 0; c < 3; ++c) gs[c] = ig[(size_t)(7 + c) * n + o];
  const int nb = d.inb[o];
  for (int p = 0; p < 3; ++p) {
    // own limiter: dot_p(m) = W_p(m) . W_p(m +- 1)
    const float g = ig[(size_t)(10 + p) * n + o];
    if (g != 0.0f) add_wave(gW, d.iw, n, (nb >> p) & 1 ? o + st : o - st, p, g);
    // limiter of m-1 dotted with m (its upwind neighbour is +1)
    if (m - 1 >= 1 && ((d.inb[o - st] >> p) & 1)) {
      const float gm = ig[(size_t)(10 + p) * n + o - st];
      if (gm != 0.0f) add_wave(gW, d.iw, n, o - st, p, gm);
    }
    // limiter of m+1 dotted with m (its upwind neighbour is -1)
    if (m + 1 <= L - 1 && !((d.inb[o + st] >> p) & 1)) {
      const float gp = ig[(size_t)(10 + p) * n + o + st];
      if (gp != 0.0f) add_wave(gW, d.iw, n, o + st, p, gp);
    }
  }
  const float* q = src_sel == 0 ? d.qb : d.qy;
  const int ol = o - st;
  fvad::RW w;
  fvad::riemann_fw(q[ol], q[(size_t)fld(axis, 1) * n + ol], q[(size_t)fld(axis, 2) * n + ol],
                   q[o], q[(size_t)fld(axis, 1) * n + o], q[(size_t)fld(axis, 2) * n + o], d.beta,
                   w);
  float gL[3] = {0.f, 0.f, 0.f}, gR[3] = {0.f, 0.f, 0.f};
  fvad::riemann_bw(w, d.beta, gW, gs, gL, gR);
  for (int c = 0; c < 3; ++c) {
    iq[(size_t)c * n + o] = gL[c];
    iq[(size_t)(3 + c) * n + o] = gR[c];
  }
}

// sweep^T part 3: lam_in = lam_out + gR(own interface) + gL(interface above)
// sel 0: x sweep (lam2 -> lamx); sel 1: y sweep (lamx -> lam2)
__global__ __launch_bounds__(256) void adj_gather(const AdjDesc* __restrict__ ds, int axis, int sel) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int n = d.nx * d.ny, m = axis ? j : i, L = axis ? d.ny : d.nx, st = axis ? d.nx : 1;
  const float* lo = sel == 0 ? d.lam2 : d.lamx;
  float* li = sel == 0 ? d.lamx : d.lam2;
  for (int c = 0; c < 3; ++c) {
    const int f = fld(axis, c);
    float v = lo[(size_t)f * n + o];
    if (m >= 1) v += d.iq[(size_t)(3 + c) * n + o];
    if (m + 1 <= L - 1) v += d.iq[(size_t)c * n + o + st];
    li[(size_t)f * n + o] = v;
  }
}

// drive parameters: gradient of the ghost-column values (before B^T, from lam2)
__global__ __launch_bounds__(256) void adj_theta(const AdjDesc* __restrict__ ds, int k) {
  const AdjDesc& d = ds[blockIdx.y];
  if (d.bc != 1) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.ny) return;
  const int nx = d.nx, n = nx * d.ny;
  const double th = d.omega * d.T_log[k] + (double)d.edge_phase[j];
  const double sn = sin(th), cs = cos(th);
  const double p = (double)d.edge_amp[j] * sn;
  double gp = 0.0;
  for (int i = 0; i <= 1; ++i) {
    const int o = i + j * nx;
    gp += (double)d.lam2[o] + (double)d.lam2[n + o] * (2.0 * p + 1.0);
  }
  d.grad_amp[j] += gp * sn;
  d.grad_phase[j] += gp * (double)d.edge_amp[j] * cs;
}

// B^T (gather form of the copy rules): lam = B^T lam2
__global__ __launch_bounds__(256) void adj_bnd_bw(const AdjDesc* __restrict__ ds) {
  const AdjDesc& d = ds[blockIdx.y];
  int o, i, j;
  if (!cell_of(d, o, i, j)) return;
  const int nx = d.nx, ny = d.ny, n = nx * ny;
  const float* lb = d.lam2;
  const bool ghost = i <= 1 || j <= 1 || i >= nx - 2 || j >= ny - 2;
  for (int f = 0; f < 3; ++f) {
    const float* g = lb + (size_t)f * n;
    float v = ghost ? 0.0f : g[o];
    if (d.bc != 2) {
      if (i == 2 && d.bc == 0) v += g[j * nx] + g[1 + j * nx];                 // left rule
      if (j == 2 && i >= 2) v += g[i] + g[i + nx];                              // bottom
      if (j == ny - 3 && i >= 2) v += g[i + (ny - 2) * nx] + g[i + (ny - 1) * nx];  // top
      if (i == nx - 3 && j >= 2 && j <= ny - 3) v += g[nx - 2 + j * nx] + g[nx - 1 + j * nx];
    }
    d.lam[(size_t)f * n + o] = v;
  }
}

}  // namespace fv

// ---------------------------------------------------------------- launches --

namespace {
inline dim3 grid_for(int max_cells, int n_inst) { return dim3((max_cells + 255) / 256, n_inst); }
}

extern "C" hipError_t fv_adj_forward_step(const fv::AdjDesc* ds, int n_inst, int max_cells, int k,
                                          int slot, int store_next, hipStream_t s) {
  const dim3 g = grid_for(max_cells, n_inst), b(256);
  hipLaunchKernelGGL(fv::adj_bnd_fw, g, b, 0, s, ds, k, slot);
  hipLaunchKernelGGL(fv::adj_waves, g, b, 0, s, ds, 1, 0);
  hipLaunchKernelGGL(fv::adj_flux, g, b, 0, s, ds, 1, k);
  hipLaunchKernelGGL(fv::adj_update, g, b, 0, s, ds, 1, k, 0);
  hipLaunchKernelGGL(fv::adj_waves, g, b, 0, s, ds, 0, 1);
  hipLaunchKernelGGL(fv::adj_flux, g, b, 0, s, ds, 0, k);
  hipLaunchKernelGGL(fv::adj_update, g, b, 0, s, ds, 0, k, 1);
  if (store_next) hipLaunchKernelGGL(fv::adj_src_fw, g, b, 0, s, ds, k, slot);
  return hipGetLastError();
}

// reverse of step k whose input state is `slot` (intermediates of
// fv_adj_forward_step(k, slot) must be current: qb, qy, qxy, iw of the x sweep)
extern "C" hipError_t fv_adj_reverse_step(const fv::AdjDesc* ds, int n_inst, int max_cells,
                                          int max_ny, int k, int slot, hipStream_t s) {
  const dim3 g = grid_for(max_cells, n_inst), b(256);
  hipLaunchKernelGGL(fv::adj_jbox, dim3(1, n_inst), b, 0, s, ds, k, slot);
  hipLaunchKernelGGL(fv::adj_src_bw, g, b, 0, s, ds, k, slot);
  hipLaunchKernelGGL(fv::adj_flux_bw, g, b, 0, s, ds, 0, k, 0);
  hipLaunchKernelGGL(fv::adj_riemann_bw, g, b, 0, s, ds, 0, 1);
  hipLaunchKernelGGL(fv::adj_gather, g, b, 0, s, ds, 0, 0);
  hipLaunchKernelGGL(fv::adj_waves, g, b, 0, s, ds, 1, 0);
  hipLaunchKernelGGL(fv::adj_flux_bw, g, b, 0, s, ds, 1, k, 1);
  hipLaunchKernelGGL(fv::adj_riemann_bw, g, b, 0, s, ds, 1, 0);
  hipLaunchKernelGGL(fv::adj_gather, g, b, 0, s, ds, 1, 1);
  hipLaunchKernelGGL(fv::adj_theta, dim3((max_ny + 255) / 256, n_inst), b, 0, s, ds, k);
  hipLaunchKernelGGL(fv::adj_bnd_bw, g, b, 0, s, ds);
  return hipGetLastError();
}
