// apps/common/boundary.h -- ghost-cell rules as device code.  The fused step
// kernel evaluates the boundary while it loads a row of the input state, so
// the ghost ring costs no separate pass.
//
// Rules restated (reads come from the input state only, like the reference,
// which reads the other texture -- so corner ghosts lag one step):
//   zero-gradient ..... common/boundary.h:27-81; later rules win:
//                       right (i >= nx-2) -> (nx-3, j), bottom (j <= 1) -> (i, 2),
//                       top (j >= ny-2) -> (i, ny-3), left (i <= 1) -> (2, j)
//   edge drive ........ apps/hifu-beam/boundary.h:77-93 with nu = (1, 0):
//                       left ghost columns hold p, p*(p+1), 0 with
//                       p = amp_j * sin(omega*T + phase_j); other edges zero-gradient
//   Taylor shock ...... apps/taylor-angle-xy/boundary.h:38-55: every ghost cell
//                       holds the closed form rho' = rhoa*tanh(c*(x cos + y sin - T))
//   HIFU app .......... apps/hifu-beam/boundary.h:20-133: focused packet on the
//                       aperture rows of the left ghost columns, mirrored beam
//                       axis (rows 0, 1 <- 4, 3 with rho v negated), zero corners
#ifndef FIVONAGI_APPS_COMMON_BOUNDARY_H
#define FIVONAGI_APPS_COMMON_BOUNDARY_H

#include <hip/hip_runtime.h>

#include "../na/solver.h"

namespace fvbc {

struct Source {       // where a ghost cell's value comes from
  int ii, jj;         // cell of the input state to copy
  bool drive;         // left-edge drive value instead of a copy
  bool closed_form;   // Taylor closed form instead of a copy
};

__device__ __forceinline__ bool is_ghost(int i, int j, int nx, int ny) {
  return i <= 1 || j <= 1 || i >= nx - 2 || j >= ny - 2;
}

// kind: 0 zero-gradient, 1 edge drive, 2 Taylor (fv::BoundaryKind)
__device__ __forceinline__ Source map(int kind, int i, int j, int nx, int ny) {
  Source s{i, j, false, false};
  if (kind == 2) {
    s.closed_form = is_ghost(i, j, nx, ny);
  } else if (i <= 1) {
    if (kind == 1) s.drive = true; else s.ii = 2;
  } else if (j >= ny - 2) {
    s.jj = ny - 3;
  } else if (j <= 1) {
    s.jj = 2;
  } else if (i >= nx - 2) {
    s.ii = nx - 3;
  }
  return s;
}

// Left-edge drive: p = amp*sin(theta + phase) = ampc*sin(theta) + amps*cos(theta)
// with ampc = amp*cos(phase), amps = amp*sin(phase) precomputed per row and
// sin/cos(theta = omega*T) computed once per step.
template <typename R>
__device__ __forceinline__ void drive(R ampc, R amps, R sin_t, R cos_t, R& q1, R& q2, R& q3) {
  const R p = fvna::fma_(ampc, sin_t, amps * cos_t);
  q1 = p;
  q2 = p * (p + R(1));
  q3 = R(0);
}

// HIFU focused packet (apps/hifu-beam/boundary.h:10-17, 84-93):
// pack = AMPL sin(ISPW (T + phi)) exp(-opT^10), opT = FREC (T + phi)/4, with
// AMPL sin(ISPW (T + phi)) = ampc sin(ISPW T) + amps cos(ISPW T) from the
// per-row ampc = AMPL cos(ISPW phi), amps = AMPL sin(ISPW phi);
// u1 = pack, (u2, u3) = nu pack (pack + 1).
template <typename R>
__device__ __forceinline__ void hifu_packet(R ampc, R amps, R sin_t, R cos_t, R opT, R nux, R nuy,
                                            R& q1, R& q2, R& q3) {
  const R o2 = opT * opT, o4 = o2 * o2, o8 = o4 * o4;
  const R p = fvna::fma_(ampc, sin_t, amps * cos_t) * exp(-(o8 * o2));
  const R pp = p * (p + R(1));
  q1 = p;
  q2 = nux * pp;
  q3 = nuy * pp;
}

template <typename R>
__device__ __forceinline__ void taylor(double rhoa, double c, double cs, double sn,
                                       double yoff, int jc, double dx, double dy,
                                       double T, int i, int j, R& q1, R& q2, R& q3) {
  const double arg = c * ((double)i * dx * cs + ((double)(j - jc) * dy + yoff) * sn - T);
  const double r = rhoa * tanh(arg);
  q1 = (R)r;
  q2 = (R)(r * cs);
  q3 = (R)(r * sn);
}

}  // namespace fvbc

#endif  // FIVONAGI_APPS_COMMON_BOUNDARY_H
