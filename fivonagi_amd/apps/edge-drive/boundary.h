// apps/edge-drive/boundary.h -- a left-edge drive through an aperture: the
// pattern of apps/hifu-beam/boundary.h:77-93 with nu = (1, 0) and a free
// amplitude and phase per row,
//   left ghost columns (i <= 1):  p = amp_j sin(omega T + phase_j),
//                                 rho' = p, rho u = p (p + 1), rho v = 0,
// while the window has not moved in x (MDX == 0, boundary.h:82); the other
// edges, and the left columns once the window moves, zero-gradient.
#ifndef FIVONAGI_APPS_EDGE_DRIVE_BOUNDARY_H
#define FIVONAGI_APPS_EDGE_DRIVE_BOUNDARY_H

#include "../common/boundary.h"

namespace fvbc {

// p = amp sin(theta + phase) = ampc sin(theta) + amps cos(theta), with
// ampc = amp cos(phase), amps = amp sin(phase) per row (host tables) and
// sin/cos(theta = omega T) once per attempt.
template <typename R>
__device__ __forceinline__ void drive(R ampc, R amps, R sin_t, R cos_t, Cell<R>& q) {
  const R p = fvna::fma_(ampc, sin_t, amps * cos_t);
  q.q1 = p;
  q.q2 = p * (p + R(1));
  q.q3 = R(0);
}

struct EdgeDrive {
  static constexpr int kId = fv::kBcEdgeDrive;
  static constexpr bool kRowData = true;
  static constexpr bool kPrepass = false;
  static constexpr bool kExtOnly = false;
  __device__ static void lane(int ic, int nx, LaneMap& m, uint32_t size) {
    ZeroGradient::lane(ic, nx, m, size);
  }
  __device__ static uint32_t offset(int rc, int ny, const LaneMap& m, uint32_t row_bytes) {
    return ZeroGradient::offset(rc, ny, m, row_bytes);
  }
  template <typename R>
  __device__ static void load_row(const Tables& tb, int rc, RowData<R>& h) {
    typedef __attribute__((address_space(4))) const R kR;   // read-only: scalar loads
    h.ec = ((kR*)tb.edge_c)[rc];
    h.es = ((kR*)tb.edge_s)[rc];
  }
  template <typename R>
  __device__ static void patch(const PatchCtx<R>& c, int, Cell<R>& q, const RowData<R>& h) {
    // wave-uniform: the drive stops once the window moves, and only the
    // waves holding the left ghost columns compute it (the empty volatile asm
    // keeps this a branch: the compiler would otherwise compute the drive on
    // every wave and select)
    if (c.mdx != 0 || !c.any_left) return;
    asm volatile("");
    Cell<R> d;
    drive(h.ec, h.es, c.sin_t, c.cos_t, d);
    q.q1 = c.left ? d.q1 : q.q1;
    q.q2 = c.left ? d.q2 : q.q2;
    q.q3 = c.left ? d.q3 : q.q3;
  }
};

}  // namespace fvbc

#endif  // FIVONAGI_APPS_EDGE_DRIVE_BOUNDARY_H
