// apps/taylor-angle-xy/boundary.h -- the Taylor shock's ghost cells
// (apps/taylor-angle-xy/boundary.h:38-55 of the reference): every ghost cell
// holds the closed form
//   rho' = rhoa tanh(c (x cos(theta) + y sin(theta) - T)),
//   (rho u, rho v) = rho' (cos(theta), sin(theta)),
// taken at (i + MDX, j + MDY) (the window's offset).  The values depend on
// the cell and T only, so they are written into the input state before the
// step (prepass, as the reference's boundary kernel) and the step copies every
// cell as it is.
#ifndef FIVONAGI_APPS_TAYLOR_ANGLE_XY_BOUNDARY_H
#define FIVONAGI_APPS_TAYLOR_ANGLE_XY_BOUNDARY_H

#include "../common/boundary.h"

namespace fvbc {

struct Taylor {
  static constexpr int kId = fv::kBcTaylor;
  static constexpr bool kRowData = false;
  static constexpr bool kPrepass = true;
  static constexpr bool kExtOnly = false;
  __device__ static void lane(int ic, int, LaneMap& m, uint32_t size) {
    m.ic = ic;
    m.left = false;
    m.bA = m.bB = size * (uint32_t)ic;
  }
  __device__ static uint32_t offset(int rc, int, const LaneMap& m, uint32_t row_bytes) {
    return (uint32_t)rc * row_bytes + m.bA;
  }
  template <typename R>
  __device__ static void load_row(const Tables&, int, RowData<R>&) {}
  template <typename R>
  __device__ static void patch(const PatchCtx<R>&, int, Cell<R>&, const RowData<R>&) {}
  // the closed form at cell (i, j) of the unmoved frame (i + MDX, j + MDY)
  template <typename R>
  __device__ static void prepass(const fv::InstDesc& D, double T, int i, int j, Cell<R>& q) {
    const double arg = D.tay_c * ((double)i * D.dx * D.tay_cos +
                                  ((double)(j - D.tay_jc) * D.dy + D.tay_yoff) * D.tay_sin - T);
    const double r = D.tay_rhoa * tanh(arg);
    q.q1 = (R)r;
    q.q2 = (R)(r * D.tay_cos);
    q.q3 = (R)(r * D.tay_sin);
  }
};

}  // namespace fvbc

#endif  // FIVONAGI_APPS_TAYLOR_ANGLE_XY_BOUNDARY_H
