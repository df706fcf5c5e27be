// apps/taylor-angle-xy/data_collect.h -- the Taylor-shock app's data
// collection (apps/taylor-angle-xy/data_collect.h:7-98 of the reference),
// after every accepted step: the numeric and closed-form cuts along the
// propagation direction.
#ifndef FIVONAGI_APPS_TAYLOR_ANGLE_XY_DATA_COLLECT_H
#define FIVONAGI_APPS_TAYLOR_ANGLE_XY_DATA_COLLECT_H

#include "../common/boundary.h"

namespace fvbc {

// Data collection of the taylor-angle-xy app (apps/taylor-angle-xy/
// data_collect.h:43-96) for one accepted attempt: rho' of the new state along
// the propagation direction through the front's centre, interpolated in y
// between the two rows nearest to the line (entry iwrite of row 2), beside the
// closed form at the nearer row (row 3).  The arithmetic follows the
// reference's types: T, dx, dy and the constants in R, the line position ja
// and the interpolation weights in float.  The state is read from the output
// buffer, where the step has already applied the window shift (sx, sy) of this
// attempt: the cell (i, j) of the unshifted frame is (i - sx, j - sy) there.
// Thread tid = entry iwrite of the cut, 0 <= tid <= 2 cut_half.
template <typename R>
__device__ void collect_taylor(const fv::InstDesc& D, const fv::Ctl& P, double Td,
                               R* __restrict__ m, int tid) {
  const int L = D.cut_half;
  if (tid > 2 * L) return;
  const int nx = D.nx, ny = D.ny;
  const R T = R(Td), dx = R(D.dx), dy = R(D.dy);
  const int icenter = (int)(T * R(D.md_vx) / dx) - P.mdx;                  // :43
  const int jcenter = ny / 2 + (int)((T - R(D.md_t)) * R(D.md_vy) / dy) - P.mdy;  // :44-46
  const int i = icenter - L + tid;                  // iwrite = tid (:87-89)
  if (i < 0 || i > nx - 1) return;
  const float ja = (float)fvna::fma_(R((float)(i - icenter)) * (dx / dy), R(D.tan_theta),
                               R((float)jcenter));  // :52-55
  int sx = 0, sy = 0;
  if (D.md_on) {
    sx = md_shift(Td, D.md_t, D.md_vx, D.dx, P.mdx);
    sy = md_shift(Td, D.md_t, D.md_vy, D.dy, P.mdy);
  }
  const R* __restrict__ u = (const R*)D.q[P.cur ^ 1][0];
  // tex2D reads clamp to the grid; the shifted frame starts at (sx, sy)
  auto at = [&](int ii, int jj) -> R {
    jj = min(max(jj, 0), ny - 1);
    const int pi = max(ii - sx, 0), pj = max(jj - sy, 0);
    return u[(size_t)pj * nx + pi];
  };
  // the thread (i, j) of the reference whose row is nearest to ja: j - ja in
  // [0, 1/2] takes rows j, j-1 (:58-68), j - ja in (-1/2, 0) rows j+1, j (:69-79)
  const int jf = (int)floorf(ja);
  for (int j = max(jf - 1, 0); j <= min(jf + 2, ny - 1); ++j) {
    const float dj = (float)j - ja;
    R val;
    if (dj >= 0.0f && dj <= 0.5f) {
      const float djm = (float)(j - 1) - ja;
      val = fvna::fma_(R(-djm), at(i, j), R(dj) * at(i, j - 1));
    } else if (dj < 0.0f && -dj < 0.5f) {
      const float djp = (float)(j + 1) - ja;
      val = fvna::fma_(R(-dj), at(i, j + 1), R(djp) * at(i, j));
    } else {
      continue;
    }
    // the closed form at (i + MDX, j + MDY) (:90-95)
    const R arg = R((float)(i + P.mdx)) * dx * R(D.tay_cos) +
                  (R(j + P.mdy - ny / 2) * dy + R(D.md_t) * R(D.tay_sin)) * R(D.tay_sin) - T;
    m[(size_t)tid + 2 * (size_t)nx] = val;
    m[(size_t)tid + 3 * (size_t)nx] = R(D.tay_rhoa) * tanh(R(D.tay_c) * arg);
    return;
  }
}

}  // namespace fvbc

#endif  // FIVONAGI_APPS_TAYLOR_ANGLE_XY_DATA_COLLECT_H
