// apps/hifu-beam/data_collect.h -- the hifu-beam app's data collection
// (apps/hifu-beam/data_collect.h:7-107 of the reference), after every
// accepted step, into the measurement grid laid out as the reference's
// measure1: the microphones and the running maximum of the central peak, on
// row JCUT of the new state before the window shift (staged by the step).
// Also the collection of every other app but the Taylor one.
#ifndef FIVONAGI_APPS_HIFU_BEAM_DATA_COLLECT_H
#define FIVONAGI_APPS_HIFU_BEAM_DATA_COLLECT_H

#include "../common/boundary.h"

namespace fvbc {

// Data collection of the hifu app (apps/hifu-beam/data_collect.h:51-106) for
// one accepted attempt: row jcut of its new state, staged by the step before
// the window shift, read with the pre-shift MDX.  Thread i = column i.
template <typename R>
__device__ void collect_hifu(const fv::InstDesc& D, const fv::Ctl& P, double T, int64_t nf,
                             R* __restrict__ m, int i) {
  const int nx = D.nx, ny = D.ny;
  if (i >= nx) return;
  const R u1 = ((const R*)D.stage)[i];
  const int mdx = P.mdx;
  for (int k = 0; k < 4; ++k) {                     // data_collect.h:51-76
    if (i + mdx == D.mic[k]) {
      R* row = m + (size_t)(2 + k) * nx;
      if (nf == 0)
        for (int ii = 0; ii < nx; ++ii) row[ii] = R(0);
      row[nf] = u1;
      break;
    }
  }
  if (i == 0 && nf == 0)                            // microphones outside the window
    for (int k = 0; k < 4; ++k)
      if (D.mic[k] - mdx < 0 || D.mic[k] - mdx >= nx)
        for (int ii = 0; ii < nx; ++ii) m[(size_t)(2 + k) * nx + ii] = R(0);
  // central peak only (data_collect.h:100-106)
  if (fabs((double)(i + mdx) / D.eta - T) > 0.5) return;
  const int64_t o = (int64_t)i + mdx + 6 * (int64_t)nx;
  if (o < (int64_t)nx * ny) m[o] = (u1 > m[o] ? u1 : m[o]);
}

}  // namespace fvbc

#endif  // FIVONAGI_APPS_HIFU_BEAM_DATA_COLLECT_H
