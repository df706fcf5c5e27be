// apps/na/source.h -- the "na" (nonlinear acoustics) diffusion source as
// device code (apps/na/source.h:84-104 of the reference): on interior cells
//   rho u += delta dt/dx^2 (u_W - 2u + u_E) + delta dt/dy^2 (u_S - 2u + u_N),
// the same for rho v with v, u = rho u / rho, and rho' unchanged.
#ifndef FIVONAGI_APPS_NA_SOURCE_H
#define FIVONAGI_APPS_NA_SOURCE_H

#include "solver.h"

namespace fvna {

// The increment of one momentum component from its velocity u at the cell,
// u_we = u_W + u_E, u_s, u_n; ddx2 = delta dt/dx^2, ddy2 = delta dt/dy^2.
// The x and y parts are summed before they meet the state (apply_source):
// q + (dx + dy) rounds once at q's scale, where (q + dx) + dy rounds twice.
// In FP32 the two roundings land low on average (by 0.1 ulp of q per update,
// over a 3000-step edge-drive run where the two forms differ), which lowers
// the states at the shocks, hence the maximum wave speed, and the dt feedback
// turns that into a drift of T.
template <typename R>
__device__ __forceinline__ R source_increment(R ddx2, R ddy2, R u, R u_we, R u_s, R u_n) {
  return fma_(ddx2, fma_(R(-2), u, u_we), ddy2 * (fma_(R(-2), u, u_s) + u_n));
}

template <typename R>
__device__ __forceinline__ R apply_source(R q, R inc) {
  return add_nc(q, inc);
}

}  // namespace fvna

#endif  // FIVONAGI_APPS_NA_SOURCE_H
