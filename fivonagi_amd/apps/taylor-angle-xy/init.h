/* apps/taylor-angle-xy/init.h -- the Taylor-shock app's initial state
 * (apps/taylor-angle-xy/init.h of the reference): the closed form at T0,
 *   rho' = rhoa tanh(C (i dx cos(theta) + ((j - jc) dy + yoff) sin(theta) - T0)),
 *   (rho u, rho v) = rho' (cos(theta), sin(theta)).
 * Host code (C99); SoA grids of nx*ny, offset i + j*nx; the double values
 * rounded to float. */
#ifndef FIVONAGI_APPS_TAYLOR_ANGLE_XY_INIT_H
#define FIVONAGI_APPS_TAYLOR_ANGLE_XY_INIT_H

#include <math.h>

#include "../../../include/fivonagi.h"

static inline void fivonagi_taylor_init(const fivonagi_params *p, float *q1, float *q2,
                                        float *q3) {
  const double cs = cos(p->tay_theta), sn = sin(p->tay_theta);
  for (int j = 0; j < p->ny; ++j)
    for (int i = 0; i < p->nx; ++i) {
      const double r = p->tay_rhoa * tanh(p->tay_c * ((double)i * p->dx * cs +
                                                      ((double)(j - p->tay_jc) * p->dy +
                                                       p->tay_yoff) * sn - p->T0));
      const size_t o = (size_t)i + (size_t)j * p->nx;
      q1[o] = (float)r;
      q2[o] = (float)(r * cs);
      q3[o] = (float)(r * sn);
    }
}

#endif /* FIVONAGI_APPS_TAYLOR_ANGLE_XY_INIT_H */
