/* apps/taylor-angle-xy/parameters.h -- the Taylor-shock app's parameters
 * (apps/taylor-angle-xy/parameters.h of the reference) in their runtime form,
 * for a C host: NX = NY = 64 ETAC, XMAX = YMAX = 12.5, BETA = 1,
 * DIFFDELTA = 1e-7, the shock rho' = rhoa tanh(C (x cos(theta) + y sin(theta)
 * - T)) with rhoa = -DIFFDELTA/BETA and C = 2.1972245773362196 (:105), the
 * window MDVX = cos(theta), MDVY = sin(theta), MDT = XMAX/2/cos(theta)
 * (:92-95), and a start at T = -TIDE with
 * TIDE = ceil(4 + (XMAX/2 - MDT sin(theta)) sin(theta)).  And the data
 * collection's options (data_collect.h:49-55): ETA = NX/XMAX, the cut's
 * half-length int(5 ETA cos(theta)) and tan(theta), in float as the FP32
 * build holds them when fp64 == 0.  Host code (C99). */
#ifndef FIVONAGI_APPS_TAYLOR_ANGLE_XY_PARAMETERS_H
#define FIVONAGI_APPS_TAYLOR_ANGLE_XY_PARAMETERS_H

#include <math.h>
#include <string.h>

#include "../../../include/fivonagi.h"

/* cflw: CFLWHISH; moving: the window moves (its MDVX, MDVY, MDT locate the
   measured cut either way). */
static inline void fivonagi_taylor_parameters(fivonagi_params *p, int etac, double theta,
                                              double cflw, int moving) {
  memset(p, 0, sizeof(*p));
  const int n = 64 * etac;
  const double xmax = 12.5, beta = 1.0, delta = 1e-7;
  p->nx = p->ny = n;
  p->dx = p->dy = xmax / (n - 1);
  p->beta = beta;
  p->delta = delta;
  p->cfl_max = 1.0;
  p->cfl_wish = cflw;
  p->dt0 = 0.005;
  const double mdt = xmax * 0.5 / cos(theta);
  p->T0 = -ceil(4.0 + (xmax * 0.5 - mdt * sin(theta)) * sin(theta));
  p->bc = FIVONAGI_BC_TAYLOR;
  p->cfl_mode = FIVONAGI_CFL_REFERENCE_FOLD;
  p->tay_rhoa = -delta / beta;
  p->tay_c = 2.1972245773362196;
  p->tay_theta = theta;
  p->tay_yoff = mdt * sin(theta);
  p->tay_jc = n / 2;
  p->md_on = moving ? 1 : 0;
  p->md_vx = cos(theta);
  p->md_vy = sin(theta);
  p->md_t = mdt;
}

static inline void fivonagi_taylor_collect_options(const fivonagi_params *p, int fp64,
                                                   fivonagi_collect_opts *o) {
  memset(o, 0, sizeof(*o));
  double eta, cs;
  if (fp64) {
    eta = p->nx / 12.5;
    cs = cos(p->tay_theta);
  } else {
    eta = (double)((float)p->nx / 12.5f);
    cs = (double)cosf((float)p->tay_theta);
  }
  o->eta = eta;
  o->cut_half = (int)(5.0 * eta * cs);
  o->tan_theta = tan(p->tay_theta);
}

#endif /* FIVONAGI_APPS_TAYLOR_ANGLE_XY_PARAMETERS_H */
