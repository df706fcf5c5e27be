/* apps/common/parameters.h -- the default app's parameters: a Gaussian
 * density pulse in a still medium with zero-gradient edges (the boundary of
 * common/boundary.h:11-82 of the reference), lengths in wavelengths of 20
 * cells (dx = dy = 1/20), c0 = 1, rho0 = 1.  The 400 x 400 pulse of
 * amplitude 0.01 and sigma = 10 cells, beta = 3.5, delta = 1e-4, is the
 * CPU-runnable workload.  Host code (C99). */
#ifndef FIVONAGI_APPS_COMMON_PARAMETERS_H
#define FIVONAGI_APPS_COMMON_PARAMETERS_H

#include <string.h>

#include "../../../include/fivonagi.h"

static inline void fivonagi_pulse_parameters(fivonagi_params *p, int nx, int ny) {
  memset(p, 0, sizeof(*p));
  p->nx = nx;
  p->ny = ny;
  p->dx = p->dy = 1.0 / 20.0;
  p->beta = 3.5;
  p->delta = 1e-4;
  p->cfl_max = 1.0;
  p->cfl_wish = 0.9;
  p->dt0 = 0.005;
  p->T0 = 0.0;
  p->bc = FIVONAGI_BC_ZEROGRAD;
  p->cfl_mode = FIVONAGI_CFL_REFERENCE_FOLD;
}

#endif /* FIVONAGI_APPS_COMMON_PARAMETERS_H */
