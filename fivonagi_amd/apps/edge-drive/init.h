/* apps/edge-drive/init.h -- the initial state of the edge-drive workload (the
 * init() of an app, apps/hifu-beam/init.h:10-23 of the reference): a
 * quiescent medium, rho' = rho u = rho v = 0.  Host code (C99); SoA grids of
 * nx*ny, offset i + j*nx. */
#ifndef FIVONAGI_APPS_EDGE_DRIVE_INIT_H
#define FIVONAGI_APPS_EDGE_DRIVE_INIT_H

#include <string.h>

#include "../../../include/fivonagi.h"

static inline void fivonagi_edge_drive_init(const fivonagi_params *p, float *q1, float *q2,
                                            float *q3) {
  const size_t n = (size_t)p->nx * (size_t)p->ny;
  memset(q1, 0, sizeof(float) * n);
  memset(q2, 0, sizeof(float) * n);
  memset(q3, 0, sizeof(float) * n);
}

#endif /* FIVONAGI_APPS_EDGE_DRIVE_INIT_H */
