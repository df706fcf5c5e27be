/* apps/common/init.h -- the default app's initial state: rho' = A
 * exp(-r^2/(2 sigma^2)), r in cells from ((nx-1)/2, (ny-1)/2), at rest.  Host
 * code (C99); SoA grids of nx*ny, offset i + j*nx; values rounded to float
 * from the double expression. */
#ifndef FIVONAGI_APPS_COMMON_INIT_H
#define FIVONAGI_APPS_COMMON_INIT_H

#include <math.h>
#include <string.h>

#include "../../../include/fivonagi.h"

static inline void fivonagi_pulse_init(const fivonagi_params *p, double amp, double sigma_cells,
                                       float *q1, float *q2, float *q3) {
  const int nx = p->nx, ny = p->ny;
  const double xc = (nx - 1) / 2.0, yc = (ny - 1) / 2.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double x = i - xc, y = j - yc;
      q1[(size_t)i + (size_t)j * nx] =
          (float)(amp * exp(-(x * x + y * y) / (2.0 * sigma_cells * sigma_cells)));
    }
  memset(q2, 0, sizeof(float) * (size_t)nx * ny);
  memset(q3, 0, sizeof(float) * (size_t)nx * ny);
}

#endif /* FIVONAGI_APPS_COMMON_INIT_H */
