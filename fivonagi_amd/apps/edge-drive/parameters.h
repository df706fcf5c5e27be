/* apps/edge-drive/parameters.h -- the edge-drive workload's parameters, the
 * runtime form of an app's parameters.h (apps/hifu-beam/parameters.h:12-104
 * of the reference), for a C host: a quiescent medium driven from the left
 * ghost columns, rho' = amp_j sin(omega T + phase_j), through a Gaussian
 * aperture amp_j = A exp(-((j - (ny-1)/2)/w)^2); lengths in wavelengths of
 * 20 cells (dx = dy = 1/20), c0 = 1, rho0 = 1 (apps/hifu-beam/parameters.h:
 * 28-30, 96-97).  The 4096 x 1024 grid with A = 0.02, w = 200 cells and 20
 * cells per wavelength is the workload the step's throughput is quoted on.
 * Host code (C99). */
#ifndef FIVONAGI_APPS_EDGE_DRIVE_PARAMETERS_H
#define FIVONAGI_APPS_EDGE_DRIVE_PARAMETERS_H

#include <math.h>
#include <string.h>

#include "../../../include/fivonagi.h"

/* p: the instance's parameters; edge_amp: ny values the caller provides,
   filled with the aperture (p->edge_amp points at it, no phase). */
static inline void fivonagi_edge_drive_parameters(fivonagi_params *p, int nx, int ny, double amp,
                                                  double halfwidth_cells,
                                                  double cells_per_wavelength, double *edge_amp) {
  const double pi = 3.14159265358979323846;
  memset(p, 0, sizeof(*p));
  p->nx = nx;
  p->ny = ny;
  p->dx = p->dy = 1.0 / 20.0;
  p->beta = 3.5;
  p->delta = 5e-5;
  p->cfl_max = 1.0;
  p->cfl_wish = 0.9;
  p->dt0 = 0.005;
  p->T0 = 0.0;
  p->bc = FIVONAGI_BC_EDGE_DRIVE;
  p->cfl_mode = FIVONAGI_CFL_REFERENCE_FOLD;
  p->omega = 2.0 * pi * 20.0 / cells_per_wavelength;
  const double jc = (ny - 1) / 2.0;
  for (int j = 0; j < ny; ++j) {
    const double s = (j - jc) / halfwidth_cells;
    edge_amp[j] = amp * exp(-s * s);
  }
  p->edge_amp = edge_amp;
  p->edge_phase = NULL;
}

#endif /* FIVONAGI_APPS_EDGE_DRIVE_PARAMETERS_H */
