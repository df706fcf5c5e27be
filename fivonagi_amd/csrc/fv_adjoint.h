// fv_adjoint.h -- per-instance table of the adjoint's reverse march kernel
// (device memory), shared by fv_adjoint_march.hip and the host runtime.
#ifndef FIVONAGI_FV_ADJOINT_H
#define FIVONAGI_FV_ADJOINT_H

#include <stdint.h>

#include "fv_types.h"

namespace fv {

struct AdjDesc {
  int32_t nx, ny, bc;
  float hbeta;                         // BETA - 1, formed as the step kernel forms it
  int32_t bx0, by0, bnx, bny;          // objective box
  const void* edge_c;                  // the step's drive tables (float, ny each)
  const void* edge_s;
  float* seg;                          // states of the current segment: slot s = 3 grids
  float* lam[2];                       // adjoint of the state (3 grids), ping-pong by step
  const Ctl* ctl_log;                  // control block accepted step n of the run ran with
  double* grad_ec;                     // dJ/d(amp_j cos phase_j), ny, accumulated
  double* grad_es;                     // dJ/d(amp_j sin phase_j)
  double* jpart;                       // per work item: sum over steps of dt sum_box rho'^2
  uint32_t* mismatch;                  // cells whose recomputed state differs (checking)
};

// One wave's work in the reverse march: the strip whose lane kHalo holds
// column x0, output columns [xo0, xo1), output rows [y0, y1).  Strip and row
// boundaries never separate column 2 from columns 0, 1, column nx-3 from
// nx-2, nx-1, row 2 from rows 0, 1 or row ny-3 from ny-2, ny-1 (the ghost
// copies' transposes stay within one wave).
struct AdjItem {
  int32_t inst, x0, xo0, xo1, y0, y1;
};

}  // namespace fv

#endif  // FIVONAGI_FV_ADJOINT_H
