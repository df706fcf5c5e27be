// fv_adjoint.h -- per-instance table of the adjoint march kernels (device
// memory), shared by fv_adjoint_march.hip and the host runtime.
#ifndef FIVONAGI_FV_ADJOINT_H
#define FIVONAGI_FV_ADJOINT_H

#include <stdint.h>

namespace fv {

struct AdjDesc {
  int32_t nx, ny, bc;
  float beta;
  double dx, dy, delta, omega;
  double tay_rhoa, tay_c, tay_cos, tay_sin, tay_yoff;
  int32_t tay_jc;
  int32_t bx0, by0, bnx, bny;          // objective box
  int32_t pad;
  const float* edge_amp;               // ny drive amplitudes
  const float* edge_phase;             // ny drive phases
  float* seg;                          // (K+1) states of the current segment, 3 grids each
  float* lam;                          // adjoint of the state (3 grids)
  float* lamx;                         // adjoint of q_y between the two reverse kernels
  double* grad_amp;                    // ny, accumulated
  double* grad_phase;                  // ny, accumulated
  double* jstep;                       // per step: dt_n sum_box rho'_{n+1}^2
  const double* dt_log;                // dt of accepted step n of the run
  const double* T_log;                 // T at the start of accepted step n
};

// One wave's work in the adjoint march kernels: a strip of one instance,
// output columns [x0, x0 + width), rows [y0, y1).
struct AdjItem {
  int32_t inst, x0, y0, y1;
};

}  // namespace fv

#endif  // FIVONAGI_FV_ADJOINT_H
