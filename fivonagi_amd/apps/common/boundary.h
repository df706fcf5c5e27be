// apps/common/boundary.h -- the boundary plug-in interface of the fused step
// and the default app's rule, zero-gradient ghost cells
// (common/boundary.h:11-82 of the reference).
//
// An app's boundary.h defines one policy: a struct with the members below,
// which the step kernel calls while it loads a row of the input state (so the
// ghost ring costs no pass of its own).  Rules read the input state only, as
// the reference reads the other texture, so corner ghosts lag one step.
//
//   kId            the instance's boundary kind (FIVONAGI_BC_*)
//   kRowData       the policy reads per-row tables (RowData) with each row
//   kPrepass       ghost cells written before the step by prepass() (their
//                  values depend on T and the cell only); the step copies them
//   kExtOnly       the app always runs with the extended step (its instances
//                  set it), so the plain step need not carry its march
//   lane(ic, nx, m)                  once per lane: the source columns
//   offset(rc, ny, m, row_bytes)     per row: byte offset of the source cell
//   load_row(tb, rc, h)              per row: the row's table entries
//   patch(c, rc, q, h)               per row: values that are not copies
//
// A new app adds a directory with its boundary.h (and parameters.h, init.h),
// and one line to apps/apps.h.
#ifndef FIVONAGI_APPS_COMMON_BOUNDARY_H
#define FIVONAGI_APPS_COMMON_BOUNDARY_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../csrc/fv_types.h"
#include "../na/solver.h"

namespace fvbc {

// Moving window (common/main.cu:109-135, :167-176): once T >= MDT the window
// sits int((T - MDT) * MDV / d) cells from its start; after an accepted step
// ending at T it moves by the part not applied yet (md = the current offset).
__device__ __forceinline__ int md_shift(double T, double md_t, double v, double d, int md) {
  return T >= md_t ? (int)((T - md_t) * v / d) - md : 0;
}

// A lane's source columns: bA on rows 2..ny-3, bB on the ghost rows (bytes).
struct LaneMap {
  uint32_t bA, bB;
  bool left;         // the lane's column is a left ghost column (i <= 1)
  int ic;            // the lane's column, clamped to the grid
};

// The instance's per-row tables (device pointers, the context's precision):
// edge_c, edge_s = ampc, amps per row; tab = 4 entries per row (hifu).
struct Tables {
  const void* edge_c;
  const void* edge_s;
  const void* tab;
};

// Per-row table entries: edge drive ampc, amps (amp_j cos/sin phase_j, or the
// aperture's AMPL cos/sin(ISPW phi_j)); aperture phi_j, direction, on/off.
template <typename R>
struct RowData {
  R ec, es, ph, nux, nuy, on;
};

// Values uniform over a wave of one attempt, for patch().
template <typename R>
struct PatchCtx {
  R sin_t, cos_t;    // sin/cos(omega T) of the attempt
  double T;          // T of the attempt
  R hk;              // hifu: FREC/4
  int mdx;           // moving-window offset MDX
  int x0;            // first output column of the wave's strip
  int nx, ny;
  int ic;            // the lane's clamped column
  bool left;         // the lane's column is a left ghost column
  bool any_left;     // some lane of the wave has a left ghost column (uniform)
};

template <typename R>
struct Cell {
  R q1, q2, q3;      // rho', rho u, rho v
};

// common/boundary.h:27-81: later rules win --
// right (i >= nx-2) -> (nx-3, j), bottom (j <= 1) -> (i, 2),
// top (j >= ny-2) -> (i, ny-3), left (i <= 1) -> (2, j).
struct ZeroGradient {
  static constexpr int kId = fv::kBcZeroGrad;
  static constexpr bool kRowData = false;
  static constexpr bool kPrepass = false;
  static constexpr bool kExtOnly = false;
  __device__ static void lane(int ic, int nx, LaneMap& m, uint32_t size) {
    m.ic = ic;
    m.left = ic <= 1;
    const int colA = m.left ? 2 : (ic >= nx - 2 ? nx - 3 : ic);
    const int colB = m.left ? 2 : ic;   // ghost rows: the left rule keeps its own row
    m.bA = size * (uint32_t)colA;
    m.bB = size * (uint32_t)colB;
  }
  __device__ static uint32_t offset(int rc, int ny, const LaneMap& m, uint32_t row_bytes) {
    const bool grow = (rc <= 1) || (rc >= ny - 2);
    const int jm = rc >= ny - 2 ? ny - 3 : (rc <= 1 ? 2 : rc);
    const uint32_t rowL = (uint32_t)rc * row_bytes, rowM = (uint32_t)jm * row_bytes;
    return grow ? (m.left ? rowL : rowM) + m.bB : rowM + m.bA;
  }
  template <typename R>
  __device__ static void load_row(const Tables&, int, RowData<R>&) {}
  template <typename R>
  __device__ static void patch(const PatchCtx<R>&, int, Cell<R>&, const RowData<R>&) {}
};

}  // namespace fvbc

#endif  // FIVONAGI_APPS_COMMON_BOUNDARY_H
