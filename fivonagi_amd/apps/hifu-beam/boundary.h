// apps/hifu-beam/boundary.h -- the hifu-beam app's ghost cells
// (apps/hifu-beam/boundary.h:20-133 of the reference), rules in its order:
//   right and top ghosts zero-gradient;
//   left ghost columns: on the aperture rows 2 <= j < 2 + int(ISPa/dy),
//     while the window has not moved in x (MDX == 0), the focused packet
//       pack = AMPL sin(ISPW (T + phi_j)) exp(-(FREC (T + phi_j)/4)^10),
//       rho' = pack, (rho u, rho v) = nu_j pack (pack + 1),
//     with phi_j = dx (j-2)^2/(2 ISPF) and nu_j = (ISPF, -(j-2))/|.|
//     (packet, :10-17, and :77-93); otherwise a copy of column 2;
//   rows 0, 1 mirror rows 4, 3 with rho v negated (the beam axis);
//   the 2x2 corners are 0.
// Per row the host tables hold AMPL cos/sin(ISPW phi_j), phi_j, nu_j and
// whether the row is an aperture row; sin/cos(ISPW T) come with the attempt.
#ifndef FIVONAGI_APPS_HIFU_BEAM_BOUNDARY_H
#define FIVONAGI_APPS_HIFU_BEAM_BOUNDARY_H

#include "../common/boundary.h"

namespace fvbc {

// AMPL sin(ISPW (T + phi)) = ampc sin(ISPW T) + amps cos(ISPW T) with the
// per-row ampc = AMPL cos(ISPW phi), amps = AMPL sin(ISPW phi);
// opT = FREC (T + phi)/4.
template <typename R>
__device__ __forceinline__ void hifu_packet(R ampc, R amps, R sin_t, R cos_t, R opT, R nux, R nuy,
                                            Cell<R>& q) {
  const R o2 = opT * opT, o4 = o2 * o2, o8 = o4 * o4;
  const R p = fvna::fma_(ampc, sin_t, amps * cos_t) * exp(-(o8 * o2));
  const R pp = p * (p + R(1));
  q.q1 = p;
  q.q2 = nux * pp;
  q.q3 = nuy * pp;
}

struct Hifu {
  static constexpr int kId = fv::kBcHifu;
  static constexpr bool kRowData = true;
  static constexpr bool kPrepass = false;
  static constexpr bool kExtOnly = true;    // the app's instances run the extended step
  __device__ static void lane(int ic, int nx, LaneMap& m, uint32_t size) {
    m.ic = ic;
    m.left = ic <= 1;
    // right ghosts copy column nx-3, left ones column 2; a ghost row's own
    // column otherwise (its corner cells are zeroed by patch)
    const int col = m.left ? 2 : (ic >= nx - 2 ? nx - 3 : ic);
    m.bA = m.bB = size * (uint32_t)col;
  }
  __device__ static uint32_t offset(int rc, int ny, const LaneMap& m, uint32_t row_bytes) {
    // bottom rows mirror rows 4, 3; top rows copy ny-3
    const int jm = rc <= 1 ? 4 - rc : (rc >= ny - 2 ? ny - 3 : rc);
    return (uint32_t)jm * row_bytes + m.bA;
  }
  template <typename R>
  __device__ static void load_row(const Tables& tb, int rc, RowData<R>& h) {
    typedef __attribute__((address_space(4))) const R kR;   // read-only: scalar loads
    const kR* tab = (const kR*)tb.tab;
    h.ec = ((const kR*)tb.edge_c)[rc];
    h.es = ((const kR*)tb.edge_s)[rc];
    h.ph = tab[4 * rc];
    h.nux = tab[4 * rc + 1];
    h.nuy = tab[4 * rc + 2];
    h.on = tab[4 * rc + 3];
  }
  template <typename R>
  __device__ static void patch(const PatchCtx<R>& c, int rc, Cell<R>& q, const RowData<R>& h) {
    q.q3 = rc <= 1 ? -q.q3 : q.q3;                  // reflective beam axis
    // the packet, on the aperture rows of the first strip's left lanes while
    // the window has not moved (boundary.h:82)
    if (c.mdx == 0 && c.x0 == 0 && h.on != R(0) && rc >= 2 && rc <= c.ny - 3) {
      Cell<R> p;
      // T + phi in FP64: the envelope's 10th power magnifies its rounding
      hifu_packet(h.ec, h.es, c.sin_t, c.cos_t, c.hk * R(c.T + (double)h.ph), h.nux, h.nuy, p);
      q.q1 = c.left ? p.q1 : q.q1;
      q.q2 = c.left ? p.q2 : q.q2;
      q.q3 = c.left ? p.q3 : q.q3;
    }
    const bool corner = (rc <= 1 || rc >= c.ny - 2) && (c.ic <= 1 || c.ic >= c.nx - 2);
    q.q1 = corner ? R(0) : q.q1;
    q.q2 = corner ? R(0) : q.q2;
    q.q3 = corner ? R(0) : q.q3;
  }
};

}  // namespace fvbc

#endif  // FIVONAGI_APPS_HIFU_BEAM_BOUNDARY_H
