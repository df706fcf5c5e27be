// apps/na/adjoint.h -- reverse-mode pieces of the "na" plug-in (device code),
// used by the adjoint march kernel (csrc/fv_adjoint_march.hip).
//
// The forward values the reverse pass branches on (waves, speeds, the
// limiter's ratios, P and M) come from the functions of apps/na/solver.h that
// the step kernel calls, with the same operands, so the upwind side, the sign
// of each speed and the limiter case of every interface are those of the
// executed forward step.  What the transposes need beyond those values is
// formed next to them here.
//
// Forward (solver.h riemann; L = cell k-1, R = cell k, components
// (rho', normal momentum n, tangential momentum t), un = n/rho, ut = t/rho):
//   s1 = r1L + r1R, s2 = unL + unR, s3 = utL + utR, S1 = s1 + 2
//   drho = r1R - r1L, dn = nR - nL, dtg = tR - tL
//   tmp1 = s2/S1, tmp3 = tmp1 (tmp1 - s2) + hbeta s1 + 1, sq = sqrt(tmp3)
//   L1,2 = tmp1 +- sq, L3 = tmp1, d_p = L_p S1 - s3, A_p = s3 (L_p - s2)/d_p
//   a1 = (dn - drho L2)/(2 sq), a2 = drho - a1, a3 = dtg - (a1 A1 + a2 A2)
//   W1 = a1 (1, L1, A1), W2 = a2 (1, L2, A2), W3 = (0, 0, a3)
// Limiter (solver.h limited_flux), per family p:
//   r_p = W_p . W_p(upwind) / |W_p|^2, h_p = clamp(min((1 + r_p)/4, r_p), 0, 1)
//   c_p = |L_p| (1 - |L_p| dt/d) h_p, al_p = max(L_p, 0) - c_p, be_p = L_p - al_p
//   P = sum_p al_p W_p, M = sum_p be_p W_p
// The derivatives are those of that branch path with dt frozen; a limiter
// ratio at a clamp (h = 0 or 1) or of a zero wave has derivative 0.
#ifndef FIVONAGI_APPS_NA_ADJOINT_H
#define FIVONAGI_APPS_NA_ADJOINT_H

#include <hip/hip_runtime.h>

#include "solver.h"

namespace fvad {

using fvna::Iface;
using fvna::Waves;

// What the limiter's transpose needs of one interface, per wave family:
// h = phi/2, qd = (dh/dr) / |W|^2 and r (both 0 where h is clamped or the
// wave is zero, so no infinity or NaN of the forward enters a product).
struct LimT {
  float h[3], qd[3], r[3];
};

__device__ __forceinline__ void lim_part(float n, float dot, float& h, float& qd, float& r) {
  const float rn = fvna::rcp(n);
  const float rr = dot * rn;
  const float t = fvna::fma_(0.25f, rr, 0.25f);
  const float mn = fvna::min_(t, rr);
  h = fvna::clamp01(mn);
  const bool live = mn > 0.0f && mn < 1.0f;
  qd = live ? (rr < t ? rn : 0.25f * rn) : 0.0f;
  r = live ? rr : 0.0f;
}

// The limited fluxes of interface f as the step kernel forms them (same
// call), and the transpose's quantities.  dl_p = W_p . W_p(k-1),
// dr_p = W_p . W_p(k+1).
__device__ __forceinline__ void limiter_rec(const Iface& f, float dl1, float dr1, float dl2,
                                            float dr2, float dl3, float dr3, float dtd,
                                            float P[3], float M[3], LimT& t) {
  fvna::limited_flux(f, dl1, dr1, dl2, dr2, dl3, dr3, dtd, P, M);
  const Waves& w = f.w;
  const float n1 = fvna::fma_(w.a1, w.a1, fvna::fma_(w.w21, w.w21, w.w31 * w.w31));
  const float n2 = fvna::fma_(w.a2, w.a2, fvna::fma_(w.w22, w.w22, w.w32 * w.w32));
  const float n3 = w.a3 * w.a3;
  lim_part(n1, f.L1 > 0.0f ? dl1 : dr1, t.h[0], t.qd[0], t.r[0]);
  lim_part(n2, f.L2 > 0.0f ? dl2 : dr2, t.h[1], t.qd[1], t.r[1]);
  lim_part(n3, f.L3 > 0.0f ? dl3 : dr3, t.h[2], t.qd[2], t.r[2]);
}

// Adjoint of one interface's waves (a1, w21, w31, a2, w22, w32, a3) and
// speeds (L1, L2, L3).
struct WaveAdj {
  float g[7];
  float gs[3];
};

// An empty accumulator.  -0 is the exact identity of IEEE addition
// (-0 + x = x for every x, +0 included; fma(a, b, -0) = a b), so the first
// term added to it needs no instruction of its own; +0 is not (+0 + -0 = +0).
__device__ __forceinline__ void zero(WaveAdj& a) {
#pragma unroll
  for (int c = 0; c < 7; ++c) a.g[c] = -0.0f;
  a.gs[0] = a.gs[1] = a.gs[2] = -0.0f;
}

// Transpose of limiter_rec: from the adjoints gP, gM of the interface's P
// and M (component order of the sweep), accumulate the adjoint of its own
// waves and speeds into ga and return the adjoint of each family's dot
// product with its upwind neighbour (gd).
__device__ __forceinline__ void limiter_bw(const Iface& f, const LimT& t, float dtd,
                                           const float gP[3], const float gM[3], WaveAdj& ga,
                                           float gd[3]) {
  const Waves& w = f.w;
  const float L[3] = {f.L1, f.L2, f.L3};
  float gal[3], gbe[3];
  gal[0] = fvna::fma_(gP[0], w.a1, fvna::fma_(gP[1], w.w21, gP[2] * w.w31));
  gbe[0] = fvna::fma_(gM[0], w.a1, fvna::fma_(gM[1], w.w21, gM[2] * w.w31));
  gal[1] = fvna::fma_(gP[0], w.a2, fvna::fma_(gP[1], w.w22, gP[2] * w.w32));
  gbe[1] = fvna::fma_(gM[0], w.a2, fvna::fma_(gM[1], w.w22, gM[2] * w.w32));
  gal[2] = gP[2] * w.a3;
  gbe[2] = gM[2] * w.a3;
  float al[3], be[3], gn2[3];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const float s = fabsf(L[p]);
    const float om = fvna::fma_(-s, dtd, 1.0f);        // 1 - |L| dt/d
    const float c = s * om * t.h[p];
    al[p] = fvna::fmax2(L[p], 0.0f) - c;
    be[p] = L[p] - al[p];
    // be = L - al; al = max(L, 0) - c; c = s (1 - s dt/d) h
    const float g2 = gal[p] - gbe[p];                  // adjoint of al, be's share folded
    const float gc = -g2;
    const float gh = gc * s * om;
    const float gsa = gc * t.h[p] * fvna::fma_(-2.0f * s, dtd, 1.0f);
    float gl = gbe[p] + (L[p] > 0.0f ? g2 : 0.0f);
    gl += L[p] > 0.0f ? gsa : (L[p] < 0.0f ? -gsa : 0.0f);
    ga.gs[p] += gl;
    gd[p] = gh * t.qd[p];                              // h(r), r = dot / |W|^2
    gn2[p] = -2.0f * gd[p] * t.r[p];                   // 2 x adjoint of |W|^2
  }
  ga.g[0] += fvna::fma_(al[0], gP[0], fvna::fma_(be[0], gM[0], gn2[0] * w.a1));
  ga.g[1] += fvna::fma_(al[0], gP[1], fvna::fma_(be[0], gM[1], gn2[0] * w.w21));
  ga.g[2] += fvna::fma_(al[0], gP[2], fvna::fma_(be[0], gM[2], gn2[0] * w.w31));
  ga.g[3] += fvna::fma_(al[1], gP[0], fvna::fma_(be[1], gM[0], gn2[1] * w.a2));
  ga.g[4] += fvna::fma_(al[1], gP[1], fvna::fma_(be[1], gM[1], gn2[1] * w.w22));
  ga.g[5] += fvna::fma_(al[1], gP[2], fvna::fma_(be[1], gM[2], gn2[1] * w.w32));
  ga.g[6] += fvna::fma_(al[2], gP[2], fvna::fma_(be[2], gM[2], gn2[2] * w.a3));
}

// ga += the dot-product adjoints of the families of a neighbouring
// interface's limiter (or of this one's) times the other interface's waves:
// kp[p] multiplies family p of W.
__device__ __forceinline__ void add_scaled(WaveAdj& ga, const float kp[3], const Waves& W) {
  ga.g[0] = fvna::fma_(kp[0], W.a1, ga.g[0]);
  ga.g[1] = fvna::fma_(kp[0], W.w21, ga.g[1]);
  ga.g[2] = fvna::fma_(kp[0], W.w31, ga.g[2]);
  ga.g[3] = fvna::fma_(kp[1], W.a2, ga.g[3]);
  ga.g[4] = fvna::fma_(kp[1], W.w22, ga.g[4]);
  ga.g[5] = fvna::fma_(kp[1], W.w32, ga.g[5]);
  ga.g[6] = fvna::fma_(kp[2], W.a3, ga.g[6]);
}

// Adjoint of the sums and differences one Riemann solve starts from:
// cell R receives (s1 + dr, dn, dt, s2, s3) and cell L (s1 - dr, -dn, -dt,
// s2, s3) as the adjoint of its (rho', n, t, un, ut).
struct SumAdj {
  float s1, s2, s3, dr, dn, dt;
};

// Transpose of solver.h riemann at one interface, from the adjoint of its
// waves and speeds (ga).  s1..dn as the forward forms them; f its result.
__device__ __forceinline__ SumAdj riemann_bw(float s1, float s2, float s3, float drho, float dn,
                                             float hbeta, const Iface& f, const WaveAdj& ga) {
  const float L1 = f.L1, L2 = f.L2, tmp1 = f.L3;
  const float a1 = f.w.a1, a2 = f.w.a2;
  const float S1 = s1 + 2.0f;
  const float rS = fvna::rcp(S1);
  const float tmp3 = fvna::fma_(tmp1, tmp1 - s2, fvna::fma_(hbeta, s1, 1.0f));
  const float hrs = 0.5f * fvna::rsq(tmp3);             // 1 / (2 sq) = 1 / (L1 - L2)
  const float d1 = fvna::fma_(L1, S1, -s3), d2 = fvna::fma_(L2, S1, -s3);
  const float rdd = fvna::rcp(d1 * d2);
  const float id1 = d2 * rdd, id2 = d1 * rdd;           // 1/d1, 1/d2
  const float A1 = s3 * (L1 - s2) * id1, A2 = s3 * (L2 - s2) * id2;
  SumAdj o;
  // a3 = dtg - (a1 A1 + a2 A2), w21 = a1 L1, w31 = a1 A1, w22 = a2 L2, w32 = a2 A2
  const float ga3 = ga.g[6];
  const float gw31 = ga.g[2] - ga3, gw32 = ga.g[5] - ga3;
  o.dt = ga3;
  float ga1 = fvna::fma_(ga.g[1], L1, fvna::fma_(gw31, A1, ga.g[0]));
  const float ga2 = fvna::fma_(ga.g[4], L2, fvna::fma_(gw32, A2, ga.g[3]));
  float gL1 = fvna::fma_(ga.g[1], a1, ga.gs[0]);
  float gL2 = fvna::fma_(ga.g[4], a2, ga.gs[1]);
  const float gA1 = gw31 * a1, gA2 = gw32 * a2;
  // a2 = drho - a1
  float gdr = ga2;
  ga1 -= ga2;
  // a1 = (dn - drho L2) hrs, hrs = 1 / (2 sq)
  const float t1 = ga1 * hrs;
  o.dn = t1;
  gdr = fvna::fma_(-t1, L2, gdr);
  gL2 = fvna::fma_(-t1, drho, gL2);
  float gsq = -2.0f * t1 * a1;
  // A_p = s3 (L_p - s2) / d_p, d_p = L_p S1 - s3
  const float e1 = gA1 * id1, e2 = gA2 * id2;
  gL1 = fvna::fma_(e1, fvna::fma_(-A1, S1, s3), gL1);
  gL2 = fvna::fma_(e2, fvna::fma_(-A2, S1, s3), gL2);
  float gs2 = -(e1 + e2) * s3;
  o.s3 = fvna::fma_(e1, (L1 - s2) + A1, e2 * ((L2 - s2) + A2));
  float gS1 = -fvna::fma_(e1 * A1, L1, e2 * A2 * L2);
  // L1,2 = tmp1 +- sq, L3 = tmp1; sq = sqrt(tmp3)
  float gt1 = gL1 + gL2 + ga.gs[2];
  gsq += gL1 - gL2;
  const float gt3 = gsq * hrs;
  // tmp3 = tmp1 (tmp1 - s2) + hbeta s1 + 1
  gt1 = fvna::fma_(gt3, 2.0f * tmp1 - s2, gt1);
  gs2 = fvna::fma_(-gt3, tmp1, gs2);
  float gs1 = gt3 * hbeta;
  // tmp1 = s2 / S1, S1 = s1 + 2
  gs2 = fvna::fma_(gt1, rS, gs2);
  gS1 = fvna::fma_(-gt1 * tmp1, rS, gS1);
  gs1 += gS1;
  o.s1 = gs1;
  o.s2 = gs2;
  o.dr = gdr;
  return o;
}

}  // namespace fvad

#endif  // FIVONAGI_APPS_NA_ADJOINT_H
