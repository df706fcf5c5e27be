// apps/na/adjoint.h -- reverse-mode pieces of the "na" plug-in (device code),
// used by the adjoint march kernels (csrc/fv_adjoint_march.hip).
//
// Forward restated (generic direction; L = cell m-1, R = cell m, components
// (rho', normal momentum, tangential momentum)), as apps/na/fv.h:88-139 and
// :579-711 with the fluctuation and correction flux folded per wave family:
//   rho = 1 + r, u = n/rho, v = t/rho
//   Z1 = (rhoR+rhoL)/2, Z2 = (uR+uL)/2, Z3 = (vR+vL)/2, tmp1 = Z2/Z1
//   tmp3 = tmp1^2 - 2 Z2^2/Z1 + 1 + 2(beta-1)(Z1-1), sq = sqrt(tmp3)
//   L1,2 = tmp1 +- sq, L3 = tmp1, A_k = (L_k Z3 - 2 Z2 Z3)/(L_k Z1 - Z3)
//   a1 = (dn - dr L2)/(L1-L2), a2 = dr - a1, a3 = dt - a1 A1 - a2 A2
//   W1 = a1 (1, L1, A1), W2 = a2 (1, L2, A2), W3 = (0, 0, a3)
//   phi_p = MC(r_p), r_p = W_p . W_p(upwind) / |W_p|^2
//   c_p = |s_p| (1 - |s_p| dt/d) phi_p / 2
//   P = sum (max(s_p,0) - c_p) W_p,  M = sum (min(s_p,0) + c_p) W_p
// The derivatives are those of the executed branch path (limiter case,
// upwind side, sign of s) with dt frozen.  Divisions are reciprocal
// multiplies (v_rcp_f32, 1 ulp): the adjoint is checked against the FP64
// host adjoint and FP64 finite differences at 1e-3.
#ifndef FIVONAGI_APPS_NA_ADJOINT_H
#define FIVONAGI_APPS_NA_ADJOINT_H

#include <hip/hip_runtime.h>

namespace fvad {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

struct RW {          // one interface: waves, speeds and what the reverse pass needs
  float W[7];        // a1, a1 L1, a1 A1, a2, a2 L2, a2 A2, a3
  float s[3];        // L1, L2, L3
  float irhoL, irhoR, uL, uR, vL, vR, iZ1, Z2, Z3, tmp1, isq, A1, A2, iE1, iE2, iD, dr;
};

__device__ __forceinline__ void riemann_fw(float rL, float nL, float tL, float rR, float nR,
                                           float tR, float beta, RW& w) {
  w.irhoL = rcp(1.0f + rL);
  w.irhoR = rcp(1.0f + rR);
  w.uL = nL * w.irhoL; w.uR = nR * w.irhoR;
  w.vL = tL * w.irhoL; w.vR = tR * w.irhoR;
  const float z1m1 = 0.5f * (rR + rL);
  const float Z1 = 1.0f + z1m1;
  w.Z2 = 0.5f * (w.uR + w.uL);
  w.Z3 = 0.5f * (w.vR + w.vL);
  w.iZ1 = rcp(Z1);
  w.tmp1 = w.Z2 * w.iZ1;
  const float tmp3 = w.tmp1 * w.tmp1 - 2.0f * w.Z2 * w.Z2 * w.iZ1 + 1.0f
                     + 2.0f * (beta - 1.0f) * z1m1;
  const float sq = __builtin_sqrtf(fmaxf(tmp3, 0.0f));
  w.isq = rcp(sq);
  w.s[0] = w.tmp1 + sq;
  w.s[1] = w.tmp1 - sq;
  w.s[2] = w.tmp1;
  w.iE1 = rcp(w.s[0] * Z1 - w.Z3);
  w.iE2 = rcp(w.s[1] * Z1 - w.Z3);
  w.A1 = (w.s[0] * w.Z3 - 2.0f * w.Z2 * w.Z3) * w.iE1;
  w.A2 = (w.s[1] * w.Z3 - 2.0f * w.Z2 * w.Z3) * w.iE2;
  w.dr = rR - rL;
  const float dn = nR - nL, dt = tR - tL;
  w.iD = rcp(w.s[0] - w.s[1]);
  const float a1 = (dn - w.dr * w.s[1]) * w.iD;
  const float a2 = w.dr - a1;
  w.W[0] = a1; w.W[1] = a1 * w.s[0]; w.W[2] = a1 * w.A1;
  w.W[3] = a2; w.W[4] = a2 * w.s[1]; w.W[5] = a2 * w.A2;
  w.W[6] = dt - a1 * w.A1 - a2 * w.A2;
}

// reverse of riemann_fw: gW (7), gs (3) -> accumulate gL, gR (3 each)
__device__ __forceinline__ void riemann_bw(const RW& w, float beta, const float gW[7],
                                           const float gs[3], float gL[3], float gR[3]) {
  const float a1 = w.W[0], a2 = w.W[3];
  const float L1 = w.s[0], L2 = w.s[1];
  const float ga3 = gW[6];
  float ga1 = gW[0] + gW[1] * L1 + gW[2] * w.A1;
  float ga2 = gW[3] + gW[4] * L2 + gW[5] * w.A2;
  float gL1 = gW[1] * a1 + gs[0];
  float gL2 = gW[4] * a2 + gs[1];
  float gA1 = gW[2] * a1, gA2 = gW[5] * a2;
  float gtmp1 = gs[2];
  const float gdt = ga3;
  ga1 -= ga3 * w.A1; gA1 -= ga3 * a1;
  ga2 -= ga3 * w.A2; gA2 -= ga3 * a2;
  float gdr = ga2;
  ga1 -= ga2;
  const float gdn = ga1 * w.iD;
  gdr -= ga1 * L2 * w.iD;
  gL2 -= ga1 * w.dr * w.iD;
  const float gD = -ga1 * a1 * w.iD;
  gL1 += gD; gL2 -= gD;
  const float gN1 = gA1 * w.iE1, gE1 = -gA1 * w.A1 * w.iE1;
  const float gN2 = gA2 * w.iE2, gE2 = -gA2 * w.A2 * w.iE2;
  const float Z1 = rcp(w.iZ1);
  gL1 += gN1 * w.Z3 + gE1 * Z1;
  gL2 += gN2 * w.Z3 + gE2 * Z1;
  const float gZ3 = gN1 * (L1 - 2.0f * w.Z2) + gN2 * (L2 - 2.0f * w.Z2) - gE1 - gE2;
  float gZ2 = -2.0f * w.Z3 * (gN1 + gN2);
  float gZ1 = gE1 * L1 + gE2 * L2;
  gtmp1 += gL1 + gL2;
  const float gsq = gL1 - gL2;
  const float gtmp3 = 0.5f * gsq * w.isq;
  gtmp1 += 2.0f * w.tmp1 * gtmp3;
  gZ2 += -4.0f * w.Z2 * w.iZ1 * gtmp3;
  gZ1 += (2.0f * w.Z2 * w.Z2 * w.iZ1 * w.iZ1 + 2.0f * (beta - 1.0f)) * gtmp3;
  gZ2 += gtmp1 * w.iZ1;
  gZ1 += -gtmp1 * w.tmp1 * w.iZ1;
  float grhoR = 0.5f * gZ1, grhoL = 0.5f * gZ1;
  const float gu = 0.5f * gZ2, gv = 0.5f * gZ3;
  grhoR += -(gu * w.uR + gv * w.vR) * w.irhoR;
  grhoL += -(gu * w.uL + gv * w.vL) * w.irhoL;
  gR[0] += gdr + grhoR;
  gL[0] += -gdr + grhoL;
  gR[1] += gdn + gu * w.irhoR;
  gL[1] += -gdn + gu * w.irhoL;
  gR[2] += gdt + gv * w.irhoR;
  gL[2] += -gdt + gv * w.irhoL;
}

__device__ __forceinline__ void wave_vec(const float W[7], int p, float v[3]) {
  if (p == 0) { v[0] = W[0]; v[1] = W[1]; v[2] = W[2]; }
  else if (p == 1) { v[0] = W[3]; v[1] = W[4]; v[2] = W[5]; }
  else { v[0] = 0.0f; v[1] = 0.0f; v[2] = W[6]; }
}

struct Lim {
  float inrm[3];     // 1/|W_p|^2 (inf for a zero wave)
  float r[3], phi[3], dphidr[3], al[3], be[3];
  int nb;            // bit p set: family p dots with interface m+1, clear: with m-1
};

// limiter and folded flux of interface m from the waves of m-1, m, m+1
__device__ __forceinline__ void limiter_fw(const float Wm[7], const float W[7], const float s[3],
                                           const float Wp[7], float dtd, Lim& l, float P[3],
                                           float M[3]) {
  P[0] = P[1] = P[2] = M[0] = M[1] = M[2] = 0.0f;
  l.nb = 0;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    float v[3], vm[3], vp[3];
    wave_vec(W, p, v);
    wave_vec(Wm, p, vm);
    wave_vec(Wp, p, vp);
    const bool right = !(s[p] > 0.0f);
    if (right) l.nb |= 1 << p;
    const float vn0 = right ? vp[0] : vm[0], vn1 = right ? vp[1] : vm[1],
                vn2 = right ? vp[2] : vm[2];
    const float nrm = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float dot = v[0] * vn0 + v[1] * vn1 + v[2] * vn2;
    const bool live = nrm != 0.0f;
    l.inrm[p] = rcp(nrm);
    const float r = live ? dot * l.inrm[p] : 0.0f;
    const float c = 0.5f * (1.0f + r);
    float wv = (c < 2.0f) ? c : 2.0f;
    float dphidr = (c < 2.0f) ? 0.5f : 0.0f;
    const float r2 = 2.0f * r;
    const bool lo = !(wv < r2);
    wv = lo ? r2 : wv;
    dphidr = lo ? 2.0f : dphidr;
    const bool neg = !(wv > 0.0f);
    wv = neg ? 0.0f : wv;
    dphidr = neg ? 0.0f : dphidr;
    l.r[p] = r;
    l.phi[p] = live ? wv : 1.0f;
    l.dphidr[p] = live ? dphidr : 0.0f;
    const float as = fabsf(s[p]);
    const float cc = 0.5f * as * (1.0f - as * dtd) * l.phi[p];
    l.al[p] = fmaxf(s[p], 0.0f) - cc;
    l.be[p] = fminf(s[p], 0.0f) + cc;
#pragma unroll
    for (int k = 0; k < 3; ++k) { P[k] += l.al[p] * v[k]; M[k] += l.be[p] * v[k]; }
  }
}

// reverse of limiter_fw: gP, gM -> gW (own, 7), gs (3), gdot (3)
__device__ __forceinline__ void limiter_bw(const float W[7], const float s[3], const Lim& l,
                                           float dtd, const float gP[3], const float gM[3],
                                           float gW[7], float gs[3], float gdot[3]) {
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    float v[3];
    wave_vec(W, p, v);
    const float as = fabsf(s[p]);
    float gv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) gv[k] = l.al[p] * gP[k] + l.be[p] * gM[k];
    const float gal = gP[0] * v[0] + gP[1] * v[1] + gP[2] * v[2];
    const float gbe = gM[0] * v[0] + gM[1] * v[1] + gM[2] * v[2];
    float gss = (s[p] > 0.0f ? gal : 0.0f) + (s[p] < 0.0f ? gbe : 0.0f);
    const float gc = -gal + gbe;
    const float gphi = gc * 0.5f * as * (1.0f - as * dtd);
    const float gas = gc * 0.5f * l.phi[p] * (1.0f - 2.0f * as * dtd);
    gss += gas * (s[p] > 0.0f ? 1.0f : (s[p] < 0.0f ? -1.0f : 0.0f));
    gs[p] += gss;
    // dphidr is 0 for a zero wave, so gr = 0 and inrm = inf never meet
    const float gr = gphi * l.dphidr[p];
    const bool live = gr != 0.0f;
    gdot[p] = live ? gr * l.inrm[p] : 0.0f;
    const float gnrm = live ? -gr * l.r[p] * l.inrm[p] : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) gv[k] += 2.0f * gnrm * v[k];
    if (p == 0) { gW[0] += gv[0]; gW[1] += gv[1]; gW[2] += gv[2]; }
    else if (p == 1) { gW[3] += gv[0]; gW[4] += gv[1]; gW[5] += gv[2]; }
    else { gW[6] += gv[2]; }
  }
}

}  // namespace fvad

#endif  // FIVONAGI_APPS_NA_ADJOINT_H
