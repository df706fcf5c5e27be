// apps/na/solver.h -- the "na" (nonlinear acoustics) plug-in as device code:
// Roe-type Riemann solver, MC limiter and correction flux, delta-diffusion
// source.  These are the per-interface / per-cell bodies that the fused step
// kernel (csrc/fv_step.hip) calls; they are not kernels themselves.
//
// Reference bodies restated here:
//   Riemann solve ........ apps/na/fv.h:88-180 (x) and :276-368 (y)
//   MC limiter ........... apps/na/fv.h:579-662 (philim.f)
//   correction flux ...... apps/na/fv.h:666-711
//   first/second order ... apps/na/fv.h:427-434, :979-986
//   source ............... apps/na/source.h:84-104
//
// Arithmetic form (all algebraically identical to the reference):
//   * rho = 1 + rho' is never differenced: drho = rho'_R - rho'_L and
//     Z1 - 1 = (rho'_R + rho'_L)/2 are taken from rho' directly, so FP32 does
//     not lose the ~7 bits the reference loses to the "+1" (fv.h:62-63).
//   * a3 = dt - (a1*A1 + a2*A2), which is fv.h:119-125 after substituting a1, a2.
//   * 1/x through v_rcp_f32 / v_rsq_f32 (1 ulp); L1 - L2 = 2*sqrt(tmp3).
//   * the fluctuations and the correction flux are folded per wave family
//     (see limited_flux).
//   * the two sweep updates (step + writeLimiters) are fused into one:
//     q - dtd*((apdq - fadd)(k) + (amdq + fadd)(k+1)).
#ifndef FIVONAGI_APPS_NA_SOLVER_H
#define FIVONAGI_APPS_NA_SOLVER_H

#include <hip/hip_runtime.h>

namespace fvna {

// The bodies below are templates on the real type R: float is the
// production path (FP32 with the v_rcp/v_rsq/v_med3 forms), double the FP64
// path (PRECISION=2 of the reference).  In double, 1/x and 1/sqrt(x) are the
// hardware estimates refined by two Newton steps each (within an ulp of the
// IEEE quotient the reference takes, at less than half the instructions of
// the correctly rounded sequences).  Their arguments here are O(1) away from
// zero (2 Z1, (L1 S1 - s3)(L2 S1 - s3), 1 + rho', the discriminant) except
// the limiter's |W|^2, where 0 gives NaN as the IEEE form's 0 * inf does.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);          // e = 1 - x y
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double r = fma(-x * y, y, 1.0);      // r = 1 - x y^2
  y = fma(0.5 * y, r, y);
  r = fma(-x * y, y, 1.0);
  return fma(0.5 * y, r, y);
}
__device__ __forceinline__ float fma_(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float min_(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double min_(double a, double b) { return fmin(a, b); }

// a + b never contracted with a product feeding it, so that when a is the
// neighbouring lane's value its cross-lane move can fold into this add
template <typename R>
__device__ __forceinline__ R add_nc(R a, R b) {
#pragma clang fp contract(off)
  return a + b;
}

// max(a, b) of finite values as one v_med3_f32 (median of a, b, FLT_MAX;
// with +inf the intrinsic is folded back into a maxnum).
// fmaxf must quiet signalling NaNs in IEEE mode, so on values the compiler
// cannot prove canonical (registers carried around the march loop) it adds a
// canonicalising v_max per operand; the med3 intrinsic is taken as is.
__device__ __forceinline__ float fmax2(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 3.4028234663852886e38f);
}
__device__ __forceinline__ double fmax2(double a, double b) { return fmax(a, b); }
// clamp to [0, 1]; a NaN gives 0
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
__device__ __forceinline__ double clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }

// Waves of one interface in (density, normal, tangential) component order.
// W1 = a1*(1, L1, A1), W2 = a2*(1, L2, A2), W3 = (0, 0, a3).
template <typename R>
struct WavesT {
  R a1, w21, w31;       // wave 1: W11, W21, W31
  R a2, w22, w32;       // wave 2: W12, W22, W32
  R a3;                 // wave 3: W33 (W13 = W23 = 0)
};

template <typename R>
struct IfaceT {
  WavesT<R> w;
  R L1, L2, L3;         // speeds s1, s2, s3
  R smax;               // max |s_p| (|L3| <= max(|L1|,|L2|) since L3 is their mean)
  bool bad;             // tmp3 < 0: imaginary eigenvalue (simError = 2)
};
typedef WavesT<float> Waves;
typedef IfaceT<float> Iface;

// Riemann solve at the interface between cell L and cell R.
// r1 = rho', n = normal momentum, t = tangential momentum,
// un/ut = n/rho, t/rho of the same cell (precomputed once per cell).
// hbeta = BETA - 1.  The fluctuations amdq/apdq (fv.h:141-180) are not
// formed here: limited_flux() folds them with the correction flux.
// Every L operand enters only sums and differences with its R partner, so
// when L is the neighbouring lane's value the cross-lane move folds into
// the add/subtract (DPP operand).
template <typename R>
__device__ __forceinline__ void riemann(R r1L, R nL, R tL, R unL, R utL, R r1R, R nR, R tR,
                                        R unR, R utR, R hbeta, IfaceT<R>& f) {
  const R s1 = add_nc(r1L, r1R);                    // 2 (Z1 - 1)   fv.h:91
  const R s2 = add_nc(unL, unR);                    // 2 Z2         fv.h:92
  const R s3 = add_nc(utL, utR);                    // 2 Z3         fv.h:93
  const R drho = r1R - r1L, dn = nR - nL, dtg = tR - tL;
  const R S1 = s1 + R(2);                           // 2 Z1
  const R tmp1 = s2 * rcp(S1);                      // Z2/Z1        fv.h:96
  // tmp1^2 - 2 Z2 tmp1 + 1 + 2 (BETA-1)(Z1-1)                      fv.h:97-101
  const R tmp3 = fma_(tmp1, tmp1 - s2, fma_(hbeta, s1, R(1)));
  f.bad = !(tmp3 >= R(0));                          // fv.h:102
  const R rs = rsq(tmp3);
  const R sq = tmp3 * rs;                           // sqrt(tmp3)
  const R L1 = tmp1 + sq, L2 = tmp1 - sq;           // fv.h:105-107
  f.L1 = L1; f.L2 = L2; f.L3 = tmp1;
  // fv.h:110-113: A_p = (L_p Z3 - 2 Z2 Z3)/(L_p Z1 - Z3) = s3 (L_p - s2)/(L_p S1 - s3),
  // with one reciprocal for both denominators (each is ~ +-2)
  const R d1 = fma_(L1, S1, -s3), d2 = fma_(L2, S1, -s3);
  const R srd = s3 * rcp(d1 * d2);
  const R A1 = (L1 - s2) * (d2 * srd);
  const R A2 = (L2 - s2) * (d1 * srd);
  const R a1 = fma_(-drho, L2, dn) * (R(0.5) * rs);             // fv.h:116-117
  const R a2 = drho - a1;                                       // fv.h:118
  f.w.a1 = a1; f.w.w21 = a1 * L1; f.w.w31 = a1 * A1;
  f.w.a2 = a2; f.w.w22 = a2 * L2; f.w.w32 = a2 * A2;
  f.w.a3 = dtg - (f.w.w31 + f.w.w32);                           // fv.h:119-125
  f.smax = fmax2(fabs(L1), fabs(L2));
}

// MC limiter phi(r) = max(0, min((1+r)/2, 2, 2r)) (fv.h:587-605), returned
// halved: phi/2 = clamp(min((1+r)/4, r), 0, 1), where the clamp is the VOP3
// output modifier of the min.  The reference sets phi = 1 where wnorm2 == 0;
// such a wave is the zero vector, so its correction phi*W is 0 for any
// finite phi.  Here wnorm2 == 0 gives r = NaN (0 * inf), which the clamp
// turns into 0 (finite), and the select is skipped.
template <typename R>
__device__ __forceinline__ R philim_half(R wnorm2, R dot) {
  const R r = dot * rcp(wnorm2);
  return clamp01(min_(fma_(R(0.25), r, R(0.25)), r));
}

// W_p(a) . W_p(b); b is the operand that may come from a neighbouring lane
// (mul, then multiply-accumulates with b as the plain multiplicand)
template <typename R>
__device__ __forceinline__ R dot1(const WavesT<R>& a, const WavesT<R>& b) {
  return fma_(b.a1, a.a1, fma_(b.w21, a.w21, b.w31 * a.w31));
}
template <typename R>
__device__ __forceinline__ R dot2(const WavesT<R>& a, const WavesT<R>& b) {
  return fma_(b.a2, a.a2, fma_(b.w22, a.w22, b.w32 * a.w32));
}

// Fluctuations (fv.h:141-180) and correction flux fadd (fv.h:666-711) of one
// interface, folded into what the two adjacent cells need:
//   P = apdq - fadd = sum_p (max(s_p,0) - c_p) W_p   (cell on the + side)
//   M = amdq + fadd = sum_p (min(s_p,0) + c_p) W_p   (cell on the - side)
// with c_p = |s_p| (1 - |s_p| dt/d) phi_p / 2 = |s_p| (1 - |s_p| dt/d) philim_half.  A wave with s_p = 0 enters
// neither fluctuation, as in the reference.
// dotl_p = W_p . W_p(k-1), dotr_p = W_p . W_p(k+1); upwind by sign(s_p).
template <typename R>
__device__ __forceinline__ void limited_flux(const IfaceT<R>& f, R dotl1, R dotr1, R dotl2,
                                             R dotr2, R dotl3, R dotr3, R dtd, R P[3],
                                             R M[3]) {
  const WavesT<R>& w = f.w;
  const R n1 = fma_(w.a1, w.a1, fma_(w.w21, w.w21, w.w31 * w.w31));
  const R n2 = fma_(w.a2, w.a2, fma_(w.w22, w.w22, w.w32 * w.w32));
  const R n3 = w.a3 * w.a3;
  const R h1 = philim_half(n1, f.L1 > R(0) ? dotl1 : dotr1);
  const R h2 = philim_half(n2, f.L2 > R(0) ? dotl2 : dotr2);
  const R h3 = philim_half(n3, f.L3 > R(0) ? dotl3 : dotr3);
  const R s1 = fabs(f.L1), s2 = fabs(f.L2), s3 = fabs(f.L3);
  const R c1 = s1 * fma_(-s1, dtd, R(1)) * h1;
  const R c2 = s2 * fma_(-s2, dtd, R(1)) * h2;
  const R c3 = s3 * fma_(-s3, dtd, R(1)) * h3;
  // al_p = max(s_p, 0) - c_p; be_p = min(s_p, 0) + c_p = s_p - al_p
  const R al1 = fmax2(f.L1, R(0)) - c1, be1 = f.L1 - al1;
  const R al2 = fmax2(f.L2, R(0)) - c2, be2 = f.L2 - al2;
  const R al3 = fmax2(f.L3, R(0)) - c3, be3 = f.L3 - al3;
  P[0] = fma_(al1, w.a1, al2 * w.a2);
  P[1] = fma_(al1, w.w21, al2 * w.w22);
  P[2] = fma_(al3, w.a3, fma_(al1, w.w31, al2 * w.w32));
  M[0] = fma_(be1, w.a1, be2 * w.a2);
  M[1] = fma_(be1, w.w21, be2 * w.w22);
  M[2] = fma_(be3, w.a3, fma_(be1, w.w31, be2 * w.w32));
}

}  // namespace fvna

#endif  // FIVONAGI_APPS_NA_SOLVER_H
