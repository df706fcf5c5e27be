/*
 * fvoracle.c -- CPU oracle for the FiVoNAGI time step.  TEST INFRASTRUCTURE:
 * only the tests, the smoke check and bench.py's CPU-baseline leg use it; it
 * is the checker of the HIP path, never a fallback for it.
 *
 * Parity status of each piece (what pins it) is listed in tests/test_oracle_pins.py.
 */
#include "fvoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL double
#define SFX(x) x##_f64
#include "fvoracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(x) x##_f32
#include "fvoracle_impl.h"
#undef REAL
#undef SFX

void orc_state_init(const orc_params *p, orc_state *st) {
  memset(st, 0, sizeof(*st));
  st->T = p->T0;
  st->dt = p->dt0;
}

int orc_run_f64(const orc_params *p, double *q1, double *q2, double *q3,
                int64_t n_accept, orc_state *st) {
  return run_impl_f64(p, q1, q2, q3, n_accept, st);
}

int orc_run_f32(const orc_params *p, float *q1, float *q2, float *q3,
                int64_t n_accept, orc_state *st) {
  return run_impl_f32(p, q1, q2, q3, n_accept, st);
}

int orc_riemann_f64(const double qL[3], const double qR[3], double beta,
                    int form, double out[18]) {
  iface_t_f64 f;
  int rc = riemann_f64(qL[0], qL[1], qL[2], qR[0], qR[1], qR[2], beta, form, &f);
  if (rc) return rc;
  for (int k = 0; k < 3; ++k) out[k] = f.s[k];
  for (int k = 0; k < 9; ++k) out[3 + k] = f.W[k];
  for (int k = 0; k < 3; ++k) { out[12 + k] = f.amdq[k]; out[15 + k] = f.apdq[k]; }
  return 0;
}

double orc_philim_f64(double wnorm2, double dotl, double dotr, double s) {
  return philim_f64(wnorm2, dotl, dotr, s);
}

void orc_cfl_subset(int32_t nx, int32_t ny, uint8_t *mask) {
  /* Track which cells the fold of common/main.cu:503-524 carries into
     cfl[0]: walk the folds backwards from {0}. */
  long N = (long)nx * (long)ny;
  long strides[128];
  int nf = 0;
  long sizered = N;
  while (sizered > 32) { sizered /= 2; strides[nf++] = (sizered / 16) * 16; }
  while (sizered > 2) { sizered /= 2; strides[nf++] = sizered; }
  memset(mask, 0, (size_t)N);
  mask[0] = 1;
  for (int f = nf - 1; f >= 0; --f) {
    long S = strides[f];
    for (long i = 0; i < S; ++i)
      if (mask[i] && i + S < N) mask[i + S] = 1;
  }
}

static int check_grid(const orc_params *p) {
  return (p && p->nx >= 8 && p->ny >= 8) ? 0 : ORC_ERR_ARG;
}

int orc_boundary_f64(const orc_params *p, const double *in1, const double *in2,
                     const double *in3, double *o1, double *o2, double *o3,
                     double T) {
  if (check_grid(p)) return ORC_ERR_ARG;
  double *src[3] = {(double *)in1, (double *)in2, (double *)in3};
  double *dst[3] = {o1, o2, o3};
  boundary_f64(p, src, dst, T, 0, 0);
  return 0;
}

int orc_source_f64(const orc_params *p, const double *in1, const double *in2,
                   const double *in3, double *o1, double *o2, double *o3,
                   double dt) {
  if (check_grid(p)) return ORC_ERR_ARG;
  double *src[3] = {(double *)in1, (double *)in2, (double *)in3};
  double *dst[3] = {o1, o2, o3};
  source_f64(p, src, dst, dt);
  return 0;
}

int orc_sweep_f64(const orc_params *p, int axis, double *q1, double *q2,
                  double *q3, double dt, double *cfl) {
  if (check_grid(p) || (axis != 0 && axis != 1)) return ORC_ERR_ARG;
  ctx_t_f64 c;
  if (ctx_alloc_f64(&c, p->nx, p->ny)) { ctx_free_f64(&c); return ORC_ERR_ARG; }
  memcpy(c.A[0], q1, c.n * sizeof(double));
  memcpy(c.A[1], q2, c.n * sizeof(double));
  memcpy(c.A[2], q3, c.n * sizeof(double));
  sweep_f64(&c, p, axis, dt);
  int rc = c.simError;
  memcpy(q1, c.A[0], c.n * sizeof(double));
  memcpy(q2, c.A[1], c.n * sizeof(double));
  memcpy(q3, c.A[2], c.n * sizeof(double));
  if (cfl) memcpy(cfl, c.cfl, c.n * sizeof(double));
  ctx_free_f64(&c);
  return rc;
}

int orc_collect_f64(const orc_params *p, const double *q1, const orc_state *st) {
  if (check_grid(p) || !p->measure || !st) return ORC_ERR_ARG;
  ctx_t_f64 c;
  if (ctx_alloc_f64(&c, p->nx, p->ny)) { ctx_free_f64(&c); return ORC_ERR_ARG; }
  memcpy(c.C[0], q1, c.n * sizeof(double));
  int rc = collect_f64(&c, p, st);
  ctx_free_f64(&c);
  return rc;
}
