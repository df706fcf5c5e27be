/* oracle_check.c -- every entry point of the CPU oracle (oracle/fvoracle.c,
 * oracle/fvadjoint.c) run on small grids of each boundary kind, built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile):
 * forward runs in FP64 and FP32 (CFL feedback and prescribed dt, moving
 * window), data collection of both apps, the single pieces, the adjoint, its
 * trajectory form and the objective.  What the results must equal is
 * checked elsewhere (tests/test_oracle*.py, tests/test_adjoint.py); here a
 * run must finish with its expected return code and no sanitizer report.
 * Prints "ok" and exits 0. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/fvoracle.h"

static int failures = 0;
#define CHECK(cond, what)                                          \
  do {                                                             \
    if (!(cond)) {                                                 \
      fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, what); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

static orc_params params(int nx, int ny, int bc) {
  orc_params p;
  memset(&p, 0, sizeof(p));
  p.nx = nx; p.ny = ny;
  p.dx = 0.05; p.dy = 0.05;
  p.beta = 3.5; p.delta = 5e-4;
  p.cfl_max = 1.0; p.cfl_wish = 0.9;
  p.dt0 = 0.01; p.T0 = 0.0;
  p.bc = bc; p.cfl_mode = ORC_CFL_REFERENCE_FOLD; p.form = ORC_FORM_RHO1; p.nthreads = 1;
  p.omega = 6.0;
  p.tay_rhoa = 0.01; p.tay_c = 20.0; p.tay_theta = 0.3; p.tay_yoff = 0.0; p.tay_jc = ny / 2;
  p.hifu_ampl = 1e-3; p.hifu_frec = 1.0; p.hifu_a = 0.5; p.hifu_f = 30.0;
  p.T_stop = INFINITY;
  return p;
}

static void init_state(int nx, int ny, double *q) {
  size_t n = (size_t)nx * ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t o = (size_t)i + (size_t)j * nx;
      q[o] = 1e-3 * sin(0.31 * i + 0.17 * j) + 5e-3 * exp(-((i - nx / 2.0) * (i - nx / 2.0) +
                                                             (j - ny / 2.0) * (j - ny / 2.0)) / 30.0);
      q[n + o] = 1e-3 * cos(0.29 * i + 0.13 * j);
      q[2 * n + o] = 1e-3 * sin(0.19 * j + 0.07 * i);
    }
}

int main(void) {
  const int nx = 40, ny = 36;
  const size_t n = (size_t)nx * ny;
  double amp[36], ph[36];
  for (int j = 0; j < ny; ++j) { amp[j] = 0.01 * exp(-(j - 18.0) * (j - 18.0) / 40.0); ph[j] = 0.1 * j; }
  double *q = malloc(sizeof(double) * 3 * n), *o = malloc(sizeof(double) * 3 * n);
  float *qf = malloc(sizeof(float) * 3 * n);
  double *measure = calloc(n, sizeof(double));
  for (int bc = 0; bc < 4; ++bc) {
    for (int md = 0; md < 2; ++md) {
      orc_params p = params(nx, ny, bc);
      if (bc == ORC_BC_EDGE_DRIVE) { p.edge_amp = amp; p.edge_phase = ph; }
      p.md_on = md; p.md_vx = 1.0; p.md_vy = 0.2; p.md_t = 0.05;
      if (bc == ORC_BC_HIFU || bc == ORC_BC_TAYLOR) {
        p.measure = measure; p.eta = 4.0; p.jcut = ny / 2;
        p.mic[0] = 5; p.mic[1] = 10; p.mic[2] = 15; p.mic[3] = 20;
        p.cut_half = 4; p.tan_theta = tan(0.3);
      }
      orc_state st;
      init_state(nx, ny, q);
      orc_state_init(&p, &st);
      int rc = orc_run_f64(&p, q, q + n, q + 2 * n, 12, &st);
      CHECK(rc == ORC_OK, "orc_run_f64");
      for (size_t k = 0; k < 3 * n; ++k) qf[k] = (float)q[k];
      orc_state_init(&p, &st);
      rc = orc_run_f32(&p, qf, qf + n, qf + 2 * n, 12, &st);
      CHECK(rc == ORC_OK, "orc_run_f32");
      if (p.measure) {
        st.T = 0.5; st.n_accepted = 3; st.last_cfl = 0.5;
        rc = orc_collect_f64(&p, q, &st);
        CHECK(rc == ORC_OK, "orc_collect_f64");
      }
      rc = orc_boundary_f64(&p, q, q + n, q + 2 * n, o, o + n, o + 2 * n, 0.3);
      CHECK(rc == ORC_OK, "orc_boundary_f64");
      rc = orc_source_f64(&p, q, q + n, q + 2 * n, o, o + n, o + 2 * n, 0.01);
      CHECK(rc == ORC_OK, "orc_source_f64");
      for (int axis = 0; axis < 2; ++axis) {
        double *cfl = malloc(sizeof(double) * n);
        rc = orc_sweep_f64(&p, axis, q, q + n, q + 2 * n, 0.01, cfl);
        CHECK(rc == ORC_OK, "orc_sweep_f64");
        free(cfl);
      }
    }
  }
  /* the adjoint, its trajectory form and the objective (prescribed dt) */
  {
    orc_params p = params(nx, ny, ORC_BC_EDGE_DRIVE);
    p.edge_amp = amp; p.edge_phase = ph;
    double dts[16];
    for (int k = 0; k < 16; ++k) dts[k] = 0.02;
    p.dt_seq = dts; p.dt_seq_len = 16;
    orc_box box = {10, 12, 8, 9};
    const int steps = 10;
    double *q0 = malloc(sizeof(double) * 3 * n);
    init_state(nx, ny, q0);
    double J = 0, J2 = 0, J3 = 0, ga[36], gp[36];
    double *gq = malloc(sizeof(double) * 3 * n);
    int rc = orc_adjoint_f64(&p, q0, steps, &box, &J, ga, gp, gq);
    CHECK(rc == ORC_OK && J > 0, "orc_adjoint_f64");
    double *states = malloc(sizeof(double) * 3 * n * (steps + 1));
    memcpy(states, q0, sizeof(double) * 3 * n);
    double T = p.T0;
    for (int k = 0; k < steps; ++k) {
      memcpy(states + (size_t)(k + 1) * 3 * n, states + (size_t)k * 3 * n, sizeof(double) * 3 * n);
      double *s = states + (size_t)(k + 1) * 3 * n;
      rc = orc_line_step_f64(&p, s, s + n, s + 2 * n, T, dts[k]);
      CHECK(rc == ORC_OK, "orc_line_step_f64");
      T += dts[k];
    }
    rc = orc_adjoint_path_f64(&p, states, steps, &box, &J2, ga, gp, gq);
    CHECK(rc == ORC_OK && J2 == J, "orc_adjoint_path_f64");
    rc = orc_objective_f64(&p, q0, steps, &box, &J3);
    CHECK(rc == ORC_OK && fabs(J3 - J) <= 1e-9 * J, "orc_objective_f64");
    /* a box outside the grid is refused */
    orc_box bad = {nx - 2, 0, 8, 8};
    CHECK(orc_adjoint_f64(&p, q0, steps, &bad, &J, ga, gp, gq) == ORC_ERR_ARG, "box refused");
    free(q0); free(gq); free(states);
  }
  {
    uint8_t *m = malloc(n);
    orc_cfl_subset(nx, ny, m);
    CHECK(m[0] == 1, "orc_cfl_subset");
    free(m);
    double out[18], qL[3] = {1e-3, 2e-4, -1e-4}, qR[3] = {2e-3, -1e-4, 3e-4};
    CHECK(orc_riemann_f64(qL, qR, 3.5, ORC_FORM_RHO1, out) == ORC_OK, "orc_riemann_f64");
    CHECK(orc_philim_f64(1.0, 0.3, 0.2, 1.0) >= 0.0, "orc_philim_f64");
  }
  free(q); free(o); free(qf); free(measure);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
