/*
 * fvoracle.h -- CPU oracle for the FiVoNAGI time step (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it.  It restates, kernel by kernel, the
 * reference step loop of common/main.cu:310-635 with the apps/na solver
 * (apps/na/fv.h, apps/na/source.h) and the ghost-cell rules of
 * common/boundary.h / apps/hifu-beam/boundary.h / apps/taylor-angle-xy/boundary.h.
 *
 * The same source is compiled twice (fvoracle_impl.h): REAL=double with the
 * suffix _f64 (the parity oracle) and REAL=float with the suffix _f32 (the CPU
 * baseline, and the "what does the reference itself get in FP32" reference).
 *
 * Decisions the reference leaves open (each one a flag, run identically on the
 * GPU path):
 *   cfl_mode  REFERENCE_FOLD: emulate the reduceCFL launch schedule of
 *             common/main.cu:503-524 exactly (max over a subset of cells);
 *             ALL: exact max over every cell.
 *   retry     a rejected step restarts from the untouched input state with the
 *             error flag cleared (the reference keeps simError=1 on the device
 *             and retries from a half-updated buffer, common/main.cu:207-209,
 *             :620-623; it livelocks).  Always on.
 *   dt_seq    prescribed dt per attempt: no accept/reject, no dt feedback.
 *   form      LITERAL: rho = rho' + 1 everywhere as written in the reference;
 *             RHO1: differences and means of rho are taken from rho' directly
 *             and source.h's (rho'+1)-1 round trip is skipped.  Identical in
 *             exact arithmetic; RHO1 avoids FP32 cancellation.
 */
#ifndef FVORACLE_H
#define FVORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_ZEROGRAD = 0, ORC_BC_EDGE_DRIVE = 1, ORC_BC_TAYLOR = 2, ORC_BC_HIFU = 3 };
enum { ORC_CFL_REFERENCE_FOLD = 0, ORC_CFL_ALL = 1 };
enum { ORC_FORM_LITERAL = 0, ORC_FORM_RHO1 = 1 };
enum { ORC_OK = 0, ORC_ERR_DT = 1, ORC_ERR_EIGEN = 2, ORC_ERR_COLLECT = 4, ORC_ERR_ARG = 8 };

typedef struct {
  int32_t nx, ny;               /* NX, NY incl. the 2-wide ghost frame        */
  double dx, dy;                /* cell sizes                                  */
  double beta, delta;           /* BETA, DIFFDELTA                             */
  double cfl_max, cfl_wish;     /* CFLMAX, CFLWHISH                            */
  double dt0, T0;               /* DTINI, initial time                         */
  int32_t bc, cfl_mode, form, nthreads; /* nthreads <= 0: OpenMP default      */
  /* ORC_BC_EDGE_DRIVE: ghost columns i in {0,1} hold
       rho' = amp[j] * sin(omega*T + phase[j]), rho*u = rho'(1+rho'), rho*v = 0
     (the pattern of apps/hifu-beam/boundary.h:77-93 with nu = (1,0));
     the other three edges are zero-gradient. */
  double omega;
  const double *edge_amp;       /* length ny, NULL = no drive                  */
  const double *edge_phase;     /* length ny, NULL = zero phase                */
  /* ORC_BC_TAYLOR (apps/taylor-angle-xy/boundary.h:38-55): every ghost cell
     holds rho' = rhoa*tanh(c*(i*dx*cos(th) + ((j-jc)*dy + yoff)*sin(th) - T)),
     rho*u = rho' cos(th), rho*v = rho' sin(th). */
  double tay_rhoa, tay_c, tay_theta, tay_yoff;
  int32_t tay_jc, pad0;
  /* prescribed dt: attempt k uses dt_seq[k]; NULL = CFL feedback */
  const double *dt_seq;
  int64_t dt_seq_len;
  /* moving window (common/main.cu:109-177): after each accepted step with
     T >= md_t the grid shifts by int((T - md_t)*md_v/d) - MD cells (moveDomain,
     cells whose source is off the far edge keep their value) and MD grows by
     that much (updateMDXY).  The edge drive runs only while MDX == 0
     (apps/hifu-beam/boundary.h:82); Taylor ghosts sit at (i+MDX, j+MDY). */
  int32_t md_on;
  /* stop_at_T: orc_run also stops once T >= T_stop (the frame loop's
     while (T < Tprint), common/main.cu:310) */
  int32_t stop_at_T;
  double md_vx, md_vy, md_t;
  double T_stop;
  /* ORC_BC_HIFU: apps/hifu-beam/boundary.h:9-133 -- focused aperture on the
     left ghost columns (rows 2 <= j < 2 + int(hifu_a/dy), while MDX == 0):
     phi_j = dx*(j-2)^2/(2F), pack = AMPL sin(2 pi FREC (T+phi_j))
     exp(-(FREC (T+phi_j)/4)^10), nu_j = (F, -(j-2))/sqrt((j-2)^2 + F^2),
     u1 = pack, (u2, u3) = nu_j pack (pack+1); other left ghosts copy column 2;
     right and top ghosts zero-gradient; rows 0,1 mirror rows 4,3 with u3
     negated (the beam axis); the 2x2 corners are 0.
     hifu_ampl = AMPL, hifu_frec = FREC, hifu_a = ISPa, hifu_f = ISPF (points). */
  double hifu_ampl, hifu_frec, hifu_a, hifu_f;
  /* data collection (apps/hifu-beam/data_collect.h:7-107), after every
     accepted step, before the window shift: measure (nx*ny, NULL = off) as
     the reference's measure1 -- row 0 T, row 1 cfl, rows 2..5 u1 at the
     microphones (i + MDX == mic[k]), from row 6 on max u1 over
     |(i + MDX)/eta - T| <= 0.5 at flat index i + MDX + 6 nx; entry
     n - 1 - n0frame of the frame; zeroed while T <= 0; more than nx steps in
     a frame stops the run with ORC_ERR_COLLECT (simError 4). */
  double *measure;
  int32_t jcut, mic[4], pad1;
  double eta;
  int64_t n0frame;
  /* ORC_BC_TAYLOR instead (apps/taylor-angle-xy/data_collect.h:7-98): rows
     0, 1 as above; for |i - icenter| <= cut_half, icenter =
     int(T md_vx/dx) - MDX, jcenter = ny/2 + int((T - md_t) md_vy/dy) - MDY,
     row 2 entry i - icenter + cut_half = rho' interpolated in y at ja =
     jcenter + (i - icenter)(dx/dy) tan_theta (float), between the rows of the
     thread whose j - ja lies in (-1/2, 1/2], row 3 = the closed form at that
     row.  md_vx, md_vy, md_t are read also when md_on = 0.  collect_f32: T,
     dx, dy, the constants and the closed form in float (the reference built
     with PRECISION=1), else double. */
  int32_t cut_half, collect_f32;
  double tan_theta;
} orc_params;

typedef struct {
  double T, dt, last_cfl;
  int64_t n_accepted, n_rejected, n_attempts;
  int32_t error, pad;
  int32_t mdx, mdy;             /* moving-window offset MDX, MDY            */
} orc_state;

/* Per-interface Roe-type solve (apps/na/fv.h:88-180), generic in the normal
   direction: q = (rho', normal momentum, tangential momentum).  out[0..2]=s,
   out[3..11] = W11,W21,W31,W12,W22,W32,W13,W23,W33 (component-major as in
   fv.h), out[12..14]=amdq, out[15..17]=apdq.  Returns 0 or ORC_ERR_EIGEN. */
int orc_riemann_f64(const double qL[3], const double qR[3], double beta,
                    int form, double out[18]);
/* MC limiter of fv.h:587-605: wnorm2, dot, speed -> phi */
double orc_philim_f64(double wnorm2, double dotl, double dotr, double s);
/* cells whose cfl reaches cfl[0] under the reduceCFL fold (1 = counted) */
void orc_cfl_subset(int32_t nx, int32_t ny, uint8_t *mask);

void orc_state_init(const orc_params *p, orc_state *st);

/* Advance until n_accept more steps are accepted (or error).  q* are SoA
   grids of nx*ny, offset i + j*nx, q1 = rho' = rho - 1. */
int orc_run_f64(const orc_params *p, double *q1, double *q2, double *q3,
                int64_t n_accept, orc_state *st);
int orc_run_f32(const orc_params *p, float *q1, float *q2, float *q3,
                int64_t n_accept, orc_state *st);

/* Single pieces, for the pin tests. */
int orc_boundary_f64(const orc_params *p, const double *in1, const double *in2,
                     const double *in3, double *o1, double *o2, double *o3,
                     double T);
int orc_source_f64(const orc_params *p, const double *in1, const double *in2,
                   const double *in3, double *o1, double *o2, double *o3,
                   double dt);
/* One y-sweep (fv.h:195-499, 722-930, 994-1054) or x-sweep on a state whose
   ghost cells are already set; returns the per-cell cfl in cfl (may be NULL). */
/* dataCollect on the state whose rho' is q1, at st's T, counts and MDX, MDY
   (st->last_cfl = the step's cfl), into p->measure. */
int orc_collect_f64(const orc_params *p, const double *q1, const orc_state *st);
int orc_sweep_f64(const orc_params *p, int axis, double *q1, double *q2,
                  double *q3, double dt, double *cfl);

/* ---- adjoint (fvadjoint.c) ------------------------------------------- */
/* objective box [x0, x0+nx) x [y0, y0+ny) */
typedef struct { int32_t x0, y0, nx, ny; } orc_box;

/* One step of the line-per-sweep formulation used by the adjoint's forward
   recompute (same arithmetic, rho'-direct), for checking it against orc_run. */
int orc_line_step_f64(const orc_params *p, double *q1, double *q2, double *q3, double T,
                      double dt);
/* J = sum_n dt_n sum_box rho'_{n+1}^2 over nsteps steps with the prescribed
   dt sequence p->dt_seq (frozen dt, no rejection); gradients with respect to
   edge_amp[j], edge_phase[j] (ny each) and the initial state (3 grids).
   Any gradient pointer may be NULL. */
int orc_adjoint_f64(const orc_params *p, const double *q0, int64_t nsteps, const orc_box *box,
                    double *J, double *grad_amp, double *grad_phase, double *grad_q0);
/* The same along a given trajectory: states = q_0 .. q_nsteps (3 grids each,
   nsteps + 1 states), the branches differentiated being those of its steps. */
int orc_adjoint_path_f64(const orc_params *p, const double *states, int64_t nsteps,
                         const orc_box *box, double *J, double *grad_amp, double *grad_phase,
                         double *grad_q0);
/* J of the oracle forward (orc_run_f64 step by step) for finite differences */
int orc_objective_f64(const orc_params *p, const double *q0, int64_t nsteps,
                      const orc_box *box, double *J);

#ifdef __cplusplus
}
#endif
#endif
