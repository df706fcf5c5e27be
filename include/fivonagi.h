/*
 * fivonagi.h -- C ABI of the MI355X-native FiVoNAGI time step.
 *
 * The reference has no library boundary: an app is compiled into the step
 * loop (apps/<app>/driver.cu:40-52 includes parameters.h, init.h, boundary.h,
 * apps/na/{source,fv}.h and common/main.cu), memory is owned by gpu_data
 * (common/data_definitions.h:17-89, gpu_init :523-871, clean_gpu :415-517)
 * and the loop is calcFrame (common/main.cu:248-756).  These entry points
 * replace that surface; each cites what it stands for.
 *
 * Conventions
 *   - State is SoA FP32, three grids per instance, offset i + j*nx, x fastest,
 *     ghost frame (2 wide) inside the nx*ny grid; q1 = rho' = rho - 1,
 *     q2 = rho*u, q3 = rho*v (apps/hifu-beam/init.h:18-20).
 *   - Host arrays are copied; the context owns every device allocation.
 *   - Return codes keep the reference's simError numbering: 0 ok,
 *     2 imaginary eigenvalue (apps/na/fv.h:102); rejected steps (simError 1,
 *     common/main.cu:207-211) are retried internally and counted.
 *     8 = bad argument, 9 = prescribed dt sequence exhausted,
 *     negative = -(hipError_t).  Nothing calls exit().
 *   - One context per stream; a context is used by one thread at a time; the
 *     library starts no threads.
 */
#ifndef FIVONAGI_H
#define FIVONAGI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fivonagi_ctx fivonagi_ctx;

enum {
  FIVONAGI_BC_ZEROGRAD = 0,   /* common/boundary.h:11-82                        */
  FIVONAGI_BC_EDGE_DRIVE = 1, /* apps/hifu-beam/boundary.h:77-93, nu = (1,0)    */
  FIVONAGI_BC_TAYLOR = 2,     /* apps/taylor-angle-xy/boundary.h:38-55          */
  FIVONAGI_BC_HIFU = 3        /* apps/hifu-beam/boundary.h:9-133 (full app)     */
};
enum {
  FIVONAGI_CFL_REFERENCE_FOLD = 0, /* reduceCFL schedule, common/main.cu:503-524 */
  FIVONAGI_CFL_ALL = 1             /* exact max over every cell                  */
};
enum {
  FIVONAGI_OK = 0, FIVONAGI_ERR_EIGEN = 2, FIVONAGI_ERR_ARG = 8,
  FIVONAGI_ERR_DTSEQ = 9, FIVONAGI_ERR_STALL = 10, FIVONAGI_ERR_COLLECT = 4,
  FIVONAGI_ERR_CHECK = 11  /* fivonagi_step_adjoint with check_recompute: a
                              recomputed state differs from the forward's */
};

/* Per-instance parameters: the runtime form of apps/<app>/parameters.h
   (e.g. apps/hifu-beam/parameters.h:12-104) plus the boundary choice. */
typedef struct {
  int32_t nx, ny;          /* NX, NY (>= 8)                                   */
  double dx, dy;           /* XMAX/(NX-1), YMAX/(NY-1) (common/main.cu:789-790) */
  double beta, delta;      /* BETA, DIFFDELTA                                 */
  double cfl_max, cfl_wish;/* CFLMAX, CFLWHISH                                */
  double dt0, T0;          /* DTINI, -TIDE (common/main.cu:788, :792)         */
  int32_t bc, cfl_mode;    /* FIVONAGI_BC_*, FIVONAGI_CFL_*                   */
  double omega;            /* edge drive: p_j = amp_j sin(omega T + phase_j)  */
  const double *edge_amp;  /* host, ny values (NULL = 0); copied by init       */
  const double *edge_phase;/* host, ny values (NULL = 0); copied by init       */
  double tay_rhoa, tay_c, tay_theta, tay_yoff; /* Taylor closed form          */
  int32_t tay_jc;
  int32_t rows_per_wave;   /* 0 = automatic; else rows per work item          */
  /* Moving window (common/main.cu:109-177): after every accepted step with
     T >= md_t the grid content shifts so that the window sits
     int((T - md_t) * md_v / d) cells from its start (MDX, MDY); cells whose
     source falls off the far edge keep their value.  md_v >= 0.  The edge
     drive runs only while MDX == 0 (apps/hifu-beam/boundary.h:82); the Taylor
     closed form is taken at (i + MDX, j + MDY). */
  int32_t md_on;           /* 0 = fixed domain                                */
  double md_vx, md_vy, md_t; /* MDVX, MDVY, MDT                               */
  /* FIVONAGI_BC_HIFU (apps/hifu-beam/boundary.h:9-133): focused aperture on
     the left ghost columns, rows 2 <= j < 2 + int(hifu_a/dy) while MDX == 0:
     phi_j = dx (j-2)^2 / (2 hifu_f), pack = hifu_ampl sin(2 pi hifu_frec (T+phi_j))
     exp(-(hifu_frec (T+phi_j)/4)^10), direction nu_j = (F, -(j-2))/|.|;
     rows 0,1 mirror rows 4,3 with rho v negated (beam axis); corners 0;
     right/top edges and the other left ghosts zero-gradient.  omega and the
     edge_* arrays are not used. */
  double hifu_ampl, hifu_frec; /* AMPL, FREC                                  */
  double hifu_a, hifu_f;   /* ISPa (length units), ISPF (points)              */
} fivonagi_params;

typedef struct {
  double T, dt, last_cfl;
  int64_t n_accepted, n_rejected, n_attempts;
  int32_t error, cur;
  int32_t n_items, rows_per_wave; /* work items per step, mean rows per item */
  double skew;             /* height skew of the work items in use            */
  int32_t rotate;          /* 1: skew levels rotated per strip                */
  int32_t phase;           /* next step launch mod 6 (captured graphs replay
                              from phase 0)                                   */
  int32_t md_x, md_y;      /* moving-window offset MDX, MDY                   */
  int32_t fp64, pad2;      /* 1: FP64 context (fivonagi_init_f64)             */
} fivonagi_info_t;

/* gpu_init (common/data_definitions.h:523-871) + init() upload
   (apps/APP/init.h): n_inst instances, p[k] and q_host[3k..3k+2] (q1,q2,q3).
   Each instance: nx, ny >= 8 and one field within 4 GiB (nx*ny <= 2^30 cells;
   2^29 for fivonagi_init_f64); larger grids return FIVONAGI_ERR_ARG. */
int fivonagi_init(fivonagi_ctx **ctx, const fivonagi_params *p, int n_inst,
                  const float *const *q_host);
/* The same with the reference's PRECISION=2 (common/main.cu:11-24): state,
   tables, arithmetic, the cfl reduction and the data collection in FP64
   (read back with the _f64 calls).  The adjoint is FP32 only. */
int fivonagi_init_f64(fivonagi_ctx **ctx, const fivonagi_params *p, int n_inst,
                      const double *const *q_host);

/* Run on the caller's HIP stream (hipStream_t) instead of the context's own. */
int fivonagi_set_stream(fivonagi_ctx *ctx, void *hip_stream);

/* Prescribed dt: attempt k of instance inst uses dt[k]; no CFL feedback and
   no rejection (NULL/0 restores the feedback of common/main.cu:214-230).
   A sequence of n entries runs attempts up to index n-1; a step call that
   needs an attempt past the end returns FIVONAGI_ERR_DTSEQ.  The sequence
   must hold an entry for the next attempt (FIVONAGI_ERR_ARG otherwise, with
   nothing changed). */
int fivonagi_set_dt_sequence(fivonagi_ctx *ctx, int inst, const double *dt, int64_t n);

/* The while (T < Tprint) body of common/main.cu:310-635, nsteps accepted
   steps for every instance (rejected attempts retried).  Blocking. */
int fivonagi_step(fivonagi_ctx *ctx, int64_t nsteps);

/* The frame loop `while (T < Tprint) { step }` of calcFrame
   (common/main.cu:310): attempts until every instance has T >= T_stop
   (rejected attempts retried).  Blocking. */
int fivonagi_step_until(fivonagi_ctx *ctx, double T_stop);

/* Asynchronous: enqueue nattempts step attempts (graph replays) on the
   context stream; no host synchronisation.  For timing loops. */
int fivonagi_enqueue(fivonagi_ctx *ctx, int64_t nattempts);
/* Wait for the stream; return the first instance error code (0 if none). */
int fivonagi_sync(fivonagi_ctx *ctx);

/* Autotuner (no counterpart in the reference): time trial_steps step
   attempts for each candidate height of the work items (rows per wave), then
   for each height skew at the best height, keep the fastest, and restore the
   state and counters as they were.  The trials run with data collection and
   prescribed dt sequences switched off (both are restored, untouched).
   rows_per_wave_out (may be NULL) receives the height; fivonagi_info reports
   the skew. */
int fivonagi_autotune(fivonagi_ctx *ctx, int trial_steps, int32_t *rows_per_wave_out);

/* Work split (no counterpart in the reference): skew of the work-item
   heights.  0 = equal heights; s in (0, 0.95]: as many height levels as
   waves each SIMD starts (at most its resident count), from (1+s)H to
   (1-s)H, the tallest launched first, which evens out when the waves sharing
   a SIMD finish; < 0 = automatic (0.5 when every SIMD starts its full
   complement of step waves, else 0).  rotate != 0 shifts the level pattern by
   one per strip (same count per level; strips no longer share cut rows).
   Which is fastest depends on the grid: fivonagi_autotune tries them.
   Results do not depend on it. */
int fivonagi_set_skew(fivonagi_ctx *ctx, double skew, int rotate);

/* dataCollect/dataExport stand-in (apps/hifu-beam/data_export.h:29-75):
   copy instance inst's state and counters to the host.  Any pointer may be NULL. */
/* (An FP64 context's state is rounded to float here; fivonagi_collect_f64
   returns it in full, and an FP32 context's widened.) */
int fivonagi_collect(fivonagi_ctx *ctx, int inst, float *q1, float *q2, float *q3,
                     double *T, double *dt, int64_t *n_accepted, int64_t *n_rejected);
int fivonagi_collect_f64(fivonagi_ctx *ctx, int inst, double *q1, double *q2, double *q3,
                         double *T, double *dt, int64_t *n_accepted, int64_t *n_rejected);
int fivonagi_info(fivonagi_ctx *ctx, int inst, fivonagi_info_t *info);

/* ---- discrete adjoint (no counterpart in the reference) -------------------
   Objective per instance: J = sum_n dt_n * sum_{box} rho'_{n+1}^2 over the
   nsteps accepted steps of the call.  Parameters: the left-edge drive
   amplitude and phase of every row (FIVONAGI_BC_EDGE_DRIVE), and the state at
   the start of the call.  The gradient is that of the executed branch path
   with the dt sequence frozen (dt of each accepted step taken as given). */
typedef struct {
  int32_t box_x0, box_y0, box_nx, box_ny;  /* cells [x0, x0+nx) x [y0, y0+ny) */
} fivonagi_objective;

typedef struct {
  int32_t checkpoint_every;  /* 1 = keep every state of the forward run (no
                                recompute); K > 1 = keep one every K steps and
                                recompute the segments between them with the
                                step kernel, each step from the control block
                                the forward step ran with; 0 = every state
                                when they fit in half the free device memory,
                                else K = ceil(sqrt(nsteps))                   */
  int32_t plain_forward;     /* 1 = forward with one launch per step and a host
                                readback of dt after each (no graph replay)   */
  int32_t check_recompute;   /* 1 = the reverse pass also recomputes every
                                q_{n+1} and compares it bit for bit with the
                                stored state; any difference returns
                                FIVONAGI_ERR_CHECK                            */
  int32_t rows_per_wave;     /* rows per work item of the reverse march, one
                                item per wave; 0 = waves of about equal cost,
                                as many as start at once, a wave taking the
                                end of one strip and the start of the next    */
} fivonagi_adjoint_opts;

/* Advance every instance nsteps accepted steps (as fivonagi_step, recording
   dt and T), then run the adjoint backwards.  obj: n_inst entries.  J: n_inst
   values.  grad_amp / grad_phase: dJ/d edge_amp[j], dJ/d edge_phase[j], the
   instances' ny values concatenated (may be NULL).  The state is left at the
   end of the forward run.  The call's device buffers (kept states or
   checkpoints, adjoint grids, logs) are one block that the context keeps for
   later calls, grown when a call needs more; fivonagi_destroy frees it.  The
   free memory that the checkpoint_every = 0 rule weighs counts that block. */
int fivonagi_step_adjoint(fivonagi_ctx *ctx, int64_t nsteps, const fivonagi_objective *obj,
                          const fivonagi_adjoint_opts *opts, double *J, double *grad_amp,
                          double *grad_phase);
/* dJ/d(state at the start of the last fivonagi_step_adjoint call), 3 grids. */
int fivonagi_adjoint_state(fivonagi_ctx *ctx, int inst, float *l1, float *l2, float *l3);

/* ---- data collection (apps/hifu-beam/data_collect.h:7-107,
   apps/taylor-angle-xy/data_collect.h:7-98) ----------------------------------
   After every accepted step, into a measurement grid of nx*ny values of the
   context's precision, laid out as the reference's measure1: row 0 = T and
   row 1 = cfl of the frame's steps (step n of a frame is entry n - 1 - n0,
   n0 = the step count when the frame started; fivonagi_step_until starts a
   frame).  While T <= 0 the grid is zeroed.  More than nx steps in one frame
   stops the call with code 4 (simError 4).
   hifu app (and every boundary but the Taylor one), on row jcut of the new
   state before the moving-window shift: rows 2..5 = u1 at the four
   microphones (x index mic[k] of the unmoved grid, i.e. i + MDX == mic[k]),
   from row 6 on = max of u1 over the steps where |(i + MDX)/eta - T| <= 0.5,
   at flat index i + MDX + 6 nx.
   taylor-angle-xy app (bc = FIVONAGI_BC_TAYLOR), along the propagation
   direction through the front's centre icenter = int(T MDVX/dx) - MDX,
   jcenter = NY/2 + int((T - MDT) MDVY/dy) - MDY, for i in
   [icenter - cut_half, icenter + cut_half]: row 2 = rho' interpolated in y at
   ja = jcenter + (i - icenter)(dx/dy) tan(theta), row 3 = the closed form at
   the nearer row, entry i - icenter + cut_half.  MDVX, MDVY, MDT are the
   params' md_vx, md_vy, md_t (also when md_on = 0). */
typedef struct {
  int32_t jcut;            /* JCUT (2 in the hifu app)                        */
  int32_t mic[4];          /* microphone x indices: round(ISPF*{.5,.9}), ISPF, round(ISPF*1.1) */
  double eta;              /* ETA = NX/XMAX cells per length unit             */
  int32_t cut_half;        /* Taylor: static_cast<int>(5.0*ETA*cos(ISPTHETA)) */
  int32_t pad;
  double tan_theta;        /* Taylor: tan(ISPTHETA)                           */
} fivonagi_collect_opts;
/* opts = NULL switches collection off for the instance. */
int fivonagi_set_collect(fivonagi_ctx *ctx, int inst, const fivonagi_collect_opts *opts);
/* Copy the measurement grid (nx*ny values) of instance inst to the host:
   floats from an FP32 context, doubles (the _f64 form) from an FP64 one. */
int fivonagi_measure(fivonagi_ctx *ctx, int inst, float *measure);
int fivonagi_measure_f64(fivonagi_ctx *ctx, int inst, double *measure);
/* The first `rows` rows only (rows*nx values), e.g. rows 0..3 of a Taylor
   instance each frame (apps/taylor-angle-xy/data_export.h:77-176 reads no more). */
int fivonagi_measure_rows(fivonagi_ctx *ctx, int inst, int rows, float *measure);
int fivonagi_measure_rows_f64(fivonagi_ctx *ctx, int inst, int rows, double *measure);

/* Host-side: the cells whose cfl reaches cfl[0] in the reference's fold
   (reduceCFL and its schedule, common/main.cu:179-212, :503-524), which
   FIVONAGI_CFL_REFERENCE_FOLD reduces over: mask[i + j*nx] = 1 for those
   cells (nx*ny bytes). */
int fivonagi_cfl_subset(int nx, int ny, uint8_t *mask);

/* Host-side: the relative L1, L2 and Linf errors of a Taylor cut
   (apps/taylor-angle-xy/data_export.h:96-122: sum|a-n|/sum|a|,
   sqrt(sum (a-n)^2/sum a^2), max|a-n|/max|a|) with the reference's float
   accumulators, over entries 0..length-1 of the numeric and analytic rows
   (rows 2 and 3 of the measurement grid; length = 2 cut_half). */
int fivonagi_cut_errors(const float *numeric, const float *analytic, int length,
                        float errors[3]);
int fivonagi_cut_errors_f64(const double *numeric, const double *analytic, int length,
                            float errors[3]);

/* ---- step without device allocation ---------------------------------------
   No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  fivonagi_step for one FP32 instance whose state the
   caller keeps on the device: the instance description, control blocks, CFL
   reduction slots, edge tables and work lists live in one fixed device buffer
   of FIVONAGI_NOALLOC_SCRATCH_BYTES per GPU, set up on the first call on a
   device and held until the process ends.  The call allocates no device
   memory (no hipMalloc, hipFree or allocating library routine).
     q[0..2]  device pointers: rho', rho u, rho v (nx*ny floats each, offset
              i + j*nx, ghost frame inside): the state, in and out;
     q[3..5]  device pointers: a work buffer of the same size (overwritten);
     st       in: T and dt of the next attempt, the counters and MDX, MDY;
              out: the same after the call, and error;
     nsteps   accepted steps (rejected attempts are retried); blocking.
   Every boundary kind and the moving window run as in fivonagi_init; the
   edge arrays of p are read during the call only.  No data collection, no
   prescribed dt.  A grid whose tables do not fit the buffer (the CFL lane
   masks take 8 bytes per 58 cells: beyond about 60 million cells, e.g.
   8192 x 8192) returns FIVONAGI_ERR_ARG.  All calls on one GPU share
   its buffer, which keeps the last call's tables: calls that overlap on one
   GPU (from several host threads, or on streams that run concurrently) race. */
#define FIVONAGI_NOALLOC_SCRATCH_BYTES (8u << 20)
typedef struct {
  double T, dt;            /* time and dt of the next attempt (dt > 0)        */
  int64_t n_accepted, n_rejected, n_attempts;
  int32_t md_x, md_y;      /* moving-window offset MDX, MDY                   */
  int32_t error, pad;      /* out: 0, or FIVONAGI_ERR_EIGEN                   */
} fivonagi_noalloc_state;
int fivonagi_step_noalloc(const fivonagi_params *p, float *const q[6], int64_t nsteps,
                          fivonagi_noalloc_state *st, void *hip_stream);

/* clean_gpu (common/data_definitions.h:415-517) */
void fivonagi_destroy(fivonagi_ctx *ctx);

/* Message of the last error returned on this thread. */
const char *fivonagi_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FIVONAGI_H */
