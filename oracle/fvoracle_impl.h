/*
 * fvoracle_impl.h -- body of the CPU oracle, included by fvoracle.c once per
 * REAL type (REAL = double, SFX(x) = x##_f64; REAL = float, SFX(x) = x##_f32).
 * TEST INFRASTRUCTURE: see fvoracle.h.
 *
 * Each function below restates one reference kernel as a loop nest over
 * (i, j) with the reference's index guards; the buffer roles of the step loop
 * (A = read/"R", B = write/"W", C = backup) follow common/main.cu:310-635.
 */

typedef struct {
  int nx, ny;
  size_t n;
  REAL *A[3], *B[3], *C[3];
  REAL *s[3];      /* s1, s2, s3                              */
  REAL *W[9];      /* W11,W21,W31,W12,W22,W32,W13,W23,W33     */
  REAL *amdq[3], *apdq[3];
  REAL *cfl;
  int simError;
} SFX(ctx_t);

static REAL SFX(getmax)(REAL x, REAL y) { return (x > y) ? x : y; } /* main.cu:63-65 */
static REAL SFX(getmin)(REAL x, REAL y) { return (x < y) ? x : y; } /* main.cu:67-69 */

static int SFX(ctx_alloc)(SFX(ctx_t) *c, int nx, int ny) {
  memset(c, 0, sizeof(*c));
  c->nx = nx; c->ny = ny; c->n = (size_t)nx * (size_t)ny;
  REAL **all[] = {&c->A[0], &c->A[1], &c->A[2], &c->B[0], &c->B[1], &c->B[2],
                  &c->C[0], &c->C[1], &c->C[2], &c->s[0], &c->s[1], &c->s[2],
                  &c->W[0], &c->W[1], &c->W[2], &c->W[3], &c->W[4], &c->W[5],
                  &c->W[6], &c->W[7], &c->W[8], &c->amdq[0], &c->amdq[1],
                  &c->amdq[2], &c->apdq[0], &c->apdq[1], &c->apdq[2], &c->cfl};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) {
    /* calloc: the reference's scratch grids start at whatever cudaMalloc
       gave; no read of an unwritten cell reaches the state (see below). */
    *all[k] = (REAL *)calloc(c->n, sizeof(REAL));
    if (!*all[k]) return -1;
  }
  return 0;
}

static void SFX(ctx_free)(SFX(ctx_t) *c) {
  REAL **all[] = {&c->A[0], &c->A[1], &c->A[2], &c->B[0], &c->B[1], &c->B[2],
                  &c->C[0], &c->C[1], &c->C[2], &c->s[0], &c->s[1], &c->s[2],
                  &c->W[0], &c->W[1], &c->W[2], &c->W[3], &c->W[4], &c->W[5],
                  &c->W[6], &c->W[7], &c->W[8], &c->amdq[0], &c->amdq[1],
                  &c->amdq[2], &c->apdq[0], &c->apdq[1], &c->apdq[2], &c->cfl};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(*all[k]);
}

/* ---------------------------------------------------------------------------
 * Riemann solve at one interface, generic in direction.
 * x: n = u2 (rho u), t = u3 (rho v)   -- apps/na/fv.h:88-180
 * y: n = u3 (rho v), t = u2 (rho u)   -- apps/na/fv.h:276-368
 * L = the cell at i-1 (x) or j-1 (y); R = the cell (i,j).
 * Outputs are in (density, normal, tangential) component order, as the
 * reference stores them (W11 = a1, W21 = a1*L1, W31 = a1*A1, ...).
 * ------------------------------------------------------------------------- */
typedef struct { REAL s[3], W[9], amdq[3], apdq[3]; REAL amax; } SFX(iface_t);

static int SFX(riemann)(REAL r1L, REAL nL, REAL tL, REAL r1R, REAL nR, REAL tR,
                        REAL beta, int form, SFX(iface_t) *o) {
  const REAL Num0 = 0, Num1 = 1, Num2 = 2;
  /* fv.h:62-63: rho = rho' + 1 at the texture read */
  REAL uR1 = r1R + Num1, uL1 = r1L + Num1;
  REAL Z1, Z2, Z3, Z1m1, drho;
  Z1 = (uR1 + uL1) / Num2;                               /* fv.h:91 */
  Z2 = (nR / uR1 + nL / uL1) / Num2;                     /* fv.h:92 */
  Z3 = (tR / uR1 + tL / uL1) / Num2;                     /* fv.h:93 */
  if (form == ORC_FORM_RHO1) {
    Z1m1 = (r1R + r1L) / Num2;                           /* = Z1 - 1 */
    drho = r1R - r1L;                                    /* = uR1 - uL1 */
  } else {
    Z1m1 = Z1 - Num1;
    drho = uR1 - uL1;
  }
  REAL tmp1 = Z2 / Z1;                                   /* fv.h:96 */
  REAL tmp2 = (Z2 * Z2) / Z1;                            /* fv.h:97 */
  REAL tmp3 = tmp1 * tmp1 - Num2 * tmp2 + Num1
            + Num2 * (beta - Num1) * Z1m1;               /* fv.h:98-101 */
  if (tmp3 < Num0) return ORC_ERR_EIGEN;                 /* fv.h:102 */
  REAL L1 = tmp1 + sqrt(tmp3);                           /* fv.h:105-107 */
  REAL L2 = tmp1 - sqrt(tmp3);
  REAL L3 = tmp1;
  REAL A1 = (L1 * Z3 - Num2 * Z2 * Z3) / (L1 * Z1 - Z3); /* fv.h:110-113 */
  REAL A2 = (L2 * Z3 - Num2 * Z2 * Z3) / (L2 * Z1 - Z3);
  REAL a1 = (nR - nL - drho * L2) / (L1 - L2);           /* fv.h:116-117 */
  REAL a2 = drho - a1;                                   /* fv.h:118 */
  REAL a3 = (((L1 * A2 - L2 * A1) * drho)
             + ((A1 - A2) * (nR - nL))) / (L2 - L1)
            + tR - tL;                                   /* fv.h:119-125 */
  o->s[0] = L1; o->s[1] = L2; o->s[2] = L3;              /* fv.h:127-129 */
  o->W[0] = a1; o->W[1] = a1 * L1; o->W[2] = a1 * A1;    /* fv.h:131-139 */
  o->W[3] = a2; o->W[4] = a2 * L2; o->W[5] = a2 * A2;
  o->W[6] = Num0; o->W[7] = Num0; o->W[8] = a3;
  for (int k = 0; k < 3; ++k) { o->amdq[k] = Num0; o->apdq[k] = Num0; }
  REAL amax = Num0;                   /* max |s| over s != 0, fv.h:148-180 */
  if (L1 < Num0) {
    o->amdq[0] += L1 * a1; o->amdq[1] += L1 * (a1 * L1); o->amdq[2] += L1 * (a1 * A1);
    amax = SFX(getmax)(amax, -L1);
  } else if (L1 > Num0) {
    o->apdq[0] += L1 * a1; o->apdq[1] += L1 * (a1 * L1); o->apdq[2] += L1 * (a1 * A1);
    amax = SFX(getmax)(amax, L1);
  }
  if (L2 < Num0) {
    o->amdq[0] += L2 * a2; o->amdq[1] += L2 * (a2 * L2); o->amdq[2] += L2 * (a2 * A2);
    amax = SFX(getmax)(amax, -L2);
  } else if (L2 > Num0) {
    o->apdq[0] += L2 * a2; o->apdq[1] += L2 * (a2 * L2); o->apdq[2] += L2 * (a2 * A2);
    amax = SFX(getmax)(amax, L2);
  }
  if (L3 < Num0) {
    o->amdq[2] += L3 * a3;
    amax = SFX(getmax)(amax, -L3);
  } else if (L3 > Num0) {
    o->apdq[2] += L3 * a3;
    amax = SFX(getmax)(amax, L3);
  }
  o->amax = amax;
  return 0;
}

/* MC limiter, fv.h:587-605 (philim.f) */
static REAL SFX(philim)(REAL wnorm2, REAL dotl, REAL dotr, REAL s) {
  const REAL Num0 = 0, Num1 = 1, Num2 = 2;
  REAL r, r2, c, w;
  if (wnorm2 != Num0) {
    r = (s > Num0) ? dotl / wnorm2 : dotr / wnorm2;
    r2 = Num2 * r;
    c = (Num1 + r) / Num2;
    w = SFX(getmin)(c, Num2);
    w = SFX(getmin)(w, r2);
    w = SFX(getmax)(w, Num0);
  } else {
    w = Num1;
  }
  return w;
}

/* ---------------------------------------------------------------------------
 * Ghost cells.  dst receives the whole grid: interior copied from src, ghost
 * cells by the rules below (reads come from src only, as the reference reads
 * the other texture, so corner ghosts lag one step).
 * ------------------------------------------------------------------------- */
static void SFX(boundary)(const orc_params *p, REAL *const src[3],
                          REAL *const dst[3], double T, int mdx, int mdy) {
  const int NX = p->nx, NY = p->ny;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      int ir = i, jr = j;
      if (p->bc == ORC_BC_TAYLOR) {
        /* apps/taylor-angle-xy/boundary.h:38-55, at (i + MDX, j + MDY) */
        if (i == NX - 1 || i == NX - 2 || j == 0 || j == 1 || j == NY - 1 ||
            j == NY - 2 || i == 0 || i == 1) {
          double th = p->tay_theta;
          double rhop = p->tay_rhoa *
              tanh(p->tay_c * ((double)(i + mdx) * p->dx * cos(th) +
                               ((double)(j + mdy - p->tay_jc) * p->dy + p->tay_yoff) * sin(th) -
                               T));
          dst[0][o] = (REAL)rhop;
          dst[1][o] = (REAL)(rhop * cos(th));
          dst[2][o] = (REAL)(rhop * sin(th));
        } else {
          for (int f = 0; f < 3; ++f) dst[f][o] = src[f][o];
        }
        continue;
      }
      if (p->bc == ORC_BC_HIFU) {
        /* apps/hifu-beam/boundary.h:20-133, rules in the reference's order
           (later ones override earlier ones) */
        int refl = 0, drive = 0;
        if (i == NX - 1 || i == NX - 2) { ir = NX - 3; jr = j; }
        if (j == NY - 1 || j == NY - 2) { ir = i; jr = NY - 3; }
        if (i == 0 || i == 1) {
          const int width = (int)(p->hifu_a / p->dy);
          const int jprime = j - 2;
          if (abs(jprime) < width && jprime >= 0 && mdx == 0) drive = 1;
          else { ir = 2; jr = j; }
        }
        if (j == 0 || j == 1) { ir = i; jr = 4 - j; refl = 1; drive = 0; }
        if ((i <= 1 && j <= 1) || (i <= 1 && j >= NY - 2) || (i >= NX - 2 && j <= 1) ||
            (i >= NX - 2 && j >= NY - 2)) {
          for (int f = 0; f < 3; ++f) dst[f][o] = (REAL)0;
          continue;
        }
        if (drive) {
          const double F = p->hifu_f, jp = (double)(j - 2), jp2 = jp * jp;
          const double nudenom = sqrt(jp2 + F * F);
          const double nux = F / nudenom, nuy = -jp / nudenom;
          const double phi = (p->dx * jp2) / (2.0 * F);
          const double opT = 2.0 * p->hifu_frec * (T + phi) / 8.0;
          const double o2 = opT * opT, o4 = o2 * o2, o8 = o4 * o4;
          const double ispw = 2.0 * 3.14159265358979323846 * p->hifu_frec;   /* ISPW */
          const double pack = p->hifu_ampl * sin(ispw * (T + phi)) * exp(-(o8 * o2));
          dst[0][o] = (REAL)pack;
          dst[1][o] = (REAL)(nux * pack * (pack + 1.0));
          dst[2][o] = (REAL)(nuy * pack * (pack + 1.0));
          continue;
        }
        size_t orr = (size_t)ir + (size_t)jr * NX;
        dst[0][o] = src[0][orr];
        dst[1][o] = src[1][orr];
        dst[2][o] = refl ? -src[2][orr] : src[2][orr];
        continue;
      }
      /* common/boundary.h:27-81: later rules override earlier ones */
      if (i == NX - 1 || i == NX - 2) { ir = NX - 3; jr = j; }
      if (j == 0 || j == 1) { ir = i; jr = 2; }
      if (j == NY - 1 || j == NY - 2) { ir = i; jr = NY - 3; }
      if (i == 0 || i == 1) {
        if (p->bc == ORC_BC_EDGE_DRIVE && mdx == 0) {
          /* apps/hifu-beam/boundary.h:77-93 with nu = (1, 0):
             u1 = pack, u2 = nux*pack*(pack+1), u3 = nuy*pack*(pack+1) */
          double amp = p->edge_amp ? p->edge_amp[j] : 0.0;
          double ph = p->edge_phase ? p->edge_phase[j] : 0.0;
          REAL pack = (REAL)(amp * sin(p->omega * T + ph));
          dst[0][o] = pack;
          dst[1][o] = pack * (pack + (REAL)1);
          dst[2][o] = (REAL)0;
          continue;
        }
        ir = 2; jr = j;
      }
      size_t orr = (size_t)ir + (size_t)jr * NX;
      for (int f = 0; f < 3; ++f) dst[f][o] = src[f][orr];
    }
  }
}

/* ---------------------------------------------------------------------------
 * getWavesSpeedsCFL_{x,y}: apps/na/fv.h:7-193 and :195-381.
 * axis 0 = x (interface (i-1/2, j), skipped at i == 0),
 * axis 1 = y (interface (i, j-1/2), skipped at j == 0).
 * The y sweep starts the cfl at 0 (fv.h:245); the x sweep starts from the
 * value the y sweep left (fv.h:57); the skipped row/column gets cfl = 0.
 * ------------------------------------------------------------------------- */
static void SFX(waves)(SFX(ctx_t) *c, const orc_params *p, int axis,
                       REAL *const u[3], REAL dt) {
  const int NX = c->nx, NY = c->ny;
  const REAL dtd = dt / (REAL)(axis == 0 ? p->dx : p->dy);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if ((axis == 0 && i == 0) || (axis == 1 && j == 0)) {
        c->cfl[o] = 0;
        continue;
      }
      size_t om = (axis == 0) ? o - 1 : o - (size_t)NX;
      REAL cflMax = (axis == 0) ? c->cfl[o] : (REAL)0;
      const REAL *un = u[axis == 0 ? 1 : 2], *ut = u[axis == 0 ? 2 : 1];
      SFX(iface_t) f;
      if (SFX(riemann)(u[0][om], un[om], ut[om], u[0][o], un[o], ut[o],
                       (REAL)p->beta, p->form, &f)) {
        err |= 1;  /* fv.h:102: *simError = 2; return (nothing written) */
        continue;
      }
      for (int k = 0; k < 3; ++k) c->s[k][o] = f.s[k];
      for (int k = 0; k < 9; ++k) c->W[k][o] = f.W[k];
      for (int k = 0; k < 3; ++k) { c->amdq[k][o] = f.amdq[k]; c->apdq[k][o] = f.apdq[k]; }
      /* cflMax = getmax(cflMax, |L|*dtx) for every nonzero L */
      for (int k = 0; k < 3; ++k) {
        if (f.s[k] < 0) cflMax = SFX(getmax)(cflMax, -f.s[k] * dtd);
        else if (f.s[k] > 0) cflMax = SFX(getmax)(cflMax, f.s[k] * dtd);
      }
      c->cfl[o] = cflMax;
    }
  }
  if (err) c->simError = 2;
}

/* step_{x,y}: fv.h:383-440, :442-499.  dst <- src - dtd*(apdq(i) + amdq(i+1))
   in (i,j) interior, copy elsewhere; the y sweep maps the normal component
   back to u3 and the tangential one to u2 (fv.h:489-493). */
static void SFX(step1)(SFX(ctx_t) *c, const orc_params *p, int axis,
                       REAL *const src[3], REAL *const dst[3], REAL dt) {
  const int NX = c->nx, NY = c->ny;
  const REAL dtd = dt / (REAL)(axis == 0 ? p->dx : p->dy);
  const int fn = axis == 0 ? 1 : 2, ft = axis == 0 ? 2 : 1; /* u-index of comps 2,3 */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if (i >= 2 && i <= NX - 3 && j >= 2 && j <= NY - 3) {
        size_t op = (axis == 0) ? o + 1 : o + (size_t)NX;
        dst[0][o] = src[0][o] - dtd * (c->apdq[0][o] + c->amdq[0][op]);
        dst[fn][o] = src[fn][o] - dtd * (c->apdq[1][o] + c->amdq[1][op]);
        dst[ft][o] = src[ft][o] - dtd * (c->apdq[2][o] + c->amdq[2][op]);
      } else {
        for (int f = 0; f < 3; ++f) dst[f][o] = src[f][o];
      }
    }
  }
}

/* calcLimiters_{x,y}: fv.h:501-720, :722-930.  fadd is written into amdq
   for 1 <= i <= NX-2, 1 <= j <= NY-2 (the reference reuses the amdq grids).
   All threads read W/s before any writes amdq in the reference (different
   grids), so a single pass is faithful. */
static void SFX(limiters)(SFX(ctx_t) *c, const orc_params *p, int axis, REAL dt) {
  const int NX = c->nx, NY = c->ny;
  const REAL dtd = dt / (REAL)(axis == 0 ? p->dx : p->dy);
  const REAL Num1 = 1, Num0p5 = 0.5;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= NY - 2; ++j) {
    for (int i = 1; i <= NX - 2; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      size_t om = (axis == 0) ? o - 1 : o - (size_t)NX;
      size_t op = (axis == 0) ? o + 1 : o + (size_t)NX;
      REAL wl[3];
      for (int w = 0; w < 3; ++w) {          /* wave w: W1w, W2w, W3w */
        const REAL *W1 = c->W[3 * w + 0], *W2 = c->W[3 * w + 1], *W3 = c->W[3 * w + 2];
        /* note the reference's W storage: W11,W21,W31 are wave 1 */
        REAL wnorm2 = W1[o] * W1[o] + W2[o] * W2[o] + W3[o] * W3[o];
        REAL dotr = W1[o] * W1[op] + W2[o] * W2[op] + W3[o] * W3[op];
        REAL dotl = W1[o] * W1[om] + W2[o] * W2[om] + W3[o] * W3[om];
        wl[w] = SFX(philim)(wnorm2, dotl, dotr, c->s[w][o]);
      }
      REAL cq[3];
      for (int m = 0; m < 3; ++m) {          /* component m, fv.h:666-694 */
        REAL acc = 0;
        for (int w = 0; w < 3; ++w) {
          REAL as = fabs(c->s[w][o]);
          REAL term = as * (Num1 - as * dtd) * c->W[3 * w + m][o] * wl[w];
          acc = (w == 0) ? term : acc + term;
        }
        cq[m] = acc;
      }
      for (int m = 0; m < 3; ++m) c->amdq[m][o] = Num0p5 * cq[m]; /* fv.h:705-711 */
    }
  }
}

/* writeLimiters_{x,y}: fv.h:932-992, :994-1054 */
static void SFX(write_limiters)(SFX(ctx_t) *c, const orc_params *p, int axis,
                                REAL *const src[3], REAL *const dst[3], REAL dt) {
  const int NX = c->nx, NY = c->ny;
  const REAL dtd = dt / (REAL)(axis == 0 ? p->dx : p->dy);
  const int fn = axis == 0 ? 1 : 2, ft = axis == 0 ? 2 : 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if (i >= 2 && i <= NX - 3 && j >= 2 && j <= NY - 3) {
        size_t op = (axis == 0) ? o + 1 : o + (size_t)NX;
        dst[0][o] = src[0][o] - dtd * (c->amdq[0][op] - c->amdq[0][o]);
        dst[fn][o] = src[fn][o] - dtd * (c->amdq[1][op] - c->amdq[1][o]);
        dst[ft][o] = src[ft][o] - dtd * (c->amdq[2][op] - c->amdq[2][o]);
      } else {
        for (int f = 0; f < 3; ++f) dst[f][o] = src[f][o];
      }
    }
  }
}

/* reduceCFL / reduceCFL2 schedule, common/main.cu:179-212 and :503-524.
   Within one launch every thread reads cfl[i + S] (>= S) and writes
   cfl[i] (< S), so the sequential loop is the same fold. */
static REAL SFX(reduce_cfl)(SFX(ctx_t) *c, int mode) {
  const long N = (long)c->n;
  REAL *cfl = c->cfl;
  if (mode == ORC_CFL_ALL) {
    REAL m = 0;
    for (long k = 0; k < N; ++k) m = SFX(getmax)(m, cfl[k]);
    return m;
  }
  long sizered = N;
  while (sizered > 32) {
    sizered = sizered / 2;
    long S = (sizered / 16) * 16;            /* <<<sizered/16, 16>>> */
    for (long i = 0; i < S; ++i)
      if (i <= N - 1) cfl[i] = SFX(getmax)(cfl[i], cfl[i + S]);
  }
  while (sizered > 2) {
    sizered = sizered / 2;
    long S = sizered;                        /* <<<1, sizered>>> */
    for (long i = 0; i < S; ++i)
      if (i <= N - 1) cfl[i] = SFX(getmax)(cfl[i], cfl[i + S]);
  }
  return cfl[0];
}

/* source: apps/na/source.h:8-117 (momentum diffusion, interior only) */
static void SFX(source)(const orc_params *p, REAL *const src[3],
                        REAL *const dst[3], REAL dt) {
  const int NX = p->nx, NY = p->ny;
  const REAL Num1 = 1, Num2 = 2;
  const REAL dx = (REAL)p->dx, dy = (REAL)p->dy, D = (REAL)p->delta;
  const REAL ddtx2 = D * dt / (dx * dx);
  const REAL ddty2 = D * dt / (dy * dy);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if (i >= 2 && i <= NX - 3 && j >= 2 && j <= NY - 3) {
        size_t oW = o - 1, oE = o + 1, oS = o - (size_t)NX, oN = o + (size_t)NX;
        const REAL *q1 = src[0];
        REAL r0 = q1[o] + Num1, rW = q1[oW] + Num1, rE = q1[oE] + Num1;
        REAL rS = q1[oS] + Num1, rN = q1[oN] + Num1;
        if (p->form == ORC_FORM_RHO1) dst[0][o] = q1[o];
        else dst[0][o] = r0 - Num1;          /* source.h:84 */
        for (int f = 1; f <= 2; ++f) {       /* source.h:86-104 */
          const REAL *q = src[f];
          dst[f][o] = q[o]
              + ddtx2 * ((q[oW] / rW) - Num2 * (q[o] / r0) + (q[oE] / rE))
              + ddty2 * ((q[oS] / rS) - Num2 * (q[o] / r0) + (q[oN] / rN));
        }
      } else {
        for (int f = 0; f < 3; ++f) dst[f][o] = src[f][o];
      }
    }
  }
}

/* One directional sweep on c->A (ghosts already set):
   waves -> step1 (B) -> limiters -> write_limiters (A), as main.cu:353-418
   (y) and :430-495 (x). */
static void SFX(sweep)(SFX(ctx_t) *c, const orc_params *p, int axis, REAL dt) {
  SFX(waves)(c, p, axis, c->A, dt);
  SFX(step1)(c, p, axis, c->A, c->B, dt);
  SFX(limiters)(c, p, axis, dt);
  SFX(write_limiters)(c, p, axis, c->B, c->A, dt);
}

/* u1 of the state as tex2D reads it: coordinates clamped to the grid */
static double SFX(tex_u1)(const SFX(ctx_t) *c, int i, int j) {
  i = i < 0 ? 0 : (i > c->nx - 1 ? c->nx - 1 : i);
  j = j < 0 ? 0 : (j > c->ny - 1 ? c->ny - 1 : j);
  return (double)c->C[0][(size_t)i + (size_t)j * c->nx];
}

/* The cut of apps/taylor-angle-xy/data_collect.h:41-96, every thread (i, j)
   of the reference's grid visited in turn (at most one j per column
   passes the test on dj).  ja and the weights are float as declared there; T,
   dx, dy, the constants and the closed form are float when the reference is
   built with PRECISION=1 (collect_f32), else double.  The interpolation
   keeps the state's own precision. */
static void SFX(collect_taylor)(SFX(ctx_t) *c, const orc_params *p, const orc_state *st,
                                double *m) {
  const int NX = c->nx, NY = c->ny, L = p->cut_half;
  const int f32 = p->collect_f32;
  int icenter, jcenter;
  if (f32) {                                           /* data_collect.h:43-46 */
    const float T = (float)st->T, dx = (float)p->dx, dy = (float)p->dy;
    icenter = (int)(T * (float)p->md_vx / dx) - st->mdx;
    jcenter = NY / 2 + (int)((T - (float)p->md_t) * (float)p->md_vy / dy) - st->mdy;
  } else {
    icenter = (int)(st->T * p->md_vx / p->dx) - st->mdx;
    jcenter = NY / 2 + (int)((st->T - p->md_t) * p->md_vy / p->dy) - st->mdy;
  }
  const double cs = cos(p->tay_theta), sn = sin(p->tay_theta);
  for (int i = 0; i < NX; ++i) {
    if (abs(i - icenter) > L) continue;                /* :48-50 */
    float ja;                                          /* :52-55 */
    if (f32)
      ja = fmaf((float)(i - icenter) * ((float)p->dx / (float)p->dy), (float)p->tan_theta,
                (float)jcenter);
    else
      ja = (float)fma((double)(float)(i - icenter) * (p->dx / p->dy), p->tan_theta,
                      (double)(float)jcenter);
    for (int j = 0; j < NY; ++j) {
      const float dj = (float)j - ja;
      double val;
      if (dj >= 0.0f && dj <= 0.5f) {                  /* :58-68 */
        const float djm = (float)(j - 1) - ja;
        val = -(double)djm * SFX(tex_u1)(c, i, j) + (double)dj * SFX(tex_u1)(c, i, j - 1);
      } else if (dj < 0.0f && -dj < 0.5f) {            /* :69-79 */
        const float djp = (float)(j + 1) - ja;
        val = -(double)dj * SFX(tex_u1)(c, i, j + 1) + (double)djp * SFX(tex_u1)(c, i, j);
      } else {
        continue;
      }
      const int iw = i - (icenter - L);                /* :87-89 */
      double an;                                       /* :90-95 */
      if (f32) {
        const float T = (float)st->T, dx = (float)p->dx, dy = (float)p->dy;
        const float arg = (float)(i + st->mdx) * dx * (float)cs +
                          ((float)(j + st->mdy - NY / 2) * dy + (float)p->md_t * (float)sn) *
                              (float)sn - T;
        an = (double)((float)p->tay_rhoa * tanhf((float)p->tay_c * arg));
      } else {
        const double arg = (double)(i + st->mdx) * p->dx * cs +
                           ((double)(j + st->mdy - NY / 2) * p->dy + p->md_t * sn) * sn - st->T;
        an = p->tay_rhoa * tanh(p->tay_c * arg);
      }
      m[(size_t)iw + 2 * (size_t)NX] = val;
      m[(size_t)iw + 3 * (size_t)NX] = an;
    }
  }
}

/* dataCollect, apps/hifu-beam/data_collect.h:7-107 (apps/taylor-angle-xy/
   data_collect.h:7-98 for a Taylor instance), on the accepted state in c->C at
   the updated T and n, before the window shift (main.cu:563-573).  Within one
   launch the threads touch disjoint entries, so the column loop below is the
   same computation.  Returns 0 or ORC_ERR_COLLECT. */
static int SFX(collect)(SFX(ctx_t) *c, const orc_params *p, const orc_state *st) {
  double *m = p->measure;
  const int NX = c->nx, NY = c->ny;
  if (st->T <= 0) {                                    /* data_collect.h:27-30 */
    for (size_t k = 0; k < c->n; ++k) m[k] = 0.0;
    return 0;
  }
  const int64_t nf = st->n_accepted - 1 - p->n0frame;
  if (nf >= NX) return ORC_ERR_COLLECT;               /* data_collect.h:96 */
  if (p->bc == ORC_BC_TAYLOR) {                        /* taylor data_collect.h:34-40 */
    m[nf] = p->collect_f32 ? (double)(float)st->T : st->T;
    m[nf + NX] = st->last_cfl;
    SFX(collect_taylor)(c, p, st, m);
    return 0;
  }
  const REAL *row = c->C[0] + (size_t)p->jcut * NX;
  for (int i = 0; i < NX; ++i) {
    const double u1 = (double)row[i];
    for (int k = 0; k < 4; ++k) {                      /* data_collect.h:51-76 */
      if (i + st->mdx == p->mic[k]) {
        double *mr = m + (size_t)(2 + k) * NX;
        if (nf == 0)
          for (int ii = 0; ii < NX; ++ii) mr[ii] = 0.0;
        mr[nf] = u1;
        break;
      }
    }
    if (i == 0) {                                      /* data_collect.h:78-94 */
      m[nf] = st->T;
      m[nf + NX] = st->last_cfl;
      if (nf == 0)
        for (int k = 0; k < 4; ++k)
          if (p->mic[k] - st->mdx < 0 || p->mic[k] - st->mdx >= NX)
            for (int ii = 0; ii < NX; ++ii) m[(size_t)(2 + k) * NX + ii] = 0.0;
    }
    if (fabs((double)(i + st->mdx) / p->eta - st->T) > 0.5) continue;   /* :100-103 */
    const int64_t o = (int64_t)i + st->mdx + 6 * (int64_t)NX;
    if (o < (int64_t)NX * NY) m[o] = u1 > m[o] ? u1 : m[o];
  }
  return 0;
}

/* moveDomain + updateMDXY, common/main.cu:109-176, on the accepted state in
   c->C at the updated T (main.cu:575-586): C(i,j) <- C(i+MDi, j+MDj) when
   that cell is on the grid (the reference checks only the upper bounds;
   md_v >= 0 keeps MDi, MDj >= 0), else C(i,j) unchanged. */
static void SFX(move_domain)(SFX(ctx_t) *c, const orc_params *p, orc_state *st) {
  if (!(st->T >= p->md_t)) return;
  const int NX = c->nx, NY = c->ny;
  const int mdi = (int)((st->T - p->md_t) * p->md_vx / p->dx) - st->mdx;
  const int mdj = (int)((st->T - p->md_t) * p->md_vy / p->dy) - st->mdy;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; ++j) {
    for (int i = 0; i < NX; ++i) {
      const int ir = i + mdi, jr = j + mdj;
      const size_t o = (size_t)i + (size_t)j * NX;
      const size_t os = (ir <= NX - 1 && jr <= NY - 1) ? (size_t)ir + (size_t)jr * NX : o;
      for (int f = 0; f < 3; ++f) c->A[f][o] = c->C[f][os];
    }
  }
  for (int f = 0; f < 3; ++f) memcpy(c->C[f], c->A[f], c->n * sizeof(REAL));
  st->mdx += mdi;
  st->mdy += mdj;
}

/* One step attempt on the state in c->C (the backup buffer holds the input,
   main.cu:334).  Returns 0 accepted, 1 rejected, 2 eigenvalue error, 4 data
   collection overflow. */
static int SFX(attempt)(SFX(ctx_t) *c, const orc_params *p, orc_state *st,
                        double dt_used, int prescribed) {
  REAL dt = (REAL)dt_used;
  c->simError = 0;
  SFX(boundary)(p, c->C, c->A, st->T, st->mdx, st->mdy);  /* main.cu:338 */
  SFX(sweep)(c, p, 1, dt);                                /* main.cu:353-418 */
  if (c->simError == 2) return 2;                         /* main.cu:361-369 */
  SFX(sweep)(c, p, 0, dt);                                /* main.cu:430-495 */
  if (c->simError == 2) return 2;                         /* main.cu:438-446 */
  REAL cflStep = SFX(reduce_cfl)(c, p->cfl_mode);         /* main.cu:503-524 */
  st->last_cfl = (double)cflStep;
  int accepted = prescribed || (double)cflStep < p->cfl_max; /* main.cu:207-211 */
  st->n_attempts++;
  if (accepted) {
    SFX(source)(p, c->A, c->B, dt);                       /* main.cu:535 */
    for (int f = 0; f < 3; ++f) memcpy(c->C[f], c->B[f], c->n * sizeof(REAL));
    st->T += (double)dt;                                  /* main.cu:232-238 */
    st->n_accepted++;
    if (p->measure && SFX(collect)(c, p, st)) return 4;   /* main.cu:563-573 */
    if (p->md_on) SFX(move_domain)(c, p, st);             /* main.cu:575-586 */
  } else {
    st->n_rejected++;                                     /* clean retry */
  }
  if (!prescribed) {                                      /* main.cu:214-230 */
    double new_dt = (double)((REAL)st->dt * (REAL)p->cfl_wish / cflStep);
    if (!accepted && new_dt > st->dt * 0.9) st->dt = (double)(REAL)(st->dt * 0.9);
    else st->dt = new_dt;
  }
  return accepted ? 0 : 1;
}

static int SFX(run_impl)(const orc_params *p, REAL *q1, REAL *q2, REAL *q3,
                         int64_t n_accept, orc_state *st) {
  if (!p || !q1 || !q2 || !q3 || !st || p->nx < 8 || p->ny < 8) return ORC_ERR_ARG;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  SFX(ctx_t) c;
  if (SFX(ctx_alloc)(&c, p->nx, p->ny)) { SFX(ctx_free)(&c); return ORC_ERR_ARG; }
  memcpy(c.C[0], q1, c.n * sizeof(REAL));
  memcpy(c.C[1], q2, c.n * sizeof(REAL));
  memcpy(c.C[2], q3, c.n * sizeof(REAL));
  int rc = 0, consecutive_rej = 0;
  int64_t target = st->n_accepted + n_accept;
  while (st->n_accepted < target && !(p->stop_at_T && !(st->T < p->T_stop))) {
    int prescribed = p->dt_seq != NULL;
    double dt_used = st->dt;
    if (prescribed) {
      if (st->n_attempts >= p->dt_seq_len) { rc = ORC_ERR_ARG; break; }
      dt_used = p->dt_seq[st->n_attempts];
      st->dt = dt_used;
    }
    int r = SFX(attempt)(&c, p, st, dt_used, prescribed);
    if (r == 2) { rc = ORC_ERR_EIGEN; st->error = rc; break; }
    if (r == 4) { rc = ORC_ERR_COLLECT; st->error = rc; break; }
    consecutive_rej = (r == 1) ? consecutive_rej + 1 : 0;
    if (consecutive_rej > 200) { rc = ORC_ERR_DT; st->error = rc; break; }
  }
  memcpy(q1, c.C[0], c.n * sizeof(REAL));
  memcpy(q2, c.C[1], c.n * sizeof(REAL));
  memcpy(q3, c.C[2], c.n * sizeof(REAL));
  SFX(ctx_free)(&c);
  return rc;
}
