/* apps/hifu-beam/parameters.h -- the hifu-beam app's parameters
 * (apps/hifu-beam/parameters.h:12-104 of the reference) in their runtime
 * form, for a C host: NX = 512 ETAC, NY = 1024 ETAC, XMAX = 25, YMAX = 50
 * wavelengths, BETA = 4.8, DIFFDELTA = 2.974e-4, AMPL = 4.217e-4, FREC = 1,
 * aperture ISPa = ISPAM wavelengths, focal distance ISPF = 50 wavelengths (in
 * points), a quiescent start at T = -TIDE with
 * TIDE = floor(2 + 4/FREC + ISPa^2 dx / (2 ISPF dy^2)), and the moving window
 * MDVX = 1, MDVY = 0, MDT = XMAX/2.  And the data collection's options
 * (data_collect.h:45-50): JCUT = 2, ETA = NX/XMAX and the four microphones
 * round(ISPF 0.5f), round(ISPF 0.9f), ISPF, round(ISPF 1.1f), with ISPF the
 * float the FP32 build holds.  Host code (C99). */
#ifndef FIVONAGI_APPS_HIFU_BEAM_PARAMETERS_H
#define FIVONAGI_APPS_HIFU_BEAM_PARAMETERS_H

#include <math.h>
#include <string.h>

#include "../../../include/fivonagi.h"

/* nx, ny: 0 for the app's grid at ETAC (other values keep the physical size);
   cflw: CFLWHISH; moving: the window moves (0: fixed domain). */
static inline void fivonagi_hifu_parameters(fivonagi_params *p, int etac, double ispam,
                                            double cflw, int nx, int ny, int moving) {
  memset(p, 0, sizeof(*p));
  p->nx = nx > 0 ? nx : 512 * etac;
  p->ny = ny > 0 ? ny : 1024 * etac;
  const double xmax = 25.0, ymax = 50.0, frec = 1.0;
  p->dx = xmax / (p->nx - 1);
  p->dy = ymax / (p->ny - 1);
  const double lam = 1.0;                            /* ISPlambda */
  const double a = ispam * lam, f = 50.0 * lam / p->dx;
  const double tide = floor(2.0 + 4.0 / frec + (a * a * p->dx) / (2.0 * f * p->dy * p->dy));
  p->beta = 4.8;
  p->delta = 2.974e-4;
  p->cfl_max = 1.0;
  p->cfl_wish = cflw;
  p->dt0 = 0.005;
  p->T0 = -tide;
  p->bc = FIVONAGI_BC_HIFU;
  p->cfl_mode = FIVONAGI_CFL_REFERENCE_FOLD;
  p->hifu_ampl = 4.217e-4;
  p->hifu_frec = frec;
  p->hifu_a = a;
  p->hifu_f = f;
  p->md_on = moving ? 1 : 0;
  p->md_vx = 1.0;
  p->md_vy = 0.0;
  p->md_t = xmax * 0.5;
}

static inline int fivonagi_hifu_round(float v) { return (int)floor((double)v + 0.5); }

static inline void fivonagi_hifu_collect_options(const fivonagi_params *p,
                                                 fivonagi_collect_opts *o) {
  memset(o, 0, sizeof(*o));
  const float ispf = (float)(50.0 * 1.0 / (double)((float)25.0f / (float)(p->nx - 1)));
  o->jcut = 2;
  o->mic[0] = fivonagi_hifu_round(ispf * 0.5f);
  o->mic[1] = fivonagi_hifu_round(ispf * 0.9f);
  o->mic[2] = (int)ispf;
  o->mic[3] = fivonagi_hifu_round(ispf * 1.1f);
  o->eta = p->nx / 25.0;
}

#endif /* FIVONAGI_APPS_HIFU_BEAM_PARAMETERS_H */
