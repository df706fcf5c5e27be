/* apps/hifu-beam/driver.c -- the hifu-beam app as a C host program over the
 * library's C ABI: what apps/hifu-beam/driver.cu of the reference assembles
 * from the app's headers and common/main.cu.  The app's parameters.h and
 * init.h set the instance up; its boundary (apps/hifu-beam/boundary.h) and
 * data collection are compiled into libfivonagi.so; the frame loop is
 * calcFrame's (common/main.cu:310: steps until T >= Tprint, then the frame's
 * output), with the moving window and the data collection on.
 *
 *   driver ETAC ISPAM FRAMES [OUT [NX NY]]
 *
 * Prints one line per frame (frame, T, accepted and rejected steps, MDX, the
 * largest u1max so far); OUT (optional) receives the final state
 * (3 x NX*NY floats) and the measurement grid (NX*NY floats).
 *
 *   gcc -std=c99 -O2 -I include fivonagi_amd/apps/hifu-beam/driver.c \
 *       -L fivonagi_amd -lfivonagi -Wl,-rpath,fivonagi_amd -lm -o driver
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../../include/fivonagi.h"
#include "init.h"
#include "parameters.h"

static int check(int rc, const char *what) {
  if (rc != 0) fprintf(stderr, "%s: error %d: %s\n", what, rc, fivonagi_last_error());
  return rc;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s ETAC ISPAM FRAMES [OUT [NX NY]]\n", argv[0]);
    return 2;
  }
  const int etac = atoi(argv[1]);
  const double ispam = atof(argv[2]);
  const int frames = atoi(argv[3]);
  const char *out = argc > 4 ? argv[4] : NULL;
  const int nx = argc > 6 ? atoi(argv[5]) : 0, ny = argc > 6 ? atoi(argv[6]) : 0;
  const double dtprint = 1.0;                        /* DTPRINT */

  fivonagi_params p;
  fivonagi_hifu_parameters(&p, etac, ispam, 0.9, nx, ny, 1);
  const size_t n = (size_t)p.nx * (size_t)p.ny;
  float *q = malloc(sizeof(float) * 4 * n);
  if (!q) return 3;
  fivonagi_hifu_init(&p, q, q + n, q + 2 * n);       /* init() */

  const float *q0[3] = {q, q + n, q + 2 * n};
  fivonagi_ctx *ctx = NULL;
  if (check(fivonagi_init(&ctx, &p, 1, q0), "fivonagi_init")) return 1;
  fivonagi_collect_opts co;
  fivonagi_hifu_collect_options(&p, &co);
  int rc = check(fivonagi_set_collect(ctx, 0, &co), "fivonagi_set_collect");

  float *m = q + 3 * n;
  double Tprint = p.T0;
  for (int f = 0; f < frames && rc == 0; ++f) {      /* calcFrame, one frame each */
    Tprint += dtprint;
    rc = check(fivonagi_step_until(ctx, Tprint), "fivonagi_step_until");
    if (rc) break;
    fivonagi_info_t info;
    rc = check(fivonagi_info(ctx, 0, &info), "fivonagi_info");
    if (rc) break;
    rc = check(fivonagi_measure(ctx, 0, m), "fivonagi_measure");
    if (rc) break;
    float peak = 0.0f;                               /* u1max, from row 6 on */
    for (size_t k = 6 * (size_t)p.nx; k < n; ++k) peak = m[k] > peak ? m[k] : peak;
    printf("frame %d T %.6f accepted %lld rejected %lld MDX %d u1max %.6e\n", f, info.T,
           (long long)info.n_accepted, (long long)info.n_rejected, info.md_x, peak);
  }
  if (rc == 0) rc = check(fivonagi_collect(ctx, 0, q, q + n, q + 2 * n, NULL, NULL, NULL, NULL),
                          "fivonagi_collect");
  if (rc == 0 && out) {
    FILE *fp = fopen(out, "wb");
    if (!fp || fwrite(q, sizeof(float), 4 * n, fp) != 4 * n) rc = 4;
    if (fp) fclose(fp);
  }
  fivonagi_destroy(ctx);
  free(q);
  return rc;
}
