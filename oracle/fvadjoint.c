/*
 * fvadjoint.c -- FP64 discrete adjoint of the FiVoNAGI step (TEST
 * INFRASTRUCTURE: the checker of the HIP adjoint; never linked into the
 * product).
 *
 * The reference has no adjoint.  This is the reverse-mode derivative of the
 * step the oracle's forward computes (fvoracle_impl.h, which restates
 * common/main.cu:310-635 with apps/na/fv.h and apps/na/source.h), for the
 * executed branch path and a frozen dt sequence (the dt of each accepted step
 * is an input, not a function of the state).  Objective:
 *     J = sum_n dt_n * sum_{box} rho'_{n+1}^2
 * Parameters: the left-edge drive amplitude amp_j and phase phase_j
 * (p_j = amp_j sin(omega T + phase_j), apps/hifu-beam/boundary.h:77-93
 * pattern) and the initial state.
 *
 * Structure: the step is boundary B, y sweep Y, x sweep X, source S.  Each
 * sweep is a set of independent 1-D lines (columns for Y, rows for X); the
 * line forward below is the same arithmetic as the oracle's kernels (waves ->
 * first-order step -> limiter -> second-order step), written per line with
 * the fluctuation and correction flux folded into P = apdq - fadd and
 * M = amdq + fadd, which is algebraically identical.  The adjoint runs S^T,
 * X^T, Y^T, B^T with the forward recomputed from the stored input state of
 * each step.  Its correctness is pinned by central finite differences of the
 * oracle forward (tests/test_adjoint.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fvoracle.h"

typedef struct {
  double a1, w21, w31, a2, w22, w32, a3;  /* waves (density, normal, tangential) */
  double L1, L2, L3;                      /* speeds */
  /* intermediates kept for the reverse pass */
  double rhoL, rhoR, uL, uR, vL, vR, Z1, Z2, Z3, tmp1, sq, A1, A2, E1, E2, D, dr;
} rw_t;

/* apps/na/fv.h:88-139, generic in the normal direction */
static int riemann_fw(double rL, double nL, double tL, double rR, double nR, double tR,
                      double beta, rw_t *w) {
  w->rhoL = 1.0 + rL;
  w->rhoR = 1.0 + rR;
  w->uL = nL / w->rhoL; w->uR = nR / w->rhoR;
  w->vL = tL / w->rhoL; w->vR = tR / w->rhoR;
  w->Z1 = (w->rhoR + w->rhoL) / 2.0;
  w->Z2 = (w->uR + w->uL) / 2.0;
  w->Z3 = (w->vR + w->vL) / 2.0;
  w->tmp1 = w->Z2 / w->Z1;
  double tmp3 = w->tmp1 * w->tmp1 - 2.0 * w->Z2 * w->Z2 / w->Z1 + 1.0
              + 2.0 * (beta - 1.0) * (w->Z1 - 1.0);
  if (tmp3 < 0) return ORC_ERR_EIGEN;
  w->sq = sqrt(tmp3);
  w->L1 = w->tmp1 + w->sq;
  w->L2 = w->tmp1 - w->sq;
  w->L3 = w->tmp1;
  double N1 = w->L1 * w->Z3 - 2.0 * w->Z2 * w->Z3, N2 = w->L2 * w->Z3 - 2.0 * w->Z2 * w->Z3;
  w->E1 = w->L1 * w->Z1 - w->Z3;
  w->E2 = w->L2 * w->Z1 - w->Z3;
  w->A1 = N1 / w->E1;
  w->A2 = N2 / w->E2;
  w->dr = rR - rL;
  double dn = nR - nL, dt = tR - tL;
  w->D = w->L1 - w->L2;
  w->a1 = (dn - w->dr * w->L2) / w->D;
  w->a2 = w->dr - w->a1;
  w->a3 = dt - w->a1 * w->A1 - w->a2 * w->A2;   /* = fv.h:119-125 after substitution */
  w->w21 = w->a1 * w->L1; w->w31 = w->a1 * w->A1;
  w->w22 = w->a2 * w->L2; w->w32 = w->a2 * w->A2;
  return 0;
}

/* reverse of riemann_fw: gW = adjoints of (a1,w21,w31,a2,w22,w32,a3), gs of
   (L1,L2,L3); accumulates into gL, gR (rho', normal, tangential). */
static void riemann_bw(const rw_t *w, double beta, const double gW[7], const double gs[3],
                       double gL[3], double gR[3]) {
  double ga3 = gW[6];
  double ga1 = gW[0] + gW[1] * w->L1 + gW[2] * w->A1;
  double ga2 = gW[3] + gW[4] * w->L2 + gW[5] * w->A2;
  double gL1 = gW[1] * w->a1 + gs[0];
  double gL2 = gW[4] * w->a2 + gs[1];
  double gA1 = gW[2] * w->a1, gA2 = gW[5] * w->a2;
  double gtmp1 = gs[2];
  double gdt = ga3;
  ga1 -= ga3 * w->A1; gA1 -= ga3 * w->a1;
  ga2 -= ga3 * w->A2; gA2 -= ga3 * w->a2;
  double gdr = ga2;
  ga1 -= ga2;
  double gdn = ga1 / w->D;
  gdr -= ga1 * w->L2 / w->D;
  gL2 -= ga1 * w->dr / w->D;
  double gD = -ga1 * w->a1 / w->D;
  gL1 += gD; gL2 -= gD;
  double gN1 = gA1 / w->E1, gE1 = -gA1 * w->A1 / w->E1;
  double gN2 = gA2 / w->E2, gE2 = -gA2 * w->A2 / w->E2;
  gL1 += gN1 * w->Z3 + gE1 * w->Z1;
  gL2 += gN2 * w->Z3 + gE2 * w->Z1;
  double gZ3 = gN1 * (w->L1 - 2.0 * w->Z2) + gN2 * (w->L2 - 2.0 * w->Z2) - gE1 - gE2;
  double gZ2 = -2.0 * w->Z3 * (gN1 + gN2);
  double gZ1 = gE1 * w->L1 + gE2 * w->L2;
  gtmp1 += gL1 + gL2;
  double gsq = gL1 - gL2;
  double gtmp3 = gsq / (2.0 * w->sq);
  gtmp1 += 2.0 * w->tmp1 * gtmp3;
  gZ2 += -4.0 * w->Z2 / w->Z1 * gtmp3;
  gZ1 += (2.0 * w->Z2 * w->Z2 / (w->Z1 * w->Z1) + 2.0 * (beta - 1.0)) * gtmp3;
  gZ2 += gtmp1 / w->Z1;
  gZ1 += -gtmp1 * w->tmp1 / w->Z1;
  double grhoR = gZ1 / 2.0, grhoL = gZ1 / 2.0;
  double guR = gZ2 / 2.0, guL = gZ2 / 2.0, gvR = gZ3 / 2.0, gvL = gZ3 / 2.0;
  grhoR += -(guR * w->uR + gvR * w->vR) / w->rhoR;
  grhoL += -(guL * w->uL + gvL * w->vL) / w->rhoL;
  gR[0] += gdr + grhoR;
  gL[0] += -gdr + grhoL;
  gR[1] += gdn + guR / w->rhoR;
  gL[1] += -gdn + guL / w->rhoL;
  gR[2] += gdt + gvR / w->rhoR;
  gL[2] += -gdt + gvL / w->rhoL;
}

static void wave_vec(const rw_t *w, int p, double v[3]) {
  if (p == 0) { v[0] = w->a1; v[1] = w->w21; v[2] = w->w31; }
  else if (p == 1) { v[0] = w->a2; v[1] = w->w22; v[2] = w->w32; }
  else { v[0] = 0.0; v[1] = 0.0; v[2] = w->a3; }
}

static double speed(const rw_t *w, int p) { return p == 0 ? w->L1 : (p == 1 ? w->L2 : w->L3); }

typedef struct {
  double nrm[3], r[3], phi[3], dphidr[3], al[3], be[3];
  int nb[3];   /* -1: dot with interface m-1, +1: with m+1 */
} lim_t;

/* limiter + folded flux at interface m (fv.h:579-711) */
static void limiter_fw(const rw_t *wm, const rw_t *w, const rw_t *wp, double dtd, lim_t *l,
                       double P[3], double M[3]) {
  P[0] = P[1] = P[2] = M[0] = M[1] = M[2] = 0.0;
  for (int p = 0; p < 3; ++p) {
    double v[3], vn[3];
    double s = speed(w, p);
    wave_vec(w, p, v);
    l->nb[p] = (s > 0) ? -1 : +1;
    wave_vec(l->nb[p] < 0 ? wm : wp, p, vn);
    double nrm = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double dot = v[0] * vn[0] + v[1] * vn[1] + v[2] * vn[2];
    l->nrm[p] = nrm;
    double phi, dphidr = 0.0, r = 0.0;
    if (nrm != 0.0) {
      r = dot / nrm;
      double c = (1.0 + r) / 2.0;
      double wv = (c < 2.0) ? c : 2.0;
      dphidr = (c < 2.0) ? 0.5 : 0.0;
      double r2 = 2.0 * r;
      if (!(wv < r2)) { wv = r2; dphidr = 2.0; }
      if (!(wv > 0.0)) { wv = 0.0; dphidr = 0.0; }
      phi = wv;
    } else {
      phi = 1.0;
    }
    l->r[p] = r;
    l->phi[p] = phi;
    l->dphidr[p] = dphidr;
    double as = fabs(s);
    double cc = 0.5 * as * (1.0 - as * dtd) * phi;
    l->al[p] = (s > 0 ? s : 0.0) - cc;
    l->be[p] = (s < 0 ? s : 0.0) + cc;
    for (int k = 0; k < 3; ++k) { P[k] += l->al[p] * v[k]; M[k] += l->be[p] * v[k]; }
  }
}

/* reverse of limiter_fw: from gP, gM accumulate gW (own interface, 7),
   gs (3) and the dot adjoints gdot[p] (distributed by the caller). */
static void limiter_bw(const rw_t *w, const lim_t *l, double dtd, const double gP[3],
                       const double gM[3], double gW[7], double gs[3], double gdot[3]) {
  for (int p = 0; p < 3; ++p) {
    double v[3];
    wave_vec(w, p, v);
    double s = speed(w, p), as = fabs(s);
    double gv[3];
    for (int k = 0; k < 3; ++k) gv[k] = l->al[p] * gP[k] + l->be[p] * gM[k];
    double gal = gP[0] * v[0] + gP[1] * v[1] + gP[2] * v[2];
    double gbe = gM[0] * v[0] + gM[1] * v[1] + gM[2] * v[2];
    double gss = (s > 0 ? gal : 0.0) + (s < 0 ? gbe : 0.0);
    double gc = -gal + gbe;
    double gphi = gc * 0.5 * as * (1.0 - as * dtd);
    double gas = gc * 0.5 * l->phi[p] * (1.0 - 2.0 * as * dtd);
    gss += gas * (s > 0 ? 1.0 : (s < 0 ? -1.0 : 0.0));
    gs[p] += gss;
    gdot[p] = 0.0;
    if (l->nrm[p] != 0.0) {
      double gr = gphi * l->dphidr[p];
      gdot[p] = gr / l->nrm[p];
      double gnrm = -gr * l->r[p] / l->nrm[p];
      for (int k = 0; k < 3; ++k) gv[k] += 2.0 * gnrm * v[k];
    }
    if (p == 0) { gW[0] += gv[0]; gW[1] += gv[1]; gW[2] += gv[2]; }
    else if (p == 1) { gW[3] += gv[0]; gW[4] += gv[1]; gW[5] += gv[2]; }
    else { gW[6] += gv[2]; }
  }
}

static void add_wave_scaled(double gW[7], const rw_t *w, int p, double s) {
  if (p == 0) { gW[0] += s * w->a1; gW[1] += s * w->w21; gW[2] += s * w->w31; }
  else if (p == 1) { gW[3] += s * w->a2; gW[4] += s * w->w22; gW[5] += s * w->w32; }
  else { gW[6] += s * w->a3; }
}

/* ---- one line: cells 0..L-1 along the sweep, arrays of (rho', normal, tangential) ---- */
typedef struct {
  int L;
  rw_t *w;       /* interfaces 1..L-1 */
  lim_t *lim;    /* interfaces 2..L-2 */
  double (*P)[3], (*M)[3];
  double (*gW)[7], (*gs)[3], (*gdot)[3];
  double (*gP)[3], (*gM)[3];
} line_ws;

static int line_alloc(line_ws *ws, int L) {
  ws->L = L;
  ws->w = calloc((size_t)L, sizeof(rw_t));
  ws->lim = calloc((size_t)L, sizeof(lim_t));
  ws->P = calloc((size_t)L, sizeof(*ws->P));
  ws->M = calloc((size_t)L, sizeof(*ws->M));
  ws->gW = calloc((size_t)L, sizeof(*ws->gW));
  ws->gs = calloc((size_t)L, sizeof(*ws->gs));
  ws->gdot = calloc((size_t)L, sizeof(*ws->gdot));
  ws->gP = calloc((size_t)L, sizeof(*ws->gP));
  ws->gM = calloc((size_t)L, sizeof(*ws->gM));
  return (ws->w && ws->lim && ws->P && ws->M && ws->gW && ws->gs && ws->gdot && ws->gP && ws->gM)
             ? 0 : -1;
}

static void line_free(line_ws *ws) {
  free(ws->w); free(ws->lim); free(ws->P); free(ws->M); free(ws->gW); free(ws->gs);
  free(ws->gdot); free(ws->gP); free(ws->gM);
}

/* forward of one line: q[k][3] -> out[k][3]; interior[k] marks updated cells */
static int line_fw(line_ws *ws, double (*q)[3], double (*out)[3], const unsigned char *interior,
                   double beta, double dtd) {
  int L = ws->L, err = 0;
  for (int m = 1; m < L; ++m)
    err |= riemann_fw(q[m - 1][0], q[m - 1][1], q[m - 1][2], q[m][0], q[m][1], q[m][2], beta,
                      &ws->w[m]);
  if (err) return ORC_ERR_EIGEN;
  for (int m = 2; m <= L - 2; ++m)
    limiter_fw(&ws->w[m - 1], &ws->w[m], &ws->w[m + 1], dtd, &ws->lim[m], ws->P[m], ws->M[m]);
  for (int k = 0; k < L; ++k)
    for (int c = 0; c < 3; ++c)
      out[k][c] = interior[k] ? q[k][c] - dtd * (ws->P[k][c] + ws->M[k + 1][c]) : q[k][c];
  return 0;
}

/* reverse of line_fw (forward state must be the one line_fw just used):
   lam_out[k] -> lam_in[k] (overwritten) */
static void line_bw(line_ws *ws, double (*lam_out)[3], double (*lam_in)[3],
                    const unsigned char *interior, double beta, double dtd) {
  int L = ws->L;
  memset(ws->gW, 0, sizeof(*ws->gW) * L);
  memset(ws->gs, 0, sizeof(*ws->gs) * L);
  memset(ws->gdot, 0, sizeof(*ws->gdot) * L);
  memset(ws->gP, 0, sizeof(*ws->gP) * L);
  memset(ws->gM, 0, sizeof(*ws->gM) * L);
  for (int k = 0; k < L; ++k) {
    for (int c = 0; c < 3; ++c) lam_in[k][c] = lam_out[k][c];
    if (interior[k])
      for (int c = 0; c < 3; ++c) {
        ws->gP[k][c] += -dtd * lam_out[k][c];
        ws->gM[k + 1][c] += -dtd * lam_out[k][c];
      }
  }
  for (int m = 2; m <= L - 2; ++m)
    limiter_bw(&ws->w[m], &ws->lim[m], dtd, ws->gP[m], ws->gM[m], ws->gW[m], ws->gs[m],
               ws->gdot[m]);
  /* dot_p(m) = W_p(m) . W_p(m + nb) */
  for (int m = 2; m <= L - 2; ++m)
    for (int p = 0; p < 3; ++p) {
      double g = ws->gdot[m][p];
      if (g == 0.0) continue;
      int nbm = m + ws->lim[m].nb[p];
      add_wave_scaled(ws->gW[m], &ws->w[nbm], p, g);
      add_wave_scaled(ws->gW[nbm], &ws->w[m], p, g);
    }
  for (int m = 1; m < L; ++m)
    riemann_bw(&ws->w[m], beta, ws->gW[m], ws->gs[m], lam_in[m - 1], lam_in[m]);
}

/* ---- 2-D step pieces on SoA grids ---- */

typedef struct {
  int nx, ny;
  size_t n;
  double *b[3], *y[3], *xy[3];     /* q_b, q_y, q_xy of the current step */
} step_ws;

static int ws_alloc(step_ws *s, int nx, int ny) {
  memset(s, 0, sizeof(*s));
  s->nx = nx; s->ny = ny; s->n = (size_t)nx * ny;
  for (int f = 0; f < 3; ++f) {
    s->b[f] = calloc(s->n, sizeof(double));
    s->y[f] = calloc(s->n, sizeof(double));
    s->xy[f] = calloc(s->n, sizeof(double));
    if (!s->b[f] || !s->y[f] || !s->xy[f]) return -1;
  }
  return 0;
}

static void ws_free(step_ws *s) {
  for (int f = 0; f < 3; ++f) { free(s->b[f]); free(s->y[f]); free(s->xy[f]); }
}

/* ghost source of cell (i,j) for the copy rules (common/boundary.h:27-81) */
static void ghost_map(const orc_params *p, int i, int j, int *si, int *sj, int *drive) {
  int NX = p->nx, NY = p->ny, ir = i, jr = j;
  *drive = 0;
  if (i == NX - 1 || i == NX - 2) { ir = NX - 3; jr = j; }
  if (j == 0 || j == 1) { ir = i; jr = 2; }
  if (j == NY - 1 || j == NY - 2) { ir = i; jr = NY - 3; }
  if (i == 0 || i == 1) {
    if (p->bc == ORC_BC_EDGE_DRIVE) *drive = 1;
    ir = 2; jr = j;
  }
  *si = ir; *sj = jr;
}

static int is_ghost(const orc_params *p, int i, int j) {
  return i <= 1 || j <= 1 || i >= p->nx - 2 || j >= p->ny - 2;
}

static void boundary_fw(const orc_params *p, double *const q[3], double *const b[3], double T) {
  int NX = p->nx, NY = p->ny;
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if (p->bc == ORC_BC_TAYLOR) {
        if (is_ghost(p, i, j)) {
          double th = p->tay_theta;
          double r = p->tay_rhoa * tanh(p->tay_c * ((double)i * p->dx * cos(th) +
                                        ((double)(j - p->tay_jc) * p->dy + p->tay_yoff) * sin(th) - T));
          b[0][o] = r; b[1][o] = r * cos(th); b[2][o] = r * sin(th);
        } else {
          for (int f = 0; f < 3; ++f) b[f][o] = q[f][o];
        }
        continue;
      }
      int si, sj, drive;
      ghost_map(p, i, j, &si, &sj, &drive);
      if (drive) {
        double amp = p->edge_amp ? p->edge_amp[j] : 0.0;
        double ph = p->edge_phase ? p->edge_phase[j] : 0.0;
        double pk = amp * sin(p->omega * T + ph);
        b[0][o] = pk; b[1][o] = pk * (pk + 1.0); b[2][o] = 0.0;
      } else {
        size_t os = (size_t)si + (size_t)sj * NX;
        for (int f = 0; f < 3; ++f) b[f][o] = q[f][os];
      }
    }
}

static void boundary_bw(const orc_params *p, double *const lb[3], double *const lq[3], double T,
                        double *g_amp, double *g_phase) {
  int NX = p->nx, NY = p->ny;
  size_t n = (size_t)NX * NY;
  for (int f = 0; f < 3; ++f) memset(lq[f], 0, n * sizeof(double));
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      if (p->bc == ORC_BC_TAYLOR) {
        if (!is_ghost(p, i, j))
          for (int f = 0; f < 3; ++f) lq[f][o] += lb[f][o];
        continue;
      }
      int si, sj, drive;
      ghost_map(p, i, j, &si, &sj, &drive);
      if (drive) {
        double amp = p->edge_amp ? p->edge_amp[j] : 0.0;
        double ph = p->edge_phase ? p->edge_phase[j] : 0.0;
        double sn = sin(p->omega * T + ph), cs = cos(p->omega * T + ph);
        double pk = amp * sn;
        double gp = lb[0][o] + lb[1][o] * (2.0 * pk + 1.0);
        if (g_amp) g_amp[j] += gp * sn;
        if (g_phase) g_phase[j] += gp * amp * cs;
      } else {
        size_t os = (size_t)si + (size_t)sj * NX;
        for (int f = 0; f < 3; ++f) lq[f][os] += lb[f][o];
      }
    }
}

/* the y sweep on columns (normal = q3, tangential = q2) or the x sweep on rows
   (normal = q2, tangential = q3); fw: src -> dst; bw: lam_out -> lam_in.
   The lines are independent: each thread works through its own lines with
   its own line workspace, so the result does not depend on the threads. */
static int sweep_lines(step_ws *s, const orc_params *p, int axis, double *const src[3],
                       double *const dst[3], double dtd, int backward, double *const lam_out[3],
                       double *const lam_in[3]) {
  int NX = s->nx, NY = s->ny;
  int nlines = axis == 1 ? NX : NY, L = axis == 1 ? NY : NX;
  int fn = axis == 1 ? 2 : 1, ft = axis == 1 ? 1 : 2;
  int err = 0;
#pragma omp parallel reduction(| : err)
  {
    line_ws ws;
    double(*lq)[3] = calloc((size_t)L, sizeof(*lq)), (*lo)[3] = calloc((size_t)L, sizeof(*lo));
    double(*la)[3] = calloc((size_t)L, sizeof(*la)), (*lb)[3] = calloc((size_t)L, sizeof(*lb));
    unsigned char *in = calloc((size_t)L, 1);
    int ok = !line_alloc(&ws, L) && lq && lo && la && lb && in;
    if (!ok) err |= 2;
#pragma omp for schedule(static)
    for (int line = 0; line < nlines; ++line) {
      if (!ok) continue;
      for (int k = 0; k < L; ++k) {
        int i = axis == 1 ? line : k, j = axis == 1 ? k : line;
        size_t o = (size_t)i + (size_t)j * NX;
        lq[k][0] = src[0][o]; lq[k][1] = src[fn][o]; lq[k][2] = src[ft][o];
        in[k] = (i >= 2 && i <= NX - 3 && j >= 2 && j <= NY - 3);
      }
      if (line_fw(&ws, lq, lo, in, p->beta, dtd)) { err |= 1; continue; }
      if (!backward) {
        for (int k = 0; k < L; ++k) {
          int i = axis == 1 ? line : k, j = axis == 1 ? k : line;
          size_t o = (size_t)i + (size_t)j * NX;
          dst[0][o] = lo[k][0]; dst[fn][o] = lo[k][1]; dst[ft][o] = lo[k][2];
        }
      } else {
        for (int k = 0; k < L; ++k) {
          int i = axis == 1 ? line : k, j = axis == 1 ? k : line;
          size_t o = (size_t)i + (size_t)j * NX;
          la[k][0] = lam_out[0][o]; la[k][1] = lam_out[fn][o]; la[k][2] = lam_out[ft][o];
        }
        line_bw(&ws, la, lb, in, p->beta, dtd);
        for (int k = 0; k < L; ++k) {
          int i = axis == 1 ? line : k, j = axis == 1 ? k : line;
          size_t o = (size_t)i + (size_t)j * NX;
          lam_in[0][o] = lb[k][0]; lam_in[fn][o] = lb[k][1]; lam_in[ft][o] = lb[k][2];
        }
      }
    }
    line_free(&ws);
    free(lq); free(lo); free(la); free(lb); free(in);
  }
  if (err & 2) return ORC_ERR_ARG;
  return err ? ORC_ERR_EIGEN : 0;
}

/* apps/na/source.h:84-104 (q1 passes through) */
static void source_fw(const orc_params *p, double *const q[3], double *const out[3], double dt) {
  int NX = p->nx, NY = p->ny;
  double ddx2 = p->delta * dt / (p->dx * p->dx), ddy2 = p->delta * dt / (p->dy * p->dy);
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) {
      size_t o = (size_t)i + (size_t)j * NX;
      out[0][o] = q[0][o];
      if (i >= 2 && i <= NX - 3 && j >= 2 && j <= NY - 3) {
        for (int f = 1; f <= 2; ++f) {
          double u0 = q[f][o] / (1.0 + q[0][o]);
          double uW = q[f][o - 1] / (1.0 + q[0][o - 1]), uE = q[f][o + 1] / (1.0 + q[0][o + 1]);
          double uS = q[f][o - NX] / (1.0 + q[0][o - NX]), uN = q[f][o + NX] / (1.0 + q[0][o + NX]);
          out[f][o] = q[f][o] + ddx2 * (uW - 2.0 * u0 + uE) + ddy2 * (uS - 2.0 * u0 + uN);
        }
      } else {
        out[1][o] = q[1][o];
        out[2][o] = q[2][o];
      }
    }
}

static void source_bw(const orc_params *p, double *const q[3], double *const lo[3],
                      double *const li[3], double dt, double *gu) {
  int NX = p->nx, NY = p->ny;
  size_t n = (size_t)NX * NY;
  double ddx2 = p->delta * dt / (p->dx * p->dx), ddy2 = p->delta * dt / (p->dy * p->dy);
  for (int f = 0; f < 3; ++f) memcpy(li[f], lo[f], n * sizeof(double));
  for (int f = 1; f <= 2; ++f) {
    memset(gu, 0, n * sizeof(double));
    for (int j = 2; j <= NY - 3; ++j)
      for (int i = 2; i <= NX - 3; ++i) {
        size_t o = (size_t)i + (size_t)j * NX;
        double g = lo[f][o];
        gu[o - 1] += ddx2 * g; gu[o + 1] += ddx2 * g;
        gu[o - NX] += ddy2 * g; gu[o + NX] += ddy2 * g;
        gu[o] += -2.0 * (ddx2 + ddy2) * g;
      }
    for (size_t o = 0; o < n; ++o) {
      double rho = 1.0 + q[0][o];
      li[f][o] += gu[o] / rho;
      li[0][o] += -gu[o] * (q[f][o] / rho) / rho;
    }
  }
}

static int step_fw(step_ws *s, const orc_params *p, double *const q[3], double *const out[3],
                   double T, double dt) {
  boundary_fw(p, q, s->b, T);
  if (sweep_lines(s, p, 1, s->b, s->y, dt / p->dy, 0, NULL, NULL)) return ORC_ERR_EIGEN;
  if (sweep_lines(s, p, 0, s->y, s->xy, dt / p->dx, 0, NULL, NULL)) return ORC_ERR_EIGEN;
  source_fw(p, s->xy, out, dt);
  return 0;
}

static double box_sum(const orc_box *bx, int nx, const double *q1) {
  double acc = 0.0;
  for (int j = bx->y0; j < bx->y0 + bx->ny; ++j)
    for (int i = bx->x0; i < bx->x0 + bx->nx; ++i) {
      double v = q1[(size_t)i + (size_t)j * nx];
      acc += v * v;
    }
  return acc;
}

static int check(const orc_params *p, int64_t nsteps, const orc_box *bx) {
  if (!p || p->nx < 8 || p->ny < 8 || nsteps < 0 || !bx) return ORC_ERR_ARG;
  if (!p->dt_seq || p->dt_seq_len < nsteps) return ORC_ERR_ARG;
  if (bx->x0 < 0 || bx->y0 < 0 || bx->nx < 0 || bx->ny < 0 || bx->x0 + bx->nx > p->nx ||
      bx->y0 + bx->ny > p->ny)
    return ORC_ERR_ARG;
  return 0;
}

int orc_line_step_f64(const orc_params *p, double *q1, double *q2, double *q3, double T,
                      double dt) {
  if (!p || p->nx < 8 || p->ny < 8) return ORC_ERR_ARG;
  step_ws s;
  if (ws_alloc(&s, p->nx, p->ny)) { ws_free(&s); return ORC_ERR_ARG; }
  double *q[3] = {q1, q2, q3};
  double *o[3];
  for (int f = 0; f < 3; ++f) o[f] = malloc(s.n * sizeof(double));
  int rc = step_fw(&s, p, q, o, T, dt);
  for (int f = 0; f < 3; ++f) { memcpy(q[f], o[f], s.n * sizeof(double)); free(o[f]); }
  ws_free(&s);
  return rc;
}

/* The reverse sweep over a stored trajectory: states holds q_0 .. q_nsteps
   (3 grids each), Ts the start time of each step. */
static int adjoint_reverse(step_ws *s, const orc_params *p, double *states, const double *Ts,
                           int64_t nsteps, const orc_box *bx, double *J, double *grad_amp,
                           double *grad_phase, double *grad_q0) {
  size_t n = s->n;
  double *lam = calloc(3 * n, sizeof(double)), *lam2 = calloc(3 * n, sizeof(double));
  double *lx = calloc(3 * n, sizeof(double)), *gu = calloc(n, sizeof(double));
  int rc = 0;
  if (!lam || !lam2 || !lx || !gu) { rc = ORC_ERR_ARG; goto out; }
  *J = 0.0;
  for (int64_t k = 0; k < nsteps; ++k)
    *J += p->dt_seq[k] * box_sum(bx, p->nx, states + (size_t)(3 * (k + 1)) * n);
  if (grad_amp) memset(grad_amp, 0, sizeof(double) * p->ny);
  if (grad_phase) memset(grad_phase, 0, sizeof(double) * p->ny);
  for (int64_t k = nsteps - 1; k >= 0; --k) {
    double dt = p->dt_seq[k];
    double *qi[3], *qo1 = states + (size_t)(3 * (k + 1)) * n;
    for (int f = 0; f < 3; ++f) qi[f] = states + (size_t)(3 * k + f) * n;
    double *L[3] = {lam, lam + n, lam + 2 * n}, *L2[3] = {lam2, lam2 + n, lam2 + 2 * n};
    double *LX[3] = {lx, lx + n, lx + 2 * n};
    /* dJ/dq_{k+1} */
    for (int j = bx->y0; j < bx->y0 + bx->ny; ++j)
      for (int i = bx->x0; i < bx->x0 + bx->nx; ++i) {
        size_t o = (size_t)i + (size_t)j * p->nx;
        L[0][o] += 2.0 * dt * qo1[o];
      }
    /* recompute the step's intermediates */
    boundary_fw(p, qi, s->b, Ts[k]);
    sweep_lines(s, p, 1, s->b, s->y, dt / p->dy, 0, NULL, NULL);
    sweep_lines(s, p, 0, s->y, s->xy, dt / p->dx, 0, NULL, NULL);
    /* S^T, X^T, Y^T, B^T */
    source_bw(p, s->xy, L, L2, dt, gu);                           /* lam2 = S^T lam   */
    sweep_lines(s, p, 0, s->y, NULL, dt / p->dx, 1, L2, LX);      /* lx   = X^T lam2  */
    sweep_lines(s, p, 1, s->b, NULL, dt / p->dy, 1, LX, L2);      /* lam2 = Y^T lx    */
    boundary_bw(p, L2, L, Ts[k], grad_amp, grad_phase);           /* lam  = B^T lam2  */
  }
  if (grad_q0) memcpy(grad_q0, lam, sizeof(double) * 3 * n);
out:
  free(lam); free(lam2); free(lx); free(gu);
  return rc;
}

int orc_adjoint_f64(const orc_params *p, const double *q0, int64_t nsteps, const orc_box *bx,
                    double *J, double *grad_amp, double *grad_phase, double *grad_q0) {
  if (check(p, nsteps, bx) || !q0 || !J) return ORC_ERR_ARG;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  step_ws s;
  if (ws_alloc(&s, p->nx, p->ny)) { ws_free(&s); return ORC_ERR_ARG; }
  size_t n = s.n;
  /* every state of the run (small test grids only) */
  double *states = malloc(sizeof(double) * 3 * n * (size_t)(nsteps + 1));
  double *Ts = malloc(sizeof(double) * (size_t)(nsteps + 1));
  int rc = 0;
  if (!states || !Ts) { rc = ORC_ERR_ARG; goto out; }
  memcpy(states, q0, sizeof(double) * 3 * n);
  Ts[0] = p->T0;
  for (int64_t k = 0; k < nsteps; ++k) {
    double *qi[3], *qo[3];
    for (int f = 0; f < 3; ++f) {
      qi[f] = states + (size_t)(3 * k + f) * n;
      qo[f] = states + (size_t)(3 * (k + 1) + f) * n;
    }
    rc = step_fw(&s, p, qi, qo, Ts[k], p->dt_seq[k]);
    if (rc) goto out;
    Ts[k + 1] = Ts[k] + p->dt_seq[k];
  }
  rc = adjoint_reverse(&s, p, states, Ts, nsteps, bx, J, grad_amp, grad_phase, grad_q0);
out:
  free(states); free(Ts);
  ws_free(&s);
  return rc;
}

/* The same adjoint linearised along a given trajectory q_0 .. q_nsteps
   (states: nsteps + 1 states of 3 grids, e.g. another implementation's run
   with the same dt sequence): the branches it differentiates are those of
   that trajectory's steps, recomputed here in FP64. */
int orc_adjoint_path_f64(const orc_params *p, const double *states, int64_t nsteps,
                         const orc_box *bx, double *J, double *grad_amp, double *grad_phase,
                         double *grad_q0) {
  if (check(p, nsteps, bx) || !states || !J) return ORC_ERR_ARG;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  step_ws s;
  if (ws_alloc(&s, p->nx, p->ny)) { ws_free(&s); return ORC_ERR_ARG; }
  size_t n = s.n;
  double *st = malloc(sizeof(double) * 3 * n * (size_t)(nsteps + 1));
  double *Ts = malloc(sizeof(double) * (size_t)(nsteps + 1));
  int rc = 0;
  if (!st || !Ts) { rc = ORC_ERR_ARG; goto out; }
  memcpy(st, states, sizeof(double) * 3 * n * (size_t)(nsteps + 1));
  Ts[0] = p->T0;
  for (int64_t k = 0; k < nsteps; ++k) Ts[k + 1] = Ts[k] + p->dt_seq[k];
  rc = adjoint_reverse(&s, p, st, Ts, nsteps, bx, J, grad_amp, grad_phase, grad_q0);
out:
  free(st); free(Ts);
  ws_free(&s);
  return rc;
}

/* J of the oracle's own forward (fvoracle_impl.h) with the prescribed dt
   sequence: the quantity the finite-difference checks differentiate. */
int orc_objective_f64(const orc_params *p, const double *q0, int64_t nsteps, const orc_box *bx,
                      double *J) {
  if (check(p, nsteps, bx) || !q0 || !J) return ORC_ERR_ARG;
  size_t n = (size_t)p->nx * p->ny;
  double *q = malloc(sizeof(double) * 3 * n);
  if (!q) return ORC_ERR_ARG;
  memcpy(q, q0, sizeof(double) * 3 * n);
  orc_state st;
  orc_state_init(p, &st);
  *J = 0.0;
  int rc = 0;
  for (int64_t k = 0; k < nsteps; ++k) {
    double dt = p->dt_seq[k];
    rc = orc_run_f64(p, q, q + n, q + 2 * n, 1, &st);
    if (rc) break;
    *J += dt * box_sum(bx, p->nx, q);
  }
  free(q);
  return rc;
}
