/*
 * Segment-length autotuner.  The x and y sweeps split each tridiagonal
 * system into segments of 4, 8 or 16 unknowns per lane (so a 256-point
 * system uses 64, 32 or 16 lanes and that many partitions in its reduced
 * system).  Shorter segments mean fewer registers and more lanes (higher
 * occupancy, deeper reduced-system PCR); longer ones mean less reduced-system
 * work per unknown.  The choice is made per (precision, numX, numY) class:
 * from the cache file named by futhark_context_config_set_cache_file when it
 * has an entry, else by timing the candidates on a small synthetic batch when
 * "lvc.autotune" is set, else from the built-in table below.
 *
 * Determinism: a size class always maps to the same choice within a context,
 * and a batch runs each problem with its own class's choice (one launch
 * pair per distinct choice and time step), so a problem's arithmetic never
 * depends on the rest of its batch.
 *
 * Cache file format: one line per class, "precision numX numY seg_x seg_y".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/locvolcalib.h"
#include "lvc_runtime.h"

static int table_find(struct futhark_context *ctx, int precision, int nx,
                      int ny) {
  for (int i = 0; i < ctx->ntuned; i++)
    if (ctx->tuned[i].precision == precision && ctx->tuned[i].nx == nx &&
        ctx->tuned[i].ny == ny)
      return i;
  return -1;
}

static void table_put(struct futhark_context *ctx, int precision, int nx,
                      int ny, int sx, int sy) {
  int i = table_find(ctx, precision, nx, ny);
  if (i < 0) {
    if (ctx->ntuned == LVC_MAX_TUNED)
      return;
    i = ctx->ntuned++;
  }
  lvc_tuned t = {precision, nx, ny, sx, sy};
  ctx->tuned[i] = t;
}

void lvc_tune_load(struct futhark_context *ctx) {
  if (!ctx->cfg || !ctx->cfg->cache_file)
    return;
  FILE *f = fopen(ctx->cfg->cache_file, "r");
  if (!f)
    return;
  int p, nx, ny, sx, sy;
  while (fscanf(f, "%d %d %d %d %d", &p, &nx, &ny, &sx, &sy) == 5)
    if (p >= LVC_FP32 && p <= LVC_FP64 && lvc_seg_supported(sx, 0) &&
        lvc_seg_supported(sy, 1))
      table_put(ctx, p, nx, ny, sx, sy);
  fclose(f);
}

static void cache_append(struct futhark_context *ctx, int precision, int nx,
                         int ny, int sx, int sy) {
  if (!ctx->cfg || !ctx->cfg->cache_file)
    return;
  FILE *f = fopen(ctx->cfg->cache_file, "a");
  if (!f)
    return;
  fprintf(f, "%d %d %d %d %d\n", precision, nx, ny, sx, sy);
  fclose(f);
}

/* Built-in choice: x 8 / y 32 (FP64 x 4 / y 32: 4 unknowns per lane keep
 * its registers in bounds), fitted to the size.  Measured on MI355X for
 * every class numX, numY in {32, 64, 128, 256, 512} in each precision, on a
 * batch of strike variants (shared x tiles, as `main`) and a batch of
 * distinct problems, 47 steps, 16-2048 problems (128 * 256^2 / (numX
 * numY)): it was the fastest pair, or within 3% of it, in 68 of the 75
 * classes.  The other seven (numX = 32 with lane-segment y solves of 4, up
 * to 26% faster there, and FP64 512 x 64 with y 64) won only on such large
 * batches of one class: on main(16, 32, 256, 256) x 8 / y 4 took 7.27 ms
 * against 4.87, and the 512-problem uneven batch 19.6 ms against 10.5
 * (split into three segment classes).  So no class departs from it; the
 * autotuner ("lvc.autotune") times a class on a small batch of its own. */
static void builtin(int precision, int nx, int ny, int *sx, int *sy) {
  (void)nx;
  (void)ny;
  *sx = precision == LVC_FP64 ? 4 : 8;
  *sy = 32;
}

/* Times one candidate: a batch of 8 problems of the class, 4 time steps:
 * four strike variants of one problem (which share their x tiles, as
 * `main`'s options) and four problems of their own. */
static double time_candidate(struct futhark_context *ctx, int precision, int nx,
                             int ny, int sx, int sy) {
  lvc_problem probs[8];
  memset(probs, 0, sizeof(probs));
  for (int i = 0; i < 8; i++) {
    probs[i].nx = nx;
    probs[i].ny = ny;
    probs[i].nt = 5;
    probs[i].s0 = i < 4 ? 0.03 : 0.03 + 1e-4 * i;
    probs[i].strike = 0.001 * i;
    probs[i].t = 5.0;
    probs[i].alpha = 0.2;
    probs[i].nu = 0.6;
    probs[i].beta = 0.5;
  }
  lvc_batch b;
  if (lvc_batch_build(ctx, &b, 8, probs, precision, sx, sy))
    return -1;
  hipEvent_t e0 = NULL, e1 = NULL;
  double best = -1;
  if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    for (int rep = 0; rep < 4; rep++) {
      (void)hipEventRecord(e0, ctx->stream);
      if (lvc_batch_run(ctx, &b))
        break;
      (void)hipEventRecord(e1, ctx->stream);
      if (hipEventSynchronize(e1) != hipSuccess)
        break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0 && (best < 0 || ms < best))
        best = ms;
    }
  }
  if (e0)
    (void)hipEventDestroy(e0);
  if (e1)
    (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  lvc_batch_release(ctx, &b);
  return best;
}

void lvc_tune_peek(struct futhark_context *ctx, int precision, int nx, int ny,
                   int *seg_x, int *seg_y) {
  int i = table_find(ctx, precision, nx, ny);
  if (i >= 0) {
    *seg_x = ctx->tuned[i].seg_x;
    *seg_y = ctx->tuned[i].seg_y;
    return;
  }
  builtin(precision, nx, ny, seg_x, seg_y);
}

void lvc_tune_pick(struct futhark_context *ctx, int precision, int nx, int ny,
                   int *seg_x, int *seg_y) {
  int i = table_find(ctx, precision, nx, ny);
  if (i >= 0) {
    *seg_x = ctx->tuned[i].seg_x;
    *seg_y = ctx->tuned[i].seg_y;
    return;
  }
  builtin(precision, nx, ny, seg_x, seg_y);
  if (!ctx->cfg || !ctx->cfg->autotune || !ctx->ok)
    return;
  static const int segs[] = {4, 8, 16, 32, 64};
  double best = -1;
  int bx = *seg_x, by = *seg_y;
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 5; c++) {
      if (lvc_lanes_for(nx, segs[a]) > 64 || !lvc_yseg_fits(ny, segs[c]))
        continue;
      double ms = time_candidate(ctx, precision, nx, ny, segs[a], segs[c]);
      if (ms > 0 && (best < 0 || ms < best)) {
        best = ms;
        bx = segs[a];
        by = segs[c];
      }
    }
  char *err = futhark_context_get_error(ctx); /* a failed trial is not fatal */
  free(err);
  *seg_x = bx;
  *seg_y = by;
  table_put(ctx, precision, nx, ny, bx, by);
  cache_append(ctx, precision, nx, ny, bx, by);
}

/* ---- the program's tuning parameters ------------------------------------ */

/* LocVolCalib.fut's two thresholds (the generated API lists them; the
 * kernels here need no threshold, so they are accepted and kept), then
 * this library's own knobs. */
static const char *const tuning_names[] = {
    "main.suff_intra_par_0", "main.suff_intra_par_1", "lvc.precision",
    "lvc.seg_x", "lvc.seg_y", "lvc.autotune", "lvc.graph", "lvc.xshare"};
static const char *const tuning_classes[] = {
    "threshold", "threshold", "precision", "segment", "segment", "flag", "flag",
    "count"};

static int set_param(struct futhark_context_config *cfg, const char *name, size_t v) {
  if (strcmp(name, "main.suff_intra_par_0") == 0) {
    cfg->suff_intra_par_0 = v;
    return 0;
  }
  if (strcmp(name, "main.suff_intra_par_1") == 0) {
    cfg->suff_intra_par_1 = v;
    return 0;
  }
  if (strcmp(name, "lvc.precision") == 0) {
    if (v > LVC_FP64)
      return 1;
    cfg->precision = (int)v;
    return 0;
  }
  if (strcmp(name, "lvc.seg_x") == 0 || strcmp(name, "lvc.seg_y") == 0) {
    if (v != 0 && !lvc_seg_supported((int)v, name[8] == 'y'))
      return 1;
    if (name[8] == 'x')
      cfg->seg_x = (int)v;
    else
      cfg->seg_y = (int)v;
    return 0;
  }
  if (strcmp(name, "lvc.xshare") == 0) {
    if (v > LVC_XSHARE_MAX)
      return 1;
    cfg->xshare = (int)v;
    return 0;
  }
  if (strcmp(name, "lvc.graph") == 0) {
    cfg->graph = v != 0;
    return 0;
  }
  if (strcmp(name, "lvc.autotune") == 0) {
    cfg->autotune = v != 0;
    return 0;
  }
  return 1;
}

const lvc_program lvc_prog = {(int)(sizeof(tuning_names) / sizeof(tuning_names[0])),
                              tuning_names, tuning_classes, set_param,
                              lvc_context_ready};
