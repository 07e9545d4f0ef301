/*
 * Entry points of the LocVolCalib library and the batch engine behind them.
 *
 * A batch is a set of independent problems (LocVolCalib.fut:211-236,
 * `value`), each with its own sizes and parameters.  The reference's `main`
 * (:238-252) is the batch of `outer` problems sharing everything but the
 * strike 0.001*i.  Per time step the engine launches the x sweep and the y
 * sweep once for the whole batch; a problem whose numT-1 steps are done
 * drops out (its workgroups return at once) until the longest has finished.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/locvolcalib.h"
#include "lvc_runtime.h"

/* i32.f32 / i32.f64 (truncation toward zero) of a value in the i32 range;
 * anything else (NaN, infinities, out of range) gives INT32_MIN, which lies
 * outside every grid. */
static int32_t trunc_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0))
    return INT32_MIN;
  return (int32_t)v;
}

/* LocVolCalib.fut:15-20.  The f32 branch performs the reference's own f32
 * operations in its order (C on this host evaluates float expressions in
 * single precision, IEEE-rounded, without contraction: sqrtf is correctly
 * rounded), so myXindex and x[0] are the reference's bits. */
void lvc_grid_scalars(lvc_problem *q, int grid32) {
  q->grid32 = grid32;
  q->yindex = q->ny / 2; /* i32.i64 numY / 2 */
  q->dy = 10.0 * q->nu * sqrt(q->t) / (double)q->ny;
  if (grid32) {
    const float s0 = (float)q->s0, t = (float)q->t, alpha = (float)q->alpha;
    const float stdX = 20.0f * alpha * s0 * sqrtf(t);
    const float dx = stdX / (float)q->nx;
    q->xindex = trunc_i32(s0 / dx);
    /* myX[0] = f32.i64 0 * dx - f32.i32 myXindex * dx + s0 (:21) */
    const float x0 = 0.0f * dx - (float)q->xindex * dx + s0;
    q->dx = dx;
    q->x0log = x0;
  } else {
    const double stdX = 20.0 * q->alpha * q->s0 * sqrt(q->t);
    q->dx = stdX / (double)q->nx;
    q->xindex = trunc_i32(q->s0 / q->dx);
    q->x0log = 0.0 * q->dx - (double)q->xindex * q->dx + q->s0;
  }
  q->status = (q->xindex >= 0 && q->xindex < q->nx) ? 0 : 3;
}

int lvc_validate_sizes(struct futhark_context *ctx, int64_t nx, int64_t ny,
                       int64_t nt) {
  /* LocVolCalib.fut:37 indexes x[i-1], x[i+1] for i in 1...n-2, and the
   * timeline needs two points (:217). */
  if (nx < 3 || ny < 3) {
    lvc_set_error(ctx, "numX and numY must be at least 3 (got %lld, %lld)",
                  (long long)nx, (long long)ny);
    return 1;
  }
  if (nx > LVC_MAX_N || ny > LVC_MAX_N) {
    lvc_set_error(ctx, "numX and numY must be at most %d (got %lld, %lld)",
                  LVC_MAX_N, (long long)nx, (long long)ny);
    return 1;
  }
  if (nt < 2) {
    lvc_set_error(ctx, "numT must be at least 2 (got %lld)", (long long)nt);
    return 1;
  }
  return 0;
}

/* Smallest supported segment length >= want that fits a system of length n:
 * x sweep {4, 8, 16} (<= 64 lanes per system); y sweep also the column
 * segments {32, 64} (<= LVC_COL_BLOCK/64 segments per column), falling back
 * to 16 when no column segment fits. */
static int fit_seg(int n, int want, int ysweep) {
  static const int segs[] = {4, 8, 16, 32, 64};
  const int ns = ysweep ? 5 : 3;
  for (int i = 0; i < ns; i++)
    if (segs[i] >= want && (ysweep ? lvc_yseg_fits(n, segs[i])
                                   : lvc_lanes_for(n, segs[i]) <= 64))
      return segs[i];
  return 16;
}

void lvc_choose_segs(struct futhark_context *ctx, int precision, int nx, int ny,
                     int *seg_x, int *seg_y) {
  int sx = ctx->cfg ? ctx->cfg->seg_x : 0, sy = ctx->cfg ? ctx->cfg->seg_y : 0;
  int tx = 0, ty = 0;
  if (!sx || !sy)
    lvc_tune_pick(ctx, precision, nx, ny, &tx, &ty);
  *seg_x = fit_seg(nx, sx ? sx : tx, 0);
  *seg_y = fit_seg(ny, sy ? sy : ty, 1);
}

static int64_t align64(int64_t v) { return (v + 63) & ~(int64_t)63; }

/* Stable insertion sort of problem indices by numT, descending (runs in
 * O(n) for the common case of equal or already ordered numT). */
static void nt_sort_desc(int *order, int n, const lvc_problem *probs) {
  for (int i = 1; i < n; i++) {
    const int v = order[i];
    int j = i - 1;
    while (j >= 0 && probs[order[j]].nt < probs[v].nt) {
      order[j + 1] = order[j];
      j--;
    }
    order[j + 1] = v;
  }
}

/* Problems per x tile sharing a factorisation when lvc.xshare is set (0
 * when it is not: the batch then chooses, below). */
int lvc_xshare(struct futhark_context *ctx, int precision) {
  (void)precision;
  return ctx->cfg && ctx->cfg->xshare > 0 ? ctx->cfg->xshare : 0;
}

/* The batch's default: FP64 shares 16 (its row factorisation is the dearest
 * part; at 128 x 256x256x64 36.8K options/s against 23.0K unshared).  The
 * FP32 forms share as few as still let every x workgroup of a step (each
 * group's row tiles and its y-factor workgroup) be resident at once, three
 * per CU (the shared kernels' occupancy), but at least 2: more groups are
 * more waves to hide latency with (at 128 x 256x256x64, FP32+refinement:
 * 6 per tile 46.9K options/s, 8 44.0K, 5 40.4K -- a second round of
 * workgroups).  runs[r]: lengths of the runs of problems that could share;
 * tiles[r]: x tiles per group of run r. */
static int batch_xshare(struct futhark_context *ctx, int precision, int nruns,
                        const int *runs, const int *tiles) {
  if (precision == LVC_FP64)
    return 16;
  const int64_t slots = 3 * (int64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256);
  for (int x = 2; x < LVC_XSHARE_MAX; x++) {
    int64_t wgs = 0;
    for (int r = 0; r < nruns; r++)
      wgs += (int64_t)((runs[r] + x - 1) / x) * (tiles[r] + 1);
    if (wgs <= slots)
      return x;
  }
  return LVC_XSHARE_MAX;
}

/* Problems whose factorisations coincide: everything but the strike (and
 * the state) equal.  The reference's main is a batch of such problems
 * (LocVolCalib.fut:244-251). */
static int same_factor(const lvc_problem *a, const lvc_problem *b) {
  return a->nx == b->nx && a->ny == b->ny && a->nt == b->nt && a->s0 == b->s0 &&
         a->t == b->t && a->alpha == b->alpha && a->nu == b->nu && a->beta == b->beta &&
         a->grid32 == b->grid32;
}

int lvc_batch_build(struct futhark_context *ctx, lvc_batch *b, int n,
                    const lvc_problem *probs, int precision, int seg_x,
                    int seg_y) {
  memset(b, 0, sizeof(*b));
  b->n = n;
  b->precision = precision;
  b->seg_x = seg_x;
  b->seg_y = seg_y;
  if (n == 0)
    return 0;
  const size_t real = precision == LVC_FP64 ? sizeof(double) : sizeof(float);
  int xshare = lvc_xshare(ctx, precision);
  b->h_probs = malloc((size_t)n * sizeof(lvc_problem));
  int *order = malloc((size_t)n * sizeof(int));
  int *gsize = malloc((size_t)n * sizeof(int)); /* by position in order */
  int32_t *leaders = malloc((size_t)n * sizeof(int32_t));
  if (!b->h_probs || !order || !gsize || !leaders) {
    free(order);
    free(gsize);
    free(leaders);
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  int64_t geo = 0, state = 0, nxt = 0, nyt = 0, steps = 0;
  for (int p = 0; p < n; p++) {
    lvc_problem q = probs[p];
    int gy, cpl, cols;
    lvc_ytile_shape(q.ny, seg_y, (int)real, &gy, &cpl, &cols);
    const int lds = lvc_ytile_lds(q.ny, seg_y, (int)real);
    if (lds > b->ylds)
      b->ylds = lds;
    q.geo_off = geo;
    q.state_off = state;
    q.k = -1;
    lvc_grid_scalars(&q, q.grid32);
    q.ydone = 0;
    q.fac_src = p;
    q.ytiles = (q.nx + cols - 1) / cols;
    geo += align64(LVC_GEO_LEN(q.nx, q.ny));
    state += align64((int64_t)q.nx * q.ny);
    nyt += q.ytiles;
    if (q.nt - 1 > steps)
      steps = q.nt - 1;
    b->h_probs[p] = q;
  }
  /* Problems by numT, longest first (stable), so that at every step the
   * tiles of the problems still running form a prefix of each list; runs of
   * up to xshare consecutive problems with the same factorisation share
   * their x tiles (and the first one's y factor table). */
  for (int p = 0; p < n; p++)
    order[p] = p;
  nt_sort_desc(order, n, b->h_probs);
  if (xshare <= 0) {
    /* the runs of consecutive problems that could share, unbounded */
    int nruns = 0;
    for (int oi = 0; oi < n;) {
      const int p = order[oi];
      int cnt = 1;
      while (oi + cnt < n && order[oi + cnt] == p + cnt &&
             same_factor(&b->h_probs[p], &b->h_probs[p + cnt]))
        cnt++;
      const int rows = LVC_BLOCK / lvc_lanes_for(b->h_probs[p].nx, seg_x);
      gsize[nruns] = cnt;                                    /* run length */
      leaders[nruns] = (b->h_probs[p].ny + rows - 1) / rows; /* tiles per group */
      nruns++;
      oi += cnt;
    }
    xshare = batch_xshare(ctx, precision, nruns, gsize, leaders);
  }
  for (int oi = 0; oi < n;) {
    const int p = order[oi];
    int cnt = 1;
    while (cnt < xshare && oi + cnt < n && order[oi + cnt] == p + cnt &&
           same_factor(&b->h_probs[p], &b->h_probs[p + cnt]))
      cnt++;
    gsize[oi] = cnt;
    if (cnt > 1)
      b->xshared = 1;
    for (int i = 0; i < cnt; i++)
      b->h_probs[p + i].fac_src = p;
    const int rows = LVC_BLOCK / lvc_lanes_for(b->h_probs[p].nx, seg_x);
    nxt += (b->h_probs[p].ny + rows - 1) / rows;
    oi += cnt;
  }
  if (nxt + nyt > (int64_t)1 << 30) {
    free(order);
    free(gsize);
    free(leaders);
    lvc_set_error(ctx, "batch too large");
    return 1;
  }
  b->nxt = (int)nxt;
  b->nyt = (int)nyt;
  b->steps = steps;
  lvc_tile *tiles = malloc((size_t)(nxt + nyt) * sizeof(lvc_tile));
  b->x_active = malloc((size_t)(steps ? steps : 1) * sizeof(int));
  b->y_active = malloc((size_t)(steps ? steps : 1) * sizeof(int));
  b->f_active = malloc((size_t)(steps ? steps : 1) * sizeof(int));
  if (!tiles || !b->x_active || !b->y_active || !b->f_active) {
    free(tiles);
    free(order);
    free(gsize);
    free(leaders);
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  int64_t tx = 0, ty = nxt;
  int nlead = 0;
  for (int64_t s = 0; s < steps; s++)
    b->x_active[s] = b->y_active[s] = b->f_active[s] = 0;
  for (int oi = 0; oi < n; oi += gsize[oi]) {
    const int p = order[oi], cnt = gsize[oi];
    const lvc_problem *q = &b->h_probs[p];
    const int gx = lvc_lanes_for(q->nx, seg_x);
    const int rows = LVC_BLOCK / gx;
    int gy, cpl, cols;
    lvc_ytile_shape(q->ny, seg_y, (int)real, &gy, &cpl, &cols);
    for (int r0 = 0; r0 < q->ny; r0 += rows) {
      lvc_tile t = {p, r0, q->ny - r0 < rows ? q->ny - r0 : rows, gx, cnt, 0};
      tiles[tx++] = t;
    }
    for (int i = 0; i < cnt; i++)
      for (int c0 = 0; c0 < q->nx; c0 += cols) {
        lvc_tile t = {p + i, c0, q->nx - c0 < cols ? q->nx - c0 : cols, gy, cpl, 0};
        tiles[ty++] = t;
      }
    leaders[nlead++] = p;
    /* these problems run steps 0 .. nt-2: the prefixes reach past their tiles */
    for (int64_t s = 0; s < q->nt - 1; s++) {
      b->x_active[s] = (int)tx;
      b->y_active[s] = (int)(ty - nxt);
      b->f_active[s] = nlead;
    }
  }
  b->probs_bytes = (size_t)n * sizeof(lvc_problem);
  b->geo_bytes = (size_t)geo * sizeof(double);
  b->state_bytes = (size_t)state * real;
  b->tiles_bytes = (size_t)(nxt + nyt) * sizeof(lvc_tile);
  b->yfac_bytes = lvc_seg_is_col(seg_y) ? (size_t)n * LVC_YFAC_LEN * real : 0;
  int rc = lvc_dev_alloc(ctx, b->probs_bytes, (void **)&b->d_probs) ||
           lvc_dev_alloc(ctx, b->geo_bytes, (void **)&b->d_geo) ||
           lvc_dev_alloc(ctx, b->state_bytes, &b->d_R) ||
           lvc_dev_alloc(ctx, b->state_bytes, &b->d_Y) ||
           lvc_dev_alloc(ctx, b->tiles_bytes, (void **)&b->d_tiles) ||
           (b->yfac_bytes &&
            (lvc_dev_alloc(ctx, b->yfac_bytes, &b->d_yfac) ||
             lvc_dev_alloc(ctx, (size_t)n * sizeof(int32_t), (void **)&b->d_fprobs)));
  if (!rc)
    rc = lvc_upload(ctx, b->d_probs, b->h_probs, b->probs_bytes) ||
         lvc_upload(ctx, b->d_tiles, tiles, b->tiles_bytes) ||
         (b->d_fprobs && lvc_upload(ctx, b->d_fprobs, leaders, (size_t)nlead * sizeof(int32_t)));
  free(tiles);
  free(order);
  free(gsize);
  free(leaders);
  b->graph = ctx->cfg && ctx->cfg->graph;
  return rc;
}

void lvc_batch_release(struct futhark_context *ctx, lvc_batch *b) {
  lvc_dev_free(ctx, b->d_probs, b->probs_bytes);
  lvc_dev_free(ctx, b->d_geo, b->geo_bytes);
  lvc_dev_free(ctx, b->d_R, b->state_bytes);
  lvc_dev_free(ctx, b->d_Y, b->state_bytes);
  lvc_dev_free(ctx, b->d_tiles, b->tiles_bytes);
  if (b->d_yfac)
    lvc_dev_free(ctx, b->d_yfac, b->yfac_bytes);
  if (b->d_fprobs)
    lvc_dev_free(ctx, b->d_fprobs, (size_t)b->n * sizeof(int32_t));
  free(b->h_probs);
  free(b->x_active);
  free(b->y_active);
  free(b->f_active);
  memset(b, 0, sizeof(*b));
}

/* One time step of every batch in bs that still runs: the x sweeps, then
 * the y sweeps (each batch is a set of problems sharing one segment pair). */
static int launch_step(struct futhark_context *ctx, lvc_batch *bs, int nb, int64_t s) {
  void *st = (void *)ctx->stream;
  for (int g = 0; g < nb; g++) {
    lvc_batch *b = &bs[g];
    if (s >= b->steps)
      continue;
    const int tok = lvc_prof_begin(ctx, "lvc_xsweep");
    const int e = lvc_launch_xsweep(st, b->precision, b->seg_x, b->d_probs, b->d_tiles,
                                    b->x_active[s], b->d_geo, b->d_R, b->d_Y, b->d_yfac,
                                    lvc_seg_is_col(b->seg_y) ? b->seg_y : 0, b->d_fprobs,
                                    b->f_active[s], b->xshared);
    lvc_prof_end(ctx, tok);
    if (lvc_check_hip(ctx, e, "x sweep kernel"))
      return 1;
  }
  for (int g = 0; g < nb; g++) {
    lvc_batch *b = &bs[g];
    if (s >= b->steps)
      continue;
    const int tok = lvc_prof_begin(ctx, "lvc_ysweep");
    const int e = lvc_launch_ysweep(st, b->precision, b->seg_y, b->d_probs,
                                    b->d_tiles + b->nxt, b->y_active[s], b->d_geo, b->d_Y,
                                    b->d_R, b->ylds, b->d_yfac);
    lvc_prof_end(ctx, tok);
    if (lvc_check_hip(ctx, e, "y sweep kernel"))
      return 1;
  }
  return 0;
}

/* FNV-1a over the arguments of every launch of a step sequence: equal keys,
 * equal launches. */
static uint64_t fnv(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = p;
  for (size_t i = 0; i < n; i++)
    h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

static uint64_t launch_key(const lvc_batch *bs, int nb) {
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, &nb, sizeof(nb));
  for (int g = 0; g < nb; g++) {
    const lvc_batch *b = &bs[g];
    const void *ptrs[7] = {b->d_probs, b->d_tiles, b->d_geo, b->d_R, b->d_Y, b->d_yfac,
                           b->d_fprobs};
    const int ints[8] = {b->precision, b->seg_x, b->seg_y, b->nxt, b->nyt, b->xshared,
                         b->ylds, (int)b->steps};
    h = fnv(h, ptrs, sizeof(ptrs));
    h = fnv(h, ints, sizeof(ints));
    h = fnv(h, b->x_active, (size_t)b->steps * sizeof(int));
    h = fnv(h, b->y_active, (size_t)b->steps * sizeof(int));
    h = fnv(h, b->f_active, (size_t)b->steps * sizeof(int));
  }
  return h;
}

/* The step sequence captured into a HIP graph and launched as one.  The
 * context keeps the instantiated graph of its last sequence: a call whose
 * launches are the same (sizes, active tiles and buffers: the memory pool
 * hands a repeated call the same blocks) launches it again without capture;
 * another sequence is captured and the kept graph updated in place
 * (hipGraphExecUpdate), or instantiated anew when the update is refused. */
static int run_graph(struct futhark_context *ctx, lvc_batch *bs, int nb, int64_t steps) {
  const uint64_t key = launch_key(bs, nb);
  if (!(ctx->graph_exec && ctx->graph_key == key)) {
    hipGraph_t graph = NULL;
    if (lvc_check_hip(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture"))
      return 1;
    int rc = 0;
    for (int64_t s = 0; !rc && s < steps; s++)
      rc = launch_step(ctx, bs, nb, s);
    const int e = (int)hipStreamEndCapture(ctx->stream, &graph);
    if (rc || lvc_check_hip(ctx, e, "hipStreamEndCapture")) {
      if (graph)
        (void)hipGraphDestroy(graph);
      return 1;
    }
    int updated = 0;
    if (ctx->graph_exec) {
      hipGraphExecUpdateResult res;
      hipGraphNode_t bad = NULL;
      updated = hipGraphExecUpdate((hipGraphExec_t)ctx->graph_exec, graph, &bad, &res) ==
                    hipSuccess && res == hipGraphExecUpdateSuccess;
      if (!updated) {
        /* a refused update (another topology) is not an error here: clear
         * the thread's last error, which the next launch check would read */
        (void)hipGetLastError();
        lvc_graph_release(ctx);
      }
    }
    if (!updated) {
      hipGraphExec_t exec = NULL;
      rc = lvc_check_hip(ctx, hipGraphInstantiate(&exec, graph, NULL, NULL, 0),
                         "hipGraphInstantiate");
      ctx->graph_exec = rc ? NULL : exec;
    }
    (void)hipGraphDestroy(graph);
    if (rc)
      return 1;
    ctx->graph_key = key;
  }
  return lvc_check_hip(ctx, hipGraphLaunch((hipGraphExec_t)ctx->graph_exec, ctx->stream),
                       "hipGraphLaunch");
}

/* Runs the batches in bs together: their setups, then the time steps, each
 * step every batch's x sweep then every batch's y sweep. */
static int run_batches(struct futhark_context *ctx, lvc_batch *bs, int nb) {
  int64_t steps = 0;
  for (int g = 0; g < nb; g++) {
    lvc_batch *b = &bs[g];
    if (b->n == 0)
      continue;
    const int tok = lvc_prof_begin(ctx, "lvc_setup");
    const int e = lvc_launch_setup((void *)ctx->stream, b->precision, b->d_probs, b->n,
                                   b->d_geo, b->d_R);
    lvc_prof_end(ctx, tok);
    if (lvc_check_hip(ctx, e, "setup kernel"))
      return 1;
    if (b->steps > steps)
      steps = b->steps;
  }
  const int profiling = ctx->cfg && ctx->cfg->profiling && !ctx->prof_paused;
  if (nb > 0 && bs[0].graph && !profiling && steps > 0)
    return run_graph(ctx, bs, nb, steps);
  for (int64_t s = 0; s < steps; s++)
    if (launch_step(ctx, bs, nb, s))
      return 1;
  return 0;
}

int lvc_batch_run(struct futhark_context *ctx, lvc_batch *b) {
  return run_batches(ctx, b, 1);
}

int lvc_batch_gather(struct futhark_context *ctx, lvc_batch *b, void *out,
                     int out_f64) {
  if (b->n == 0)
    return 0;
  int tok = lvc_prof_begin(ctx, "lvc_gather");
  int e = lvc_launch_gather((void *)ctx->stream, b->precision, b->d_probs, b->n,
                            b->d_R, out, out_f64);
  lvc_prof_end(ctx, tok);
  return lvc_check_hip(ctx, e, "gather kernel");
}

static int download(struct futhark_context *ctx, void *dst, const void *src,
                    size_t bytes);

/* Runs a batch end to end into a new device array of n prices.  Each
 * problem takes its own size class's segment pair; the problems of each
 * distinct pair form a sub-batch, the sub-batches advance together, and
 * each problem's price goes to its own place in the output. */
#define LVC_MAX_GROUPS 16
static int run_batch(struct futhark_context *ctx, int n, const lvc_problem *probs,
                     int precision, int out_f64, void **out_mem) {
  const size_t ob = (size_t)n * (out_f64 ? sizeof(double) : sizeof(float));
  int gsx[LVC_MAX_GROUPS], gsy[LVC_MAX_GROUPS], gcnt[LVC_MAX_GROUPS] = {0}, ng = 0;
  int *grp = malloc((size_t)(n ? n : 1) * sizeof(int));
  lvc_problem *sub = malloc((size_t)(n ? n : 1) * sizeof(lvc_problem));
  if (!grp || !sub) {
    free(grp);
    free(sub);
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  for (int p = 0; p < n; p++) {
    int sx, sy, g;
    lvc_choose_segs(ctx, precision, probs[p].nx, probs[p].ny, &sx, &sy);
    for (g = 0; g < ng && (gsx[g] != sx || gsy[g] != sy); g++)
      ;
    if (g == ng) {
      if (ng == LVC_MAX_GROUPS) { /* cannot happen: 3 x 5 segment pairs */
        free(grp);
        free(sub);
        lvc_set_error(ctx, "too many segment classes");
        return 1;
      }
      gsx[ng] = sx;
      gsy[ng] = sy;
      ng++;
    }
    grp[p] = g;
    gcnt[g]++;
  }
  lvc_batch bs[LVC_MAX_GROUPS];
  memset(bs, 0, sizeof(bs));
  int rc = lvc_dev_alloc(ctx, ob, out_mem);
  for (int g = 0; !rc && g < ng; g++) {
    int m = 0;
    for (int p = 0; p < n; p++)
      if (grp[p] == g) {
        sub[m] = probs[p];
        sub[m].out_index = p;
        m++;
      }
    rc = lvc_batch_build(ctx, &bs[g], m, sub, precision, gsx[g], gsy[g]);
  }
  if (!rc)
    rc = run_batches(ctx, bs, ng);
  for (int g = 0; !rc && g < ng; g++)
    rc = lvc_batch_gather(ctx, &bs[g], *out_mem, out_f64);
  if (!rc && n > 0 && ctx->cfg && ctx->cfg->debugging) {
    /* the index the gather read: problem 0's device descriptor (the first
     * problem of the first sub-batch) */
    lvc_problem q;
    rc = download(ctx, &q, bs[0].d_probs, sizeof(q));
    if (!rc) {
      ctx->dbg_index[0] = q.yindex;
      ctx->dbg_index[1] = q.xindex;
      ctx->dbg_index_valid = 1;
    }
  }
  for (int g = 0; g < ng; g++)
    lvc_batch_release(ctx, &bs[g]);
  free(grp);
  free(sub);
  if (rc) {
    lvc_dev_free(ctx, *out_mem, ob);
    *out_mem = NULL;
  }
  return rc;
}

static void fill_main(lvc_problem *probs, int outer, int nx, int ny, int nt,
                      double s0, double t, double alpha, double nu, double beta,
                      int grid32) {
  memset(probs, 0, (size_t)outer * sizeof(*probs));
  for (int i = 0; i < outer; i++) {
    probs[i].nx = nx;
    probs[i].ny = ny;
    probs[i].nt = nt;
    probs[i].s0 = s0;
    probs[i].t = t;
    probs[i].alpha = alpha;
    probs[i].nu = nu;
    probs[i].beta = beta;
    probs[i].grid32 = grid32;
  }
}

/* The price index of main's options (all share initGrid's scalars): the
 * reference's `myResult[myYindex, myXindex]` (:236) is bounds-checked, so an
 * index outside the grid fails the call, as the generated code does, before
 * any device work. */
static int check_main_index(struct futhark_context *ctx, int outer, int nx, int ny,
                            double s0, double t, double alpha, double nu,
                            int grid32) {
  if (outer == 0)
    return 0;
  lvc_problem q;
  memset(&q, 0, sizeof(q));
  q.nx = nx;
  q.ny = ny;
  q.s0 = s0;
  q.t = t;
  q.alpha = alpha;
  q.nu = nu;
  lvc_grid_scalars(&q, grid32);
  if (q.status == 0)
    return 0;
  lvc_set_error(ctx, "Index [%d, %d] out of bounds for array of shape [%d][%d].",
                q.yindex, q.xindex, ny, nx);
  return 1;
}

/* LocVolCalib.fut:238-252 */
int futhark_entry_main(struct futhark_context *ctx, struct futhark_f32_1d **out0,
                       const int32_t outer, const int32_t numX,
                       const int32_t numY, const int32_t numT, const float s0,
                       const float t, const float alpha, const float nu,
                       const float beta) {
  if (!out0 || !ctx)
    return 1;
  *out0 = NULL;
  if (outer < 0) {
    lvc_set_error(ctx, "outer_loop_count must be non-negative (got %d)", outer);
    return 1;
  }
  if (lvc_validate_sizes(ctx, numX, numY, numT) || !lvc_ready(ctx) ||
      check_main_index(ctx, outer, numX, numY, s0, t, alpha, nu, 1))
    return 1;
  lvc_problem *probs = malloc((size_t)(outer ? outer : 1) * sizeof(*probs));
  if (!probs) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  /* the f32 entry: initGrid's scalars in f32, as the reference */
  fill_main(probs, outer, numX, numY, numT, s0, t, alpha, nu, beta, 1);
  for (int i = 0; i < outer; i++)
    probs[i].strike = (double)(0.001f * (float)(int64_t)i); /* :247, in f32 */
  pthread_mutex_lock(&ctx->lock);
  void *mem = NULL;
  int rc = run_batch(ctx, outer, probs, ctx->cfg ? ctx->cfg->precision
                                                 : LVC_FP32_REFINE, 0, &mem);
  pthread_mutex_unlock(&ctx->lock);
  free(probs);
  if (rc)
    return 1;
  *out0 = lvc_array_adopt_f32(mem, outer);
  if (!*out0) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  return 0;
}

int futhark_entry_main_f64(struct futhark_context *ctx,
                           struct futhark_f64_1d **out0, const int32_t outer,
                           const int32_t numX, const int32_t numY,
                           const int32_t numT, const double s0, const double t,
                           const double alpha, const double nu,
                           const double beta) {
  if (!out0 || !ctx)
    return 1;
  *out0 = NULL;
  if (outer < 0) {
    lvc_set_error(ctx, "outer_loop_count must be non-negative (got %d)", outer);
    return 1;
  }
  if (lvc_validate_sizes(ctx, numX, numY, numT) || !lvc_ready(ctx) ||
      check_main_index(ctx, outer, numX, numY, s0, t, alpha, nu, 0))
    return 1;
  lvc_problem *probs = malloc((size_t)(outer ? outer : 1) * sizeof(*probs));
  if (!probs) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  fill_main(probs, outer, numX, numY, numT, s0, t, alpha, nu, beta, 0);
  for (int i = 0; i < outer; i++)
    probs[i].strike = 0.001 * (double)(int64_t)i;
  pthread_mutex_lock(&ctx->lock);
  void *mem = NULL;
  int rc = run_batch(ctx, outer, probs, LVC_FP64, 1, &mem);
  pthread_mutex_unlock(&ctx->lock);
  free(probs);
  if (rc)
    return 1;
  *out0 = lvc_array_adopt_f64(mem, outer);
  if (!*out0) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  return 0;
}

void lvc_context_ready(struct futhark_context *ctx) {
  lvc_tune_load(ctx);
  /* the device is current here (futhark_context_new selected it) */
  ctx->scratch_err = lvc_scratch_address(ctx->device, &ctx->scratch);
}

static int64_t align256(int64_t v) { return (v + 255) & ~(int64_t)255; }

/* futhark_entry_main without device allocation: the same options, segment
 * lengths, tiles and kernels as futhark_entry_main (so the same bits), with
 * every buffer carved from the static device scratch and the descriptors and
 * tiles written there by a kernel; prices into the caller's device array. */
int futhark_entry_main_noalloc(struct futhark_context *ctx, float *out,
                               const int32_t outer, const int32_t numX,
                               const int32_t numY, const int32_t numT,
                               const float s0, const float t, const float alpha,
                               const float nu, const float beta) {
  if (!ctx)
    return 1;
  if (outer < 0) {
    lvc_set_error(ctx, "outer_loop_count must be non-negative (got %d)", outer);
    return 1;
  }
  if (lvc_validate_sizes(ctx, numX, numY, numT) || !lvc_ready(ctx) ||
      check_main_index(ctx, outer, numX, numY, s0, t, alpha, nu, 1))
    return 1;
  if (outer == 0)
    return 0;
  if (!out) {
    lvc_set_error(ctx, "main_noalloc: NULL output array");
    return 1;
  }
  if (!ctx->scratch) {
    lvc_set_error(ctx, "main_noalloc: the static device scratch of device %d is unavailable "
                       "(%s)", ctx->device, hipGetErrorString((hipError_t)ctx->scratch_err));
    return 1;
  }
  const int precision = ctx->cfg ? ctx->cfg->precision : LVC_FP32_REFINE;
  const int64_t real = precision == LVC_FP64 ? 8 : 4;
  int sx = ctx->cfg ? ctx->cfg->seg_x : 0, sy = ctx->cfg ? ctx->cfg->seg_y : 0, tx = 0, ty = 0;
  if (!sx || !sy)
    lvc_tune_peek(ctx, precision, numX, numY, &tx, &ty); /* never times (that allocates) */
  sx = fit_seg(numX, sx ? sx : tx, 0);
  sy = fit_seg(numY, sy ? sy : ty, 1);

  lvc_problem tmpl;
  memset(&tmpl, 0, sizeof(tmpl));
  fill_main(&tmpl, 1, numX, numY, numT, s0, t, alpha, nu, beta, 1);
  lvc_grid_scalars(&tmpl, 1);
  lvc_main_layout lay;
  memset(&lay, 0, sizeof(lay));
  lay.outer = outer;
  lay.gx = lvc_lanes_for(numX, sx);
  lay.rows = LVC_BLOCK / lay.gx;
  lay.xtpg = (numY + lay.rows - 1) / lay.rows;
  lvc_ytile_shape(numY, sy, (int)real, &lay.gy, &lay.ycpl, &lay.cols);
  lay.ytpo = (numX + lay.cols - 1) / lay.cols;
  int xshare = lvc_xshare(ctx, precision);
  if (xshare <= 0) {
    const int runs[1] = {outer}, tiles[1] = {lay.xtpg};
    xshare = batch_xshare(ctx, precision, 1, runs, tiles);
  }
  lay.xshare = xshare;
  lay.groups = (outer + xshare - 1) / xshare;
  lay.strike_f32 = 1;
  lay.geo_stride = align64(LVC_GEO_LEN(numX, numY));
  lay.state_stride = align64((int64_t)numX * numY);
  const int64_t nxt = (int64_t)lay.groups * lay.xtpg, nyt = (int64_t)outer * lay.ytpo;
  const int col = lvc_seg_is_col(sy);

  /* the scratch, in order: descriptors, tiles, leaders, geometry, R, Y, y factors */
  const int64_t o_tiles = align256((int64_t)outer * (int64_t)sizeof(lvc_problem));
  const int64_t o_lead = o_tiles + align256((nxt + nyt) * (int64_t)sizeof(lvc_tile));
  const int64_t o_geo = o_lead + align256((int64_t)lay.groups * 4);
  const int64_t o_R = o_geo + align256((int64_t)outer * lay.geo_stride * 8);
  const int64_t o_Y = o_R + align256((int64_t)outer * lay.state_stride * real);
  const int64_t o_fac = o_Y + align256((int64_t)outer * lay.state_stride * real);
  const int64_t need = o_fac + (col ? align256((int64_t)outer * LVC_YFAC_LEN * real) : 0);
  if (need > LVC_SCRATCH_BYTES) {
    lvc_set_error(ctx, "main_noalloc: main(%d, %d, %d, %d) needs %lld bytes of device scratch; "
                       "the static scratch holds %lld",
                  outer, numX, numY, numT, (long long)need, (long long)LVC_SCRATCH_BYTES);
    return 1;
  }
  char *base = (char *)ctx->scratch;
  lvc_problem *probs = (lvc_problem *)base;
  lvc_tile *tl = (lvc_tile *)(base + o_tiles);
  int32_t *lead = (int32_t *)(base + o_lead);
  double *geo = (double *)(base + o_geo);
  void *R = base + o_R, *Y = base + o_Y, *yfac = col ? base + o_fac : NULL;
  const int ylds = lvc_ytile_lds(numY, sy, (int)real);
  const int xshared = (xshare < outer ? xshare : outer) > 1;

  void *st = (void *)ctx->stream;
  pthread_mutex_lock(&ctx->lock);
  int e = lvc_launch_main_init(st, &tmpl, &lay, probs, tl, lead);
  if (!e)
    e = lvc_launch_setup(st, precision, probs, outer, geo, R);
  for (int s = 0; !e && s < numT - 1; s++) {
    e = lvc_launch_xsweep(st, precision, sx, probs, tl, (int)nxt, geo, R, Y, yfac,
                          col ? sy : 0, lead, lay.groups, xshared);
    if (!e)
      e = lvc_launch_ysweep(st, precision, sy, probs, tl + nxt, (int)nyt, geo, Y, R, ylds,
                            yfac);
  }
  if (!e)
    e = lvc_launch_gather(st, precision, probs, outer, R, out, 0);
  const int rc = lvc_check_hip(ctx, e, "main_noalloc kernel launch");
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}

static int download(struct futhark_context *ctx, void *dst, const void *src,
                    size_t bytes) {
  if (bytes == 0)
    return 0;
  if (lvc_check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                                        ctx->stream),
                    "hipMemcpyAsync"))
    return 1;
  return lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),
                       "hipStreamSynchronize");
}

int futhark_entry_main_batch(struct futhark_context *ctx,
                             struct futhark_f64_1d **out0,
                             const struct futhark_i32_1d *numX,
                             const struct futhark_i32_1d *numY,
                             const struct futhark_i32_1d *numT,
                             const struct futhark_f64_1d *s0,
                             const struct futhark_f64_1d *strike,
                             const struct futhark_f64_1d *t,
                             const struct futhark_f64_1d *alpha,
                             const struct futhark_f64_1d *nu,
                             const struct futhark_f64_1d *beta,
                             int32_t precision) {
  if (!out0)
    return 1;
  *out0 = NULL;
  if (!lvc_ready(ctx))
    return 1;
  if (!numX || !numY || !numT || !s0 || !strike || !t || !alpha || !nu || !beta) {
    lvc_set_error(ctx, "NULL input array");
    return 1;
  }
  if (precision < LVC_FP32 || precision > LVC_FP64) {
    lvc_set_error(ctx, "unknown precision %d", precision);
    return 1;
  }
  const int64_t n = lvc_array_len_i32(numX);
  const struct futhark_f64_1d *reals[6] = {s0, strike, t, alpha, nu, beta};
  if (lvc_array_len_i32(numY) != n || lvc_array_len_i32(numT) != n) {
    lvc_set_error(ctx, "input arrays differ in length");
    return 1;
  }
  for (int k = 0; k < 6; k++)
    if (lvc_array_len_f64(reals[k]) != n) {
      lvc_set_error(ctx, "input arrays differ in length");
      return 1;
    }
  if (n > (1 << 24)) {
    lvc_set_error(ctx, "too many problems (%lld)", (long long)n);
    return 1;
  }
  int32_t *iv = malloc((size_t)(n ? n : 1) * 3 * sizeof(int32_t));
  double *rv = malloc((size_t)(n ? n : 1) * 6 * sizeof(double));
  lvc_problem *probs = calloc((size_t)(n ? n : 1), sizeof(*probs));
  if (!iv || !rv || !probs) {
    free(iv);
    free(rv);
    free(probs);
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  pthread_mutex_lock(&ctx->lock);
  int rc = 0;
  rc = rc || download(ctx, iv, lvc_array_mem_i32(numX), (size_t)n * 4);
  rc = rc || download(ctx, iv + n, lvc_array_mem_i32(numY), (size_t)n * 4);
  rc = rc || download(ctx, iv + 2 * n, lvc_array_mem_i32(numT), (size_t)n * 4);
  for (int k = 0; k < 6 && !rc; k++)
    rc = download(ctx, rv + k * n, lvc_array_mem_f64(reals[k]), (size_t)n * 8);
  for (int64_t p = 0; p < n && !rc; p++) {
    probs[p].nx = iv[p];
    probs[p].ny = iv[n + p];
    probs[p].nt = iv[2 * n + p];
    probs[p].s0 = rv[p];
    probs[p].strike = rv[n + p];
    probs[p].t = rv[2 * n + p];
    probs[p].alpha = rv[3 * n + p];
    probs[p].nu = rv[4 * n + p];
    probs[p].beta = rv[5 * n + p];
    if (lvc_validate_sizes(ctx, probs[p].nx, probs[p].ny, probs[p].nt)) {
      char *inner = futhark_context_get_error(ctx);
      lvc_set_error(ctx, "problem %lld: %s", (long long)p, inner ? inner : "");
      free(inner);
      rc = 1;
    }
  }
  void *mem = NULL;
  if (!rc)
    rc = run_batch(ctx, (int)n, probs, precision, 1, &mem);
  pthread_mutex_unlock(&ctx->lock);
  free(iv);
  free(rv);
  free(probs);
  if (rc)
    return 1;
  *out0 = lvc_array_adopt_f64(mem, n);
  if (!*out0) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  return 0;
}
