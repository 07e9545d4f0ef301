/*
 * Entry points of finpar/InterestCalib.fut (`main`, :929-936) for
 * libinterestcalib.so.  This file checks what the reference's indexing and
 * sizes impose, fixes the move schedule (which depends only on Sobol
 * numbers at data-independent offsets), lays out the swaption tables and
 * the wavefront work lists, and launches csrc/ic_kernels.hip on the
 * context's stream: per call one setup, one initial evaluation, then three
 * launches per MCMC iteration (evaluate, accept, propose), the winner and
 * its summary, with one synchronisation at the end for the six scalar
 * results.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/interestcalib.h"
#include "ic_internal.h"
#include "lvc_ndarray.h"
#include "lvc_runtime.h"

/* The kernels take no tuning parameters. */
const lvc_program lvc_prog = {0, NULL, NULL, NULL, NULL};

/* sobolInd (:190-202) on the host, for the move schedule: the same integer
 * and f32 operations as the kernels. */
static float host_sobol(const ic_dirvec *dv, int32_t n) {
  const float norm_fact = 1.0f / ((float)((int64_t)1 << dv->bits) + 1.0f);
  const int32_t n_gray = (n >> 1) ^ n;
  int32_t res = 0;
  for (int i = 0; i < dv->bits; i++) {
    const int32_t t = (int32_t)(1u << i);
    if ((n_gray & t) == t) res ^= dv->dir[i];
  }
  return (float)res * norm_fact;
}

/* The moves of iterations 0..conv-1 (:648-690) and the Sobol offsets of
 * their acceptance draws (:737); the offset advances by move and pop only. */
static void move_schedule(int32_t pop, int32_t conv, const ic_dirvec *dv, ic_move *moves,
                          int32_t *acc, int *has_de) {
  int32_t offs = 5 * pop + 1;
  *has_de = 0;
  for (int32_t it = 0; it < conv; it++) {
    const float r01 = host_sobol(dv, offs);
    ic_move m = {r01 <= 0.2f ? 1 : r01 <= 0.5f ? 2 : 3, 0, 0, 0}; /* selectMoveType, :240 */
    offs = offs + 1;
    if (m.type == 1) {
      m.base = offs;
      offs = offs + 5 * pop;
    } else if (m.type == 2) {
      m.dim = (int32_t)(host_sobol(dv, offs) * (float)5);
      m.base = offs + 1;
      offs = offs + 5 * pop + 1;
    } else {
      m.base = offs;
      offs = offs + 8 * pop;
      *has_de = 1;
    }
    moves[it] = m;
    acc[it] = offs;
    offs = offs + pop;
  }
}

/* The lanes of one evaluation and the schedule periods each lane holds.
 * An iteration's evaluations fill only a few wavefronts per SIMD, so a
 * launch lasts about as long as its slowest wavefront: 2 periods a lane (1
 * for a one-period swaption) measured fastest on the reference's datasets
 * (against 1, 4 and 8), up to 128 periods; longer schedules take 4 or 8
 * periods a lane in the second launch. */
static void layout_of(int32_t n, int32_t *lanes, int32_t *per_lane) {
  const int32_t per = n <= 1 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : 8;
  int32_t l = 1;
  while (l * per < n) l <<= 1;
  *lanes = l;
  *per_lane = per;
}

/* Everything one call keeps on the device. */
typedef struct {
  int32_t nswap, pop, nherm;
  ic_swaption *sw;
  ic_period *per;
  ic_work *work;          /* the work lists below, back to back */
  ic_worklists pop_lists; /* for `pop` genomes */
  ic_worklists one_lists; /* for one genome (the winner's summary) */
  float *state;  /* genomes, props [pop][5]; logliks, fb [pop]; out [8], winner [8] */
  float *terms;  /* [nswap][pop] */
  size_t sw_bytes, per_bytes, work_bytes, state_bytes, terms_bytes;
} ic_call;

static void call_release(struct futhark_context *ctx, ic_call *c) {
  if (c->sw) lvc_dev_free(ctx, c->sw, c->sw_bytes);
  if (c->per) lvc_dev_free(ctx, c->per, c->per_bytes);
  if (c->work) lvc_dev_free(ctx, c->work, c->work_bytes);
  if (c->state) lvc_dev_free(ctx, c->state, c->state_bytes);
  if (c->terms) lvc_dev_free(ctx, c->terms, c->terms_bytes);
  memset(c, 0, sizeof(*c));
}

/* Insertion sort of swaption indices by descending schedule length, ties
 * in index order (stable). */
static void sort_by_periods(int64_t *order, int64_t n, const ic_swaption *sw) {
  for (int64_t i = 1; i < n; i++) {
    const int64_t v = order[i];
    int64_t j = i;
    while (j > 0 && sw[order[j - 1]].n < sw[v].n) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = v;
  }
}

static int64_t add_work(ic_work *w, int64_t k, int32_t s, int32_t lanes, int32_t npop) {
  const int32_t per_wave = 64 / lanes;
  for (int32_t g = 0; g < npop; g += per_wave) {
    w[k].swaption = s;
    w[k].genome0 = g;
    k++;
  }
  return k;
}

/* Reads the swaptions back to size their schedules, checks them, and builds
 * and uploads the tables and work lists for `pop` genomes (and one, for the
 * summary); enqueues the setup kernel.  Called with ctx->lock held. */
static int call_prepare(struct futhark_context *ctx, ic_call *c, const lvc_ndarray *swp,
                        int32_t pop, int32_t nherm) {
  memset(c, 0, sizeof(*c));
  const int64_t nswap = swp->shape[0], cols = swp->shape[1];
  if (nswap > INT32_MAX / 2) {
    lvc_set_error(ctx, "main: %lld swaptions", (long long)nswap);
    return 1;
  }
  c->nswap = (int32_t)nswap;
  c->pop = pop;
  c->nherm = nherm;
  if (nswap == 0) return 0;
  if (cols < 4) {
    lvc_set_error(ctx, "main: a swaption has %lld fields; evalGenomeOnSwap reads swaption[3] "
                  "(:489-523): index out of bounds", (long long)cols);
    return 1;
  }
  float *host = malloc((size_t)(nswap * cols) * sizeof(float));
  ic_swaption *sw = calloc((size_t)nswap, sizeof(ic_swaption));
  if (!host || !sw) {
    free(host);
    free(sw);
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  int rc = lvc_check_hip(ctx, hipMemcpyAsync(host, swp->mem, (size_t)(nswap * cols) * sizeof(float),
                                             hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpyAsync") ||
           lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  int64_t nper = 0, nwp[IC_LAUNCHES] = {0}, nwo[IC_LAUNCHES] = {0};
  for (int64_t s = 0; !rc && s < nswap; s++) {
    const float *row = host + s * cols;
    const float periods = 12.0f * row[2] / row[1]; /* i64.f32 (:491) */
    if (!(periods > -1.0f)) {
      lvc_set_error(ctx, "main: swaption %lld has %g schedule periods: iota of a negative "
                    "count (:491, :530)", (long long)s, (double)periods);
      rc = 1;
    } else if (!(periods < (float)(IC_MAX_PERIODS + 1))) {
      lvc_set_error(ctx, "main: swaption %lld has %g schedule periods (at most %d)",
                    (long long)s, (double)periods, IC_MAX_PERIODS);
      rc = 1;
    } else {
      const int32_t n = (int32_t)(int64_t)periods;
      int32_t lanes, per;
      layout_of(n, &lanes, &per);
      sw[s].n = n;
      sw[s].off = (int32_t)nper;
      sw[s].lanes = lanes;
      sw[s].per_lane = per;
      nper += n;
      nwp[per > 2] += (pop + 64 / lanes - 1) / (64 / lanes);
      nwo[per > 2] += 1;
    }
  }
  ic_work *work = NULL;
  int64_t *order = NULL;
  int64_t total = 0;
  for (int cl = 0; cl < IC_LAUNCHES; cl++) total += nwp[cl] + nwo[cl];
  if (!rc) {
    work = malloc((size_t)(total > 0 ? total : 1) * sizeof(ic_work));
    if (!work) {
      lvc_set_error(ctx, "host allocation failed");
      rc = 1;
    }
  }
  if (!rc) {
    order = malloc((size_t)nswap * sizeof(int64_t));
    if (!order) {
      lvc_set_error(ctx, "host allocation failed");
      rc = 1;
    }
  }
  int64_t start[2][IC_LAUNCHES];
  if (!rc) {
    /* the longest schedules first: their wavefronts run longest */
    for (int64_t s = 0; s < nswap; s++) order[s] = s;
    sort_by_periods(order, nswap, sw);
    int64_t k = 0;
    for (int one = 0; one < 2; one++) /* pop's lists by launch, then one genome's */
      for (int cl = 0; cl < IC_LAUNCHES; cl++) {
        start[one][cl] = k;
        for (int64_t r = 0; r < nswap; r++) {
          const int64_t s = order[r];
          if ((sw[s].per_lane > 2) == cl)
            k = add_work(work, k, (int32_t)s, sw[s].lanes, one ? 1 : pop);
        }
      }
    for (int cl = 0; cl < IC_LAUNCHES; cl++) {
      c->pop_lists.nwork[cl] = (int32_t)nwp[cl];
      c->one_lists.nwork[cl] = (int32_t)nwo[cl];
    }
    c->sw_bytes = (size_t)nswap * sizeof(ic_swaption);
    c->per_bytes = (size_t)(nper > 0 ? nper : 1) * sizeof(ic_period);
    c->work_bytes = (size_t)k * sizeof(ic_work);
    rc = lvc_dev_alloc(ctx, c->sw_bytes, (void **)&c->sw) ||
         lvc_dev_alloc(ctx, c->per_bytes, (void **)&c->per) ||
         lvc_dev_alloc(ctx, c->work_bytes, (void **)&c->work) ||
         lvc_upload(ctx, c->sw, sw, c->sw_bytes) || lvc_upload(ctx, c->work, work, c->work_bytes);
    for (int cl = 0; !rc && cl < IC_LAUNCHES; cl++) {
      c->pop_lists.work[cl] = c->work + start[0][cl];
      c->one_lists.work[cl] = c->work + start[1][cl];
    }
  }
  if (!rc)
    rc = lvc_check_hip(ctx, ic_launch_setup((void *)ctx->stream, swp->mem, cols, (int32_t)nswap,
                                            c->sw, c->per),
                       "InterestCalib setup kernel");
  free(host);
  free(sw);
  free(work);
  free(order);
  return rc;
}

static int alloc_state(struct futhark_context *ctx, ic_call *c) {
  const size_t pop = (size_t)c->pop;
  c->state_bytes = (pop * 12 + 16) * sizeof(float);
  c->terms_bytes = ((size_t)c->nswap * pop > 0 ? (size_t)c->nswap * pop : 1) * sizeof(float);
  return lvc_dev_alloc(ctx, c->state_bytes, (void **)&c->state) ||
         lvc_dev_alloc(ctx, c->terms_bytes, (void **)&c->terms);
}

/* The arrays main takes, checked; fills dv. */
static int check_inputs(struct futhark_context *ctx, const struct futhark_f32_2d *swaptions,
                        const struct futhark_f32_1d *herm_coefs,
                        const struct futhark_f32_1d *herm_weights,
                        const struct futhark_i32_1d *sob_dir_vct, ic_dirvec *dv) {
  if (!swaptions || !herm_coefs || !herm_weights || !sob_dir_vct) {
    lvc_set_error(ctx, "main: NULL input array");
    return 1;
  }
  const int64_t nh = lvc_array_len_f32(herm_coefs);
  if (lvc_array_len_f32(herm_weights) != nh) {
    lvc_set_error(ctx, "main: %lld Hermite coefficients and %lld weights (zip, :933)",
                  (long long)nh, (long long)lvc_array_len_f32(herm_weights));
    return 1;
  }
  if (nh > INT32_MAX) {
    lvc_set_error(ctx, "main: too many Hermite points");
    return 1;
  }
  const int64_t bits = lvc_array_len_i32(sob_dir_vct);
  if (bits > IC_MAX_BITS) {
    lvc_set_error(ctx, "main: %lld direction-vector bits (at most %d: sobolInd's i32 shifts, "
                  ":199-200)", (long long)bits, IC_MAX_BITS);
    return 1;
  }
  memset(dv, 0, sizeof(*dv));
  dv->bits = (int32_t)bits;
  if (bits > 0 &&
      (lvc_check_hip(ctx, hipMemcpyAsync(dv->dir, lvc_array_mem_i32(sob_dir_vct),
                                         (size_t)bits * sizeof(int32_t), hipMemcpyDeviceToHost,
                                         ctx->stream),
                     "hipMemcpyAsync") ||
       lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")))
    return 1;
  return 0;
}

/* main's whole calibration (:627-849) with optional trajectory recording
 * into device buffers.  Called with ctx->lock held; c prepared. */
static int run_calibration(struct futhark_context *ctx, ic_call *c, int32_t conv,
                           const ic_dirvec *dv, const ic_move *moves, const int32_t *acc,
                           const float *hx, const float *hw, const ic_trace *trace,
                           float *summary) {
  void *st = (void *)ctx->stream;
  const int32_t pop = c->pop;
  float *genomes = c->state, *props = genomes + 5 * (size_t)pop;
  float *logliks = props + 5 * (size_t)pop, *fb = logliks + pop;
  float *out = fb + pop, *winner = out + 8;
  const ic_move none = {0, 0, 0, 0};
  /* with profiling on, each launch is timed under its name (events on the
   * context's stream): ic_eval (the population's evaluations), ic_step
   * (accept + propose), ic_other (init, winner, summary) */
  int tok = lvc_prof_begin(ctx, "ic_other");
  int e = ic_launch_init(st, pop, dv, genomes, props);
  lvc_prof_end(ctx, tok);
  if (!e) {
    tok = lvc_prof_begin(ctx, "ic_eval");
    e = ic_launch_eval(st, c->sw, c->per, &c->pop_lists, genomes, pop, hx, hw, c->nherm,
                       c->nswap, IC_OUT_TERMS, c->terms, NULL);
    lvc_prof_end(ctx, tok);
  }
  if (!e) {
    tok = lvc_prof_begin(ctx, "ic_step");
    e = ic_launch_step(st, pop, c->nswap, c->terms, genomes, props, logliks, fb, 1, 0,
                       conv > 0 ? &moves[0] : &none, dv, -1, conv, trace);
    lvc_prof_end(ctx, tok);
  }
  for (int32_t it = 0; !e && it < conv; it++) {
    tok = lvc_prof_begin(ctx, "ic_eval");
    e = ic_launch_eval(st, c->sw, c->per, &c->pop_lists, props, pop, hx, hw, c->nherm,
                       c->nswap, IC_OUT_TERMS, c->terms, NULL);
    lvc_prof_end(ctx, tok);
    if (!e) {
      tok = lvc_prof_begin(ctx, "ic_step");
      e = ic_launch_step(st, pop, c->nswap, c->terms, genomes, props, logliks, fb, 0, acc[it],
                         it + 1 < conv ? &moves[it + 1] : &none, dv, it, conv, trace);
      lvc_prof_end(ctx, tok);
    }
  }
  tok = lvc_prof_begin(ctx, "ic_other");
  if (!e) e = ic_launch_winner(st, pop, genomes, logliks, out, winner);
  if (!e && c->nswap > 0)
    e = ic_launch_eval(st, c->sw, c->per, &c->one_lists, winner, 1, hx, hw, c->nherm, c->nswap,
                       IC_OUT_SUMMARY, summary, NULL);
  lvc_prof_end(ctx, tok);
  return lvc_check_hip(ctx, e, "InterestCalib kernels");
}

/* Validation shared by the entry points; fills dv, the schedule (moves,
 * acc: conv entries each, malloc'd) and the device tables. */
static int begin_call(struct futhark_context *ctx, ic_call *c, int32_t pop, int32_t conv,
                      const struct futhark_f32_2d *swaptions,
                      const struct futhark_f32_1d *herm_coefs,
                      const struct futhark_f32_1d *herm_weights,
                      const struct futhark_i32_1d *sob_dir_vct, ic_dirvec *dv, ic_move **moves,
                      int32_t **acc) {
  *moves = NULL;
  *acc = NULL;
  memset(c, 0, sizeof(*c));
  if (check_inputs(ctx, swaptions, herm_coefs, herm_weights, sob_dir_vct, dv)) return 1;
  if (pop < 1) {
    lvc_set_error(ctx, "main: population %d: the winner indexes genomes[0] of an empty "
                  "population (:752-758)", pop);
    return 1;
  }
  if (pop > (1 << 24) || conv > (1 << 24)) {
    lvc_set_error(ctx, "main: population %d or %d iterations (at most 2^24 each)", pop, conv);
    return 1;
  }
  const int32_t n = conv > 0 ? conv : 0;
  /* Sobol offsets grow by at most 10 pop + 1 per iteration */
  if ((int64_t)(10 * (int64_t)pop + 2) * (n + 1) > INT32_MAX) {
    lvc_set_error(ctx, "main: population %d x %d iterations overflows the i32 Sobol offsets",
                  pop, conv);
    return 1;
  }
  *moves = malloc((size_t)(n > 0 ? n : 1) * sizeof(ic_move));
  *acc = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  if (!*moves || !*acc) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  int has_de = 0;
  move_schedule(pop, n, dv, *moves, *acc, &has_de);
  if (pop == 1 && has_de) {
    lvc_set_error(ctx, "main: a differential-evolution move with one genome indexes "
                  "genomes[-1] (:705-712)");
    return 1;
  }
  const int64_t nh = lvc_array_len_f32(herm_coefs);
  return call_prepare(ctx, c, &swaptions->a, pop, (int32_t)nh) || alloc_state(ctx, c);
}

static void end_call(struct futhark_context *ctx, ic_call *c, ic_move *moves, int32_t *acc) {
  call_release(ctx, c);
  free(moves);
  free(acc);
}

static struct futhark_f32_2d *new_summary(struct futhark_context *ctx, int32_t nswap) {
  struct futhark_f32_2d *arr = calloc(1, sizeof(*arr));
  if (!arr) {
    lvc_set_error(ctx, "host allocation failed");
    return NULL;
  }
  arr->a.shape[0] = nswap;
  arr->a.shape[1] = 3;
  arr->a.shape[2] = 1;
  arr->a.bytes = (size_t)nswap * 3 * sizeof(float);
  if (nswap == 0) {
    arr->a.raw = 1;
    return arr;
  }
  if (lvc_dev_alloc(ctx, arr->a.bytes, &arr->a.mem)) {
    free(arr);
    return NULL;
  }
  return arr;
}

int futhark_entry_main(struct futhark_context *ctx, float *out0, float *out1, float *out2,
                       float *out3, float *out4, float *out5, struct futhark_f32_2d **out6,
                       const int32_t pop, const int32_t mcmc_conv, const int32_t l,
                       const struct futhark_f32_2d *swaptions, const int32_t ll,
                       const struct futhark_f32_1d *herm_coefs,
                       const struct futhark_f32_1d *herm_weights, const int32_t lll,
                       const struct futhark_i32_1d *sob_dir_vct) {
  (void)l;
  (void)ll;
  (void)lll; /* unused by the reference (:929-931) */
  if (!out0 || !out1 || !out2 || !out3 || !out4 || !out5 || !out6)
    return 1;
  *out6 = NULL;
  if (!lvc_ready(ctx))
    return 1;
  pthread_mutex_lock(&ctx->lock);
  ic_call c;
  ic_dirvec dv;
  ic_move *moves = NULL;
  int32_t *acc = NULL;
  struct futhark_f32_2d *summary = NULL;
  float res[8];
  int rc = begin_call(ctx, &c, pop, mcmc_conv, swaptions, herm_coefs, herm_weights, sob_dir_vct,
                      &dv, &moves, &acc);
  if (!rc) {
    summary = new_summary(ctx, c.nswap);
    rc = !summary;
  }
  if (!rc)
    rc = run_calibration(ctx, &c, mcmc_conv, &dv, moves, acc,
                         lvc_array_mem_f32(herm_coefs), lvc_array_mem_f32(herm_weights), NULL,
                         summary->a.mem);
  if (!rc)
    rc = lvc_check_hip(ctx, hipMemcpyAsync(res, c.state + 12 * (size_t)c.pop, 8 * sizeof(float),
                                           hipMemcpyDeviceToHost, ctx->stream),
                       "hipMemcpyAsync") ||
         lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (!rc) lvc_prof_collect(ctx);
  end_call(ctx, &c, moves, acc);
  if (rc && summary) {
    if (!summary->a.raw) lvc_dev_free(ctx, summary->a.mem, summary->a.bytes);
    free(summary);
    summary = NULL;
  }
  pthread_mutex_unlock(&ctx->lock);
  if (rc)
    return 1;
  *out0 = res[0];
  *out1 = res[1];
  *out2 = res[2];
  *out3 = res[3];
  *out4 = res[4];
  *out5 = res[5];
  *out6 = summary;
  return 0;
}

int interestcalib_trace(struct futhark_context *ctx, const int32_t pop, const int32_t mcmc_conv,
                        const struct futhark_f32_2d *swaptions,
                        const struct futhark_f32_1d *herm_coefs,
                        const struct futhark_f32_1d *herm_weights,
                        const struct futhark_i32_1d *sob_dir_vct, float *genomes, float *logliks,
                        float *props, float *fb, float *newl, float *out, float *summary) {
  if (!genomes || !logliks || !props || !fb || !newl || !out || !summary) {
    if (ctx) lvc_set_error(ctx, "interestcalib_trace: NULL output");
    return 1;
  }
  if (!lvc_ready(ctx))
    return 1;
  pthread_mutex_lock(&ctx->lock);
  ic_call c;
  ic_dirvec dv;
  ic_move *moves = NULL;
  int32_t *acc = NULL;
  float *tr = NULL, *sm = NULL;
  const size_t conv = mcmc_conv > 0 ? (size_t)mcmc_conv : 0;
  const size_t p = pop > 0 ? (size_t)pop : 0;
  const size_t g_n = (conv + 1) * p * 5, l_n = (conv + 1) * p, pr_n = conv * p * 5, f_n = conv * p;
  const size_t tr_bytes = (g_n + l_n + pr_n + 2 * f_n + 1) * sizeof(float);
  size_t sm_bytes = 0;
  int rc = begin_call(ctx, &c, pop, mcmc_conv, swaptions, herm_coefs, herm_weights, sob_dir_vct,
                      &dv, &moves, &acc);
  if (!rc) {
    sm_bytes = ((size_t)c.nswap * 3 + 1) * sizeof(float);
    rc = lvc_dev_alloc(ctx, tr_bytes, (void **)&tr) || lvc_dev_alloc(ctx, sm_bytes, (void **)&sm);
  }
  if (!rc) {
    ic_trace t = {tr, tr + g_n, tr + g_n + l_n, tr + g_n + l_n + pr_n,
                  tr + g_n + l_n + pr_n + f_n};
    rc = run_calibration(ctx, &c, mcmc_conv, &dv, moves, acc, lvc_array_mem_f32(herm_coefs),
                         lvc_array_mem_f32(herm_weights), &t, sm);
    struct {
      void *dst;
      const float *src;
      size_t n;
    } cp[] = {{genomes, t.genomes, g_n}, {logliks, t.logliks, l_n}, {props, t.props, pr_n},
              {fb, t.fb, f_n}, {newl, t.newl, f_n}, {out, c.state + 12 * p, 6},
              {summary, sm, (size_t)c.nswap * 3}};
    for (size_t i = 0; !rc && i < sizeof(cp) / sizeof(cp[0]); i++)
      if (cp[i].n)
        rc = lvc_check_hip(ctx, hipMemcpyAsync(cp[i].dst, cp[i].src, cp[i].n * sizeof(float),
                                               hipMemcpyDeviceToHost, ctx->stream),
                           "hipMemcpyAsync");
    if (!rc)
      rc = lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  }
  if (tr) lvc_dev_free(ctx, tr, tr_bytes);
  if (sm) lvc_dev_free(ctx, sm, sm_bytes);
  end_call(ctx, &c, moves, acc);
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}

int interestcalib_eval(struct futhark_context *ctx, const int32_t ngenomes, const float *genomes,
                       const struct futhark_f32_2d *swaptions,
                       const struct futhark_f32_1d *herm_coefs,
                       const struct futhark_f32_1d *herm_weights, float *quotes, float *prices,
                       float *logliks) {
  if (!genomes || !quotes || !prices || !logliks) {
    if (ctx) lvc_set_error(ctx, "interestcalib_eval: NULL argument");
    return 1;
  }
  if (!lvc_ready(ctx))
    return 1;
  if (!swaptions || !herm_coefs || !herm_weights) {
    lvc_set_error(ctx, "interestcalib_eval: NULL input array");
    return 1;
  }
  if (ngenomes < 1 || ngenomes > (1 << 24)) {
    lvc_set_error(ctx, "interestcalib_eval: %d genomes (1 .. 2^24)", ngenomes);
    return 1;
  }
  const int64_t nh = lvc_array_len_f32(herm_coefs);
  if (lvc_array_len_f32(herm_weights) != nh || nh > INT32_MAX) {
    lvc_set_error(ctx, "interestcalib_eval: %lld Hermite coefficients and %lld weights",
                  (long long)nh, (long long)lvc_array_len_f32(herm_weights));
    return 1;
  }
  pthread_mutex_lock(&ctx->lock);
  ic_call c;
  float *pq = NULL;
  const size_t np = (size_t)ngenomes;
  size_t pq_bytes = 0;
  int rc = call_prepare(ctx, &c, &swaptions->a, ngenomes, (int32_t)nh) || alloc_state(ctx, &c);
  const size_t ns = (size_t)c.nswap;
  if (!rc) {
    pq_bytes = (2 * np * ns + 1) * sizeof(float);
    rc = lvc_dev_alloc(ctx, pq_bytes, (void **)&pq) ||
         lvc_upload(ctx, c.state, genomes, np * 5 * sizeof(float));
  }
  if (!rc) {
    void *st = (void *)ctx->stream;
    const float *hx = lvc_array_mem_f32(herm_coefs), *hw = lvc_array_mem_f32(herm_weights);
    float *lg = c.state + 10 * np;
    const ic_move none = {0, 0, 0, 0};
    ic_dirvec dv;
    memset(&dv, 0, sizeof(dv));
    int e = 0;
    if (ns > 0) {
      e = ic_launch_eval(st, c.sw, c.per, &c.pop_lists, c.state, ngenomes, hx, hw, c.nherm,
                         c.nswap, IC_OUT_PRICES, pq, pq + np * ns);
      if (!e)
        e = ic_launch_eval(st, c.sw, c.per, &c.pop_lists, c.state, ngenomes, hx, hw, c.nherm,
                           c.nswap, IC_OUT_TERMS, c.terms, NULL);
    }
    if (!e)
      e = ic_launch_step(st, ngenomes, c.nswap, c.terms, c.state, c.state + 5 * np, lg,
                         lg + np, 1, 0, &none, &dv, -1, 0, NULL);
    rc = lvc_check_hip(ctx, e, "InterestCalib kernels");
    if (!rc && ns > 0)
      rc = lvc_check_hip(ctx, hipMemcpyAsync(quotes, pq, np * ns * sizeof(float),
                                             hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpyAsync") ||
           lvc_check_hip(ctx, hipMemcpyAsync(prices, pq + np * ns, np * ns * sizeof(float),
                                             hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpyAsync");
    if (!rc)
      rc = lvc_check_hip(ctx, hipMemcpyAsync(logliks, lg, np * sizeof(float),
                                             hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpyAsync") ||
           lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  }
  if (pq) lvc_dev_free(ctx, pq, pq_bytes);
  call_release(ctx, &c);
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}
