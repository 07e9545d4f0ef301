/* Host-side runtime shared by the C sources of the library (not exported). */
#ifndef LVC_RUNTIME_H
#define LVC_RUNTIME_H

#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include <hip/hip_runtime_api.h>

#include "../../include/futhark_context.h"
#include "lvc_internal.h"

#define LVC_MAX_TUNED 64
#define LVC_XSHARE_MAX 16
#define LVC_MAX_PROF 16

struct futhark_context_config {
  int device;
  char *cache_file;
  int debugging, profiling, logging;
  int precision; /* enum lvc_precision for futhark_entry_main */
  int seg_x, seg_y; /* 0 = tuned */
  int autotune;
  int graph; /* replay the step launches from a HIP graph */
  int xshare; /* problems per x tile sharing a factorisation, 1..LVC_XSHARE_MAX;
                0 = the default of the precision */
  size_t suff_intra_par_0, suff_intra_par_1;
};

typedef struct {
  int precision, nx, ny;
  int seg_x, seg_y;
} lvc_tuned;

typedef struct {
  void *ptr;
  size_t size;
} lvc_block;

typedef struct {
  const char *name;
  int64_t count;
  double total_us;
} lvc_prof_entry;

typedef struct {
  int which; /* index into prof[] */
  hipEvent_t start, end;
} lvc_prof_pending;

struct futhark_context {
  struct futhark_context_config *cfg;
  int device;
  int ok; /* device opened */
  hipStream_t stream;
  pthread_mutex_t lock;
  char *error;
  FILE *log;
  /* device memory pool: blocks returned by lvc_dev_free are kept for reuse
   * (stream-ordered on ctx->stream) until clear_caches / context_free */
  lvc_block *pool;
  int npool, cappool;
  size_t pool_bytes;
  /* pinned staging for host -> device uploads */
  void *stage;
  size_t stage_size, stage_used;
  hipEvent_t stage_done;
  int stage_pending;
  /* profiling */
  int prof_paused;
  lvc_prof_entry prof[LVC_MAX_PROF];
  int nprof;
  lvc_prof_pending *pend;
  int npend, cappend;
  hipEvent_t *event_pool;
  int nevents, capevents;
  /* tuning table */
  lvc_tuned tuned[LVC_MAX_TUNED];
  int num_cus; /* compute units of the device */
  int ntuned;
  /* debugging: the price index (myYindex, myXindex) the last entry call's
   * gather read, from the device descriptor of its first problem */
  int dbg_index_valid;
  int32_t dbg_index[2];
  /* device address of the static scratch of the allocation-free entry
   * (looked up when the context opens its device), or NULL with the
   * hipError_t of the failed lookup */
  void *scratch;
  int scratch_err;
  /* the instantiated HIP graph of the last step sequence replayed with
   * "lvc.graph" (a hipGraphExec_t) and the key of its launches */
  void *graph_exec;
  uint64_t graph_key;
};

/* What differs between the programs built on this runtime: the tuning
 * parameters futhark_get_tuning_param_* list, how a configuration takes
 * them (0: accepted, 1: unknown name or bad value), and a hook run once a
 * context has opened its device.  Each library defines `lvc_prog`. */
typedef struct {
  int ntuning;
  const char *const *names, *const *classes;
  int (*set_param)(struct futhark_context_config *cfg, const char *name, size_t v);
  void (*context_ready)(struct futhark_context *ctx);
} lvc_program;
extern const lvc_program lvc_prog;

/* errors: printf-style; keeps the first message until it is read */
void lvc_set_error(struct futhark_context *ctx, const char *fmt, ...);
int lvc_check_hip(struct futhark_context *ctx, int err, const char *what);
int lvc_ready(struct futhark_context *ctx);

/* drops the kept graph (its launches name pool memory) */
void lvc_graph_release(struct futhark_context *ctx);

/* device memory pool */
int lvc_dev_alloc(struct futhark_context *ctx, size_t bytes, void **ptr);
void lvc_dev_free(struct futhark_context *ctx, void *ptr, size_t bytes);

/* stream-ordered upload through pinned staging: the device copy is enqueued
 * on ctx->stream and the host data may be reused on return */
int lvc_upload(struct futhark_context *ctx, void *dst, const void *src,
               size_t bytes);

/* the 1-D arrays' device memory and length; adopt wraps device memory from
 * lvc_dev_alloc as an array that futhark_free_*_1d returns to the pool */
void *lvc_array_mem_f32(const struct futhark_f32_1d *a);
void *lvc_array_mem_f64(const struct futhark_f64_1d *a);
int64_t lvc_array_len_f32(const struct futhark_f32_1d *a);
void *lvc_array_mem_i32(const struct futhark_i32_1d *a);
int64_t lvc_array_len_f64(const struct futhark_f64_1d *a);
int64_t lvc_array_len_i32(const struct futhark_i32_1d *a);
struct futhark_f32_1d *lvc_array_adopt_f32(void *mem, int64_t n);
struct futhark_f64_1d *lvc_array_adopt_f64(void *mem, int64_t n);

/* profiling brackets around a launch (no-ops unless profiling is on) */
int lvc_prof_begin(struct futhark_context *ctx, const char *name);
void lvc_prof_end(struct futhark_context *ctx, int token);
void lvc_prof_collect(struct futhark_context *ctx);

/* tuning (lvc_tune.c) */
void lvc_tune_pick(struct futhark_context *ctx, int precision, int nx, int ny,
                   int *seg_x, int *seg_y);
/* the tuned choice without timing anything: the table (cache file, earlier
 * tuning) or the built-in choice */
void lvc_tune_peek(struct futhark_context *ctx, int precision, int nx, int ny,
                   int *seg_x, int *seg_y);
void lvc_tune_load(struct futhark_context *ctx);

/* engine (lvc_engine.c) */
typedef struct lvc_batch {
  int n, precision, seg_x, seg_y;
  int64_t steps; /* max(nt) - 1 */
  lvc_problem *h_probs;
  lvc_problem *d_probs;
  size_t probs_bytes;
  double *d_geo;
  size_t geo_bytes;
  void *d_R, *d_Y;
  size_t state_bytes;
  lvc_tile *d_tiles; /* x tiles then y tiles, each by problem numT, longest first */
  size_t tiles_bytes;
  int nxt, nyt;
  /* tiles of the problems still running at step s: a prefix of each list */
  int *x_active, *y_active; /* [steps] */
  int graph;
  int ylds; /* LDS bytes of the y-sweep launches */
  void *d_yfac; /* y factor tables (column y sweep only), LVC_YFAC_LEN each */
  size_t yfac_bytes;
  int32_t *d_fprobs; /* problems by numT, longest first (factor workgroups) */
  int xshared;       /* some x tiles cover several problems */
  int *f_active;     /* [steps] problems still running at step s */
} lvc_batch;

int lvc_batch_build(struct futhark_context *ctx, lvc_batch *b, int n,
                    const lvc_problem *probs, int precision, int seg_x,
                    int seg_y);
int lvc_batch_run(struct futhark_context *ctx, lvc_batch *b);
int lvc_batch_gather(struct futhark_context *ctx, lvc_batch *b, void *out,
                     int out_f64);
void lvc_batch_release(struct futhark_context *ctx, lvc_batch *b);
int lvc_validate_sizes(struct futhark_context *ctx, int64_t nx, int64_t ny,
                       int64_t nt);
void lvc_choose_segs(struct futhark_context *ctx, int precision, int nx, int ny,
                     int *seg_x, int *seg_y);
int lvc_xshare(struct futhark_context *ctx, int precision);
/* LocVolCalib's context-ready hook: the tuning cache and the scratch address */
void lvc_context_ready(struct futhark_context *ctx);

#endif /* LVC_RUNTIME_H */
