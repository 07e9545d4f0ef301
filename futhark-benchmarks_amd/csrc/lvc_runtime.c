/*
 * Configuration, context, device-memory pool, staged uploads, arrays, errors
 * and profiling of the LocVolCalib library: the runtime half of the C API
 * that `futhark <backend> --library` generates (exercised by the reference's
 * lys frontends: accelerate/julia/lib/github.com/diku-dk/lys/shared.c:15-49,
 * shared.h:20-28, console/liblys.c:243-251, console/main.c:176-180).
 */
#define _GNU_SOURCE
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lvc_runtime.h"

/* ---- errors ------------------------------------------------------------- */

void lvc_set_error(struct futhark_context *ctx, const char *fmt, ...) {
  if (!ctx || ctx->error)
    return; /* keep the first error until it is read */
  va_list ap;
  va_start(ap, fmt);
  char *msg = NULL;
  if (vasprintf(&msg, fmt, ap) < 0)
    msg = NULL;
  va_end(ap);
  ctx->error = msg;
  if (ctx->cfg && ctx->cfg->logging && msg)
    fprintf(ctx->log ? ctx->log : stderr, "locvolcalib: %s\n", msg);
}

int lvc_check_hip(struct futhark_context *ctx, int err, const char *what) {
  if (err == 0)
    return 0;
  lvc_set_error(ctx, "%s failed: %s (%d)", what,
                hipGetErrorString((hipError_t)err), err);
  return 1;
}

int lvc_ready(struct futhark_context *ctx) {
  if (!ctx)
    return 0;
  if (!ctx->ok) {
    lvc_set_error(ctx, "no usable HIP device: the context could not be created");
    return 0;
  }
  return 1;
}

/* ---- configuration ------------------------------------------------------ */

/* The tuning parameters belong to the program (lvc_prog, defined beside
 * its entry points). */
int futhark_get_tuning_param_count(void) { return lvc_prog.ntuning; }
const char *futhark_get_tuning_param_name(int i) {
  return (i >= 0 && i < lvc_prog.ntuning) ? lvc_prog.names[i] : NULL;
}
const char *futhark_get_tuning_param_class(int i) {
  return (i >= 0 && i < lvc_prog.ntuning) ? lvc_prog.classes[i] : NULL;
}

struct futhark_context_config *futhark_context_config_new(void) {
  struct futhark_context_config *cfg = calloc(1, sizeof(*cfg));
  if (!cfg)
    return NULL;
  cfg->device = 0;
  cfg->precision = LVC_FP32_REFINE;
  cfg->graph = 0;
  cfg->xshare = 0;
  cfg->suff_intra_par_0 = 32; /* finpar/LocVolCalib.fut.tuning:1-2 */
  cfg->suff_intra_par_1 = 32;
  return cfg;
}

void futhark_context_config_free(struct futhark_context_config *cfg) {
  if (!cfg)
    return;
  free(cfg->cache_file);
  free(cfg);
}

void futhark_context_config_set_device(struct futhark_context_config *cfg,
                                       const char *s) {
  if (!cfg || !s)
    return;
  const char *p = s;
  if (strncmp(p, "hip:", 4) == 0)
    p += 4;
  if (*p == '#')
    p++;
  cfg->device = atoi(p);
}

void futhark_context_config_set_cache_file(struct futhark_context_config *cfg,
                                           const char *f) {
  if (!cfg)
    return;
  free(cfg->cache_file);
  cfg->cache_file = f ? strdup(f) : NULL;
}

int futhark_context_config_set_tuning_param(struct futhark_context_config *cfg,
                                            const char *name, size_t v) {
  if (!cfg || !name || !lvc_prog.set_param)
    return 1;
  return lvc_prog.set_param(cfg, name, v);
}

void futhark_context_config_set_debugging(struct futhark_context_config *cfg,
                                          int flag) {
  if (cfg)
    cfg->debugging = flag;
}
void futhark_context_config_set_profiling(struct futhark_context_config *cfg,
                                          int flag) {
  if (cfg)
    cfg->profiling = flag;
}
void futhark_context_config_set_logging(struct futhark_context_config *cfg,
                                        int flag) {
  if (cfg)
    cfg->logging = flag;
}

/* ---- context ------------------------------------------------------------ */

struct futhark_context *futhark_context_new(struct futhark_context_config *cfg) {
  struct futhark_context *ctx = calloc(1, sizeof(*ctx));
  if (!ctx)
    return NULL;
  ctx->cfg = cfg;
  ctx->device = cfg ? cfg->device : 0;
  pthread_mutex_init(&ctx->lock, NULL);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    lvc_set_error(ctx, "no HIP device available (%s)",
                  hipGetErrorString(e != hipSuccess ? e : hipErrorNoDevice));
    return ctx;
  }
  if (ctx->device < 0 || ctx->device >= n) {
    lvc_set_error(ctx, "device %d out of range (%d devices)", ctx->device, n);
    return ctx;
  }
  if (lvc_check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice"))
    return ctx;
  /* default-priority, non-blocking stream: no implicit sync with stream 0 */
  if (lvc_check_hip(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking),
                    "hipStreamCreate"))
    return ctx;
  if (lvc_check_hip(ctx, hipEventCreateWithFlags(&ctx->stage_done,
                                                 hipEventDisableTiming),
                    "hipEventCreate"))
    return ctx;
  if (hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount,
                            ctx->device) != hipSuccess || ctx->num_cus <= 0)
    ctx->num_cus = 256;
  ctx->ok = 1;
  if (lvc_prog.context_ready)
    lvc_prog.context_ready(ctx);
  return ctx;
}

void lvc_graph_release(struct futhark_context *ctx) {
  if (ctx->graph_exec)
    (void)hipGraphExecDestroy((hipGraphExec_t)ctx->graph_exec);
  ctx->graph_exec = NULL;
  ctx->graph_key = 0;
}

static void pool_release_all(struct futhark_context *ctx) {
  lvc_graph_release(ctx);
  for (int i = 0; i < ctx->npool; i++)
    (void)hipFree(ctx->pool[i].ptr);
  ctx->npool = 0;
  ctx->pool_bytes = 0;
}

void futhark_context_free(struct futhark_context *ctx) {
  if (!ctx)
    return;
  if (ctx->ok) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pool_release_all(ctx);
    if (ctx->stage)
      (void)hipHostFree(ctx->stage);
    for (int i = 0; i < ctx->npend; i++) {
      (void)hipEventDestroy(ctx->pend[i].start);
      (void)hipEventDestroy(ctx->pend[i].end);
    }
    for (int i = 0; i < ctx->nevents; i++)
      (void)hipEventDestroy(ctx->event_pool[i]);
    (void)hipEventDestroy(ctx->stage_done);
    (void)hipStreamDestroy(ctx->stream);
  }
  free(ctx->pool);
  free(ctx->pend);
  free(ctx->event_pool);
  free(ctx->error);
  pthread_mutex_destroy(&ctx->lock);
  free(ctx);
}

int futhark_context_sync(struct futhark_context *ctx) {
  if (!lvc_ready(ctx))
    return 1;
  pthread_mutex_lock(&ctx->lock);
  int rc = lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),
                         "hipStreamSynchronize");
  if (!rc)
    lvc_prof_collect(ctx);
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}

char *futhark_context_get_error(struct futhark_context *ctx) {
  if (!ctx)
    return NULL;
  char *e = ctx->error;
  ctx->error = NULL;
  return e;
}

void futhark_context_set_logging_file(struct futhark_context *ctx, FILE *f) {
  if (ctx)
    ctx->log = f;
}

void futhark_context_pause_profiling(struct futhark_context *ctx) {
  if (ctx)
    ctx->prof_paused = 1;
}
void futhark_context_unpause_profiling(struct futhark_context *ctx) {
  if (ctx)
    ctx->prof_paused = 0;
}

int futhark_context_clear_caches(struct futhark_context *ctx) {
  if (!lvc_ready(ctx))
    return 1;
  pthread_mutex_lock(&ctx->lock);
  int rc = lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),
                         "hipStreamSynchronize");
  pool_release_all(ctx);
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}

/* ---- device memory pool ------------------------------------------------- */

int lvc_dev_alloc(struct futhark_context *ctx, size_t bytes, void **ptr) {
  *ptr = NULL;
  if (bytes == 0)
    bytes = 256;
  /* best fit among cached blocks no more than twice the request */
  int best = -1;
  for (int i = 0; i < ctx->npool; i++) {
    size_t s = ctx->pool[i].size;
    if (s >= bytes && s <= 2 * bytes &&
        (best < 0 || s < ctx->pool[best].size))
      best = i;
  }
  if (best >= 0) {
    *ptr = ctx->pool[best].ptr;
    ctx->pool_bytes -= ctx->pool[best].size;
    ctx->pool[best] = ctx->pool[--ctx->npool];
    return 0;
  }
  hipError_t e = hipMalloc(ptr, bytes);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    /* give cached blocks back and retry once */
    (void)hipStreamSynchronize(ctx->stream);
    pool_release_all(ctx);
    e = hipMalloc(ptr, bytes);
  }
  if (e != hipSuccess) {
    *ptr = NULL;
    lvc_set_error(ctx, "device allocation of %zu bytes failed: %s", bytes,
                  hipGetErrorString(e));
    return 1;
  }
  return 0;
}

/* `bytes` must be the size the block was requested with; the pool keeps the
 * block for stream-ordered reuse (every kernel runs on ctx->stream). */
void lvc_dev_free(struct futhark_context *ctx, void *ptr, size_t bytes) {
  if (!ptr)
    return;
  if (bytes == 0)
    bytes = 256;
  if (ctx->npool == ctx->cappool) {
    int nc = ctx->cappool ? 2 * ctx->cappool : 32;
    lvc_block *np = realloc(ctx->pool, (size_t)nc * sizeof(*np));
    if (!np) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ptr);
      return;
    }
    ctx->pool = np;
    ctx->cappool = nc;
  }
  ctx->pool[ctx->npool].ptr = ptr;
  ctx->pool[ctx->npool].size = bytes;
  ctx->npool++;
  ctx->pool_bytes += bytes;
}

/* ---- staged uploads ----------------------------------------------------- */

int lvc_upload(struct futhark_context *ctx, void *dst, const void *src,
               size_t bytes) {
  if (bytes == 0)
    return 0;
  /* The staging buffer is a bump arena: uploads take consecutive regions and
   * only a wrap-around waits, for the copies enqueued so far (they run in
   * stream order, so the event after the latest one covers them all). */
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (ctx->stage_used + need > ctx->stage_size) {
    if (ctx->stage_pending) {
      if (lvc_check_hip(ctx, hipEventSynchronize(ctx->stage_done),
                        "hipEventSynchronize"))
        return 1;
      ctx->stage_pending = 0;
    }
    ctx->stage_used = 0;
    if (need > ctx->stage_size) {
      if (ctx->stage)
        (void)hipHostFree(ctx->stage);
      ctx->stage = NULL;
      size_t sz = need < (4u << 20) ? (4u << 20) : need;
      if (lvc_check_hip(ctx, hipHostMalloc(&ctx->stage, sz, hipHostMallocDefault),
                        "hipHostMalloc")) {
        ctx->stage_size = 0;
        return 1;
      }
      ctx->stage_size = sz;
    }
  }
  char *region = (char *)ctx->stage + ctx->stage_used;
  memcpy(region, src, bytes);
  if (lvc_check_hip(ctx, hipMemcpyAsync(dst, region, bytes,
                                        hipMemcpyHostToDevice, ctx->stream),
                    "hipMemcpyAsync"))
    return 1;
  if (lvc_check_hip(ctx, hipEventRecord(ctx->stage_done, ctx->stream),
                    "hipEventRecord"))
    return 1;
  ctx->stage_used += need;
  ctx->stage_pending = 1;
  return 0;
}

/* ---- profiling ---------------------------------------------------------- */

static hipEvent_t event_get(struct futhark_context *ctx) {
  if (ctx->nevents > 0)
    return ctx->event_pool[--ctx->nevents];
  hipEvent_t e = NULL;
  if (hipEventCreate(&e) != hipSuccess)
    return NULL;
  return e;
}

static void event_put(struct futhark_context *ctx, hipEvent_t e) {
  if (ctx->nevents == ctx->capevents) {
    int nc = ctx->capevents ? 2 * ctx->capevents : 64;
    hipEvent_t *np = realloc(ctx->event_pool, (size_t)nc * sizeof(*np));
    if (!np) {
      (void)hipEventDestroy(e);
      return;
    }
    ctx->event_pool = np;
    ctx->capevents = nc;
  }
  ctx->event_pool[ctx->nevents++] = e;
}

int lvc_prof_begin(struct futhark_context *ctx, const char *name) {
  if (!ctx->cfg || !ctx->cfg->profiling || ctx->prof_paused)
    return -1;
  int which = -1;
  for (int i = 0; i < ctx->nprof; i++)
    if (strcmp(ctx->prof[i].name, name) == 0)
      which = i;
  if (which < 0) {
    if (ctx->nprof == LVC_MAX_PROF)
      return -1;
    which = ctx->nprof++;
    ctx->prof[which].name = name;
    ctx->prof[which].count = 0;
    ctx->prof[which].total_us = 0;
  }
  if (ctx->npend == ctx->cappend) {
    int nc = ctx->cappend ? 2 * ctx->cappend : 256;
    lvc_prof_pending *np = realloc(ctx->pend, (size_t)nc * sizeof(*np));
    if (!np)
      return -1;
    ctx->pend = np;
    ctx->cappend = nc;
  }
  hipEvent_t s = event_get(ctx), e = event_get(ctx);
  if (!s || !e)
    return -1;
  (void)hipEventRecord(s, ctx->stream);
  int tok = ctx->npend++;
  ctx->pend[tok].which = which;
  ctx->pend[tok].start = s;
  ctx->pend[tok].end = e;
  return tok;
}

void lvc_prof_end(struct futhark_context *ctx, int token) {
  if (token < 0)
    return;
  (void)hipEventRecord(ctx->pend[token].end, ctx->stream);
}

/* Called with the stream synchronised. */
void lvc_prof_collect(struct futhark_context *ctx) {
  for (int i = 0; i < ctx->npend; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->pend[i].start, ctx->pend[i].end) ==
        hipSuccess) {
      ctx->prof[ctx->pend[i].which].count++;
      ctx->prof[ctx->pend[i].which].total_us += 1000.0 * ms;
    }
    event_put(ctx, ctx->pend[i].start);
    event_put(ctx, ctx->pend[i].end);
  }
  ctx->npend = 0;
}

char *futhark_context_report(struct futhark_context *ctx) {
  if (!ctx)
    return NULL;
  if (ctx->ok) {
    pthread_mutex_lock(&ctx->lock);
    if (hipStreamSynchronize(ctx->stream) == hipSuccess)
      lvc_prof_collect(ctx);
    pthread_mutex_unlock(&ctx->lock);
  }
  size_t cap = 320 + 128 * (size_t)ctx->nprof;
  char *out = malloc(cap);
  if (!out)
    return NULL;
  size_t len = (size_t)snprintf(out, cap, "{");
  for (int i = 0; i < ctx->nprof; i++)
    len += (size_t)snprintf(out + len, cap - len,
                            "%s\"%s\": {\"count\": %lld, \"total_us\": %.3f}",
                            i ? ", " : "", ctx->prof[i].name,
                            (long long)ctx->prof[i].count, ctx->prof[i].total_us);
  if (ctx->dbg_index_valid)
    len += (size_t)snprintf(out + len, cap - len, "%s\"price_index\": [%d, %d]",
                            ctx->nprof ? ", " : "", ctx->dbg_index[0], ctx->dbg_index[1]);
  snprintf(out + len, cap - len, "}");
  return out;
}

/* ---- arrays ------------------------------------------------------------- */

struct lvc_array {
  void *mem;
  int64_t shape[1];
  size_t bytes;
  int raw; /* wraps caller memory: not freed */
};
struct futhark_f32_1d { struct lvc_array a; };
struct futhark_f64_1d { struct lvc_array a; };
struct futhark_i32_1d { struct lvc_array a; };

static int array_new(struct futhark_context *ctx, struct lvc_array *a,
                     const void *data, int64_t n, size_t elem) {
  if (n < 0) {
    lvc_set_error(ctx, "negative array size %lld", (long long)n);
    return 1;
  }
  a->shape[0] = n;
  a->bytes = (size_t)n * elem;
  a->raw = 0;
  if (lvc_dev_alloc(ctx, a->bytes, &a->mem))
    return 1;
  if (data && a->bytes &&
      lvc_check_hip(ctx, hipMemcpyAsync(a->mem, data, a->bytes,
                                        hipMemcpyHostToDevice, ctx->stream),
                    "hipMemcpyAsync") == 0)
    return lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),
                         "hipStreamSynchronize");
  return data && a->bytes ? 1 : 0;
}

static int array_values(struct futhark_context *ctx, struct lvc_array *a,
                        void *data) {
  if (a->bytes == 0)
    return 0;
  if (lvc_check_hip(ctx, hipMemcpyAsync(data, a->mem, a->bytes,
                                        hipMemcpyDeviceToHost, ctx->stream),
                    "hipMemcpyAsync"))
    return 1;
  int rc = lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),
                         "hipStreamSynchronize");
  if (!rc)
    lvc_prof_collect(ctx);
  return rc;
}

#define DEFINE_ARRAY(T, name, ctype)                                           \
  struct futhark_##name *futhark_new_##name(struct futhark_context *ctx,       \
                                            const ctype *data, int64_t dim0) { \
    if (!lvc_ready(ctx))                                                       \
      return NULL;                                                             \
    struct futhark_##name *arr = calloc(1, sizeof(*arr));                      \
    if (!arr)                                                                  \
      return NULL;                                                             \
    pthread_mutex_lock(&ctx->lock);                                            \
    int rc = array_new(ctx, &arr->a, data, dim0, sizeof(ctype));               \
    pthread_mutex_unlock(&ctx->lock);                                          \
    if (rc) {                                                                  \
      if (arr->a.mem)                                                          \
        lvc_dev_free(ctx, arr->a.mem, arr->a.bytes);                           \
      free(arr);                                                               \
      return NULL;                                                             \
    }                                                                          \
    return arr;                                                                \
  }                                                                            \
  int futhark_free_##name(struct futhark_context *ctx,                         \
                          struct futhark_##name *arr) {                        \
    if (!ctx || !arr)                                                          \
      return 0;                                                                \
    pthread_mutex_lock(&ctx->lock);                                            \
    if (!arr->a.raw && ctx->ok)                                                \
      lvc_dev_free(ctx, arr->a.mem, arr->a.bytes);                             \
    pthread_mutex_unlock(&ctx->lock);                                          \
    free(arr);                                                                 \
    return 0;                                                                  \
  }                                                                            \
  int futhark_values_##name(struct futhark_context *ctx,                       \
                            struct futhark_##name *arr, ctype *data) {         \
    if (!lvc_ready(ctx))                                                       \
      return 1;                                                                \
    if (!arr) {                                                                \
      lvc_set_error(ctx, "NULL array");                                        \
      return 1;                                                                \
    }                                                                          \
    pthread_mutex_lock(&ctx->lock);                                            \
    int rc = array_values(ctx, &arr->a, data);                                 \
    pthread_mutex_unlock(&ctx->lock);                                          \
    return rc;                                                                 \
  }                                                                            \
  const int64_t *futhark_shape_##name(struct futhark_context *ctx,             \
                                      struct futhark_##name *arr) {            \
    (void)ctx;                                                                 \
    return arr ? arr->a.shape : NULL;                                          \
  }

DEFINE_ARRAY(float, f32_1d, float)
DEFINE_ARRAY(double, f64_1d, double)
DEFINE_ARRAY(int32_t, i32_1d, int32_t)

#define DEFINE_RAW(name, ctype)                                                \
  struct futhark_##name *futhark_new_raw_##name(struct futhark_context *ctx,   \
                                                void *data, int64_t dim0) {    \
    if (!lvc_ready(ctx))                                                       \
      return NULL;                                                             \
    if (dim0 < 0) {                                                            \
      lvc_set_error(ctx, "negative array size %lld", (long long)dim0);         \
      return NULL;                                                             \
    }                                                                          \
    struct futhark_##name *arr = calloc(1, sizeof(*arr));                      \
    if (!arr)                                                                  \
      return NULL;                                                             \
    arr->a.mem = data;                                                         \
    arr->a.shape[0] = dim0;                                                    \
    arr->a.bytes = (size_t)dim0 * sizeof(ctype);                               \
    arr->a.raw = 1;                                                            \
    return arr;                                                                \
  }                                                                            \
  void *futhark_values_raw_##name(struct futhark_context *ctx,                 \
                                  struct futhark_##name *arr) {                \
    (void)ctx;                                                                 \
    return arr ? arr->a.mem : NULL;                                            \
  }                                                                            \
  int futhark_index_##name(struct futhark_context *ctx, ctype *out,            \
                           struct futhark_##name *arr, int64_t i0) {           \
    if (!lvc_ready(ctx))                                                       \
      return 1;                                                                \
    if (!arr || i0 < 0 || i0 >= arr->a.shape[0]) {                             \
      lvc_set_error(ctx, "index %lld out of bounds", (long long)i0);           \
      return 1;                                                                \
    }                                                                          \
    pthread_mutex_lock(&ctx->lock);                                            \
    int rc = lvc_check_hip(                                                    \
        ctx,                                                                   \
        hipMemcpyAsync(out, (char *)arr->a.mem + (size_t)i0 * sizeof(ctype),   \
                       sizeof(ctype), hipMemcpyDeviceToHost, ctx->stream),     \
        "hipMemcpyAsync");                                                     \
    if (!rc)                                                                   \
      rc = lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream),               \
                         "hipStreamSynchronize");                              \
    pthread_mutex_unlock(&ctx->lock);                                          \
    return rc;                                                                 \
  }

DEFINE_RAW(f32_1d, float)
DEFINE_RAW(f64_1d, double)

/* Internal accessors for the engine. */
void *lvc_array_mem_f32(const struct futhark_f32_1d *a) { return a->a.mem; }
void *lvc_array_mem_f64(const struct futhark_f64_1d *a) { return a->a.mem; }
int64_t lvc_array_len_f32(const struct futhark_f32_1d *a) { return a->a.shape[0]; }
void *lvc_array_mem_i32(const struct futhark_i32_1d *a) { return a->a.mem; }
int64_t lvc_array_len_f64(const struct futhark_f64_1d *a) { return a->a.shape[0]; }
int64_t lvc_array_len_i32(const struct futhark_i32_1d *a) { return a->a.shape[0]; }

struct futhark_f32_1d *lvc_array_adopt_f32(void *mem, int64_t n) {
  struct futhark_f32_1d *arr = calloc(1, sizeof(*arr));
  if (!arr)
    return NULL;
  arr->a.mem = mem;
  arr->a.shape[0] = n;
  arr->a.bytes = (size_t)n * sizeof(float);
  return arr;
}
struct futhark_f64_1d *lvc_array_adopt_f64(void *mem, int64_t n) {
  struct futhark_f64_1d *arr = calloc(1, sizeof(*arr));
  if (!arr)
    return NULL;
  arr->a.mem = mem;
  arr->a.shape[0] = n;
  arr->a.bytes = (size_t)n * sizeof(double);
  return arr;
}
