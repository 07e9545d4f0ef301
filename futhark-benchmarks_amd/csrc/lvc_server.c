/*
 * Serving form: continuous batching of independent calibration requests
 * (each one LocVolCalib.fut:211-236 `value` with its own sizes and
 * parameters).  No counterpart in the reference.
 *
 * The server owns `max_slots` device slots and one arena for the resident
 * grids.  A slot holds a resident request's descriptor, geometry and y factor
 * table (fixed size, for the largest accepted grid); its two state grids
 * (R, Y) take an extent of the arena of exactly their own size, found first
 * fit, so admission and completion never move another request's data.  All
 * of it is allocated when the server is made, and the memory budget bounds
 * the total: the slot tables first, the arena from what remains (at most
 * what max_slots grids of 512 x 512 would take).  One engine tick
 * (lvc_server_step):
 *   1. admits queued requests, in arrival order, while a slot is free and the
 *      arena has an extent for the request's grids (in-order: the first
 *      request that does not fit waits for memory, and the ones behind it
 *      with it); each admission uploads its descriptor and runs the setup
 *      kernel for that slot;
 *   2. launches the x sweep and the y sweep once over every resident request
 *      (each advances one time step);
 *   3. gathers the price of every request that has just done its last step,
 *      copies those prices back and frees their slots.
 * Every kernel runs on the context's stream, so the slot reuse in step 1 is
 * ordered after the completed request's last use.  A request's arithmetic
 * depends only on its own inputs and the server's fixed segment lengths,
 * never on which other requests are resident.
 */
#include <stdlib.h>
#include <string.h>

#include "../../include/locvolcalib.h"
#include "lvc_runtime.h"

#define SLOT_MAX_N 512 /* largest numX / numY a server slot holds */
/* tiles of one request at most: x tiles ny / (rows per tile >= 4), y tiles
 * nx / (columns per tile >= 4) */
#define SLOT_MAX_TILES (2 * SLOT_MAX_N / 4)

/* a free extent of the grid arena, in elements */
typedef struct {
  int64_t off, len;
} extent;

typedef struct {
  int64_t id;
  struct lvc_request req;
} queued_req;

typedef struct {
  int state; /* 0 queued, 1 resident, 3 finishing (price copy in flight), 2 done */
  double price;
  int64_t batch; /* finishing: sequence number of its completion batch */
} result_rec;

/* Completions of one tick: their prices are copied to pinned host memory
 * behind the tick's launches and published when the copy has landed. */
#define LVC_SRV_PENDING 4
typedef struct {
  hipEvent_t done;
  double *host;  /* pinned, max_slots prices */
  int64_t *ids;  /* request ids, in gather order */
  int count;
  int64_t seq;
} completion_batch;

struct lvc_server {
  struct futhark_context *ctx;
  int max_slots, precision, seg_x, seg_y;
  int64_t budget, resident_bytes;
  size_t real;
  int64_t slot_geo;
  /* grid arena: R and Y at the same element offsets of d_R and d_Y */
  int64_t arena_elems, device_bytes;
  extent *free_ext; /* sorted by offset, disjoint, not adjacent */
  int nfree;
  int64_t *slot_len; /* arena elements held by each slot */
  /* slots */
  lvc_problem *h_slots;
  int64_t *slot_id;  /* request id, -1 when free */
  int *steps_left;
  int nresident;
  lvc_problem *d_slots;
  double *d_geo;
  void *d_R, *d_Y;
  void *d_yfac; /* y factor tables per slot (column y sweep only) */
  int32_t *h_fprobs, *d_fprobs; /* resident slots (factor workgroups) */
  int nfac;
  double *d_out;
  /* admission: the tick's new descriptors and their slots, staged on the
   * device for one setup launch */
  lvc_problem *h_adm, *d_adm;
  int32_t *h_admslot, *d_admslot;
  int nadm;
  /* completions: the tick's finished slots, one gather launch */
  int32_t *h_done, *d_done;
  completion_batch pend[LVC_SRV_PENDING];
  int64_t pend_head, pend_tail; /* batches pend_head .. pend_tail-1 in flight */
  /* tiles of the resident set */
  lvc_tile *h_tiles, *d_tiles;
  int nxt, nyt, tiles_cap, tiles_dirty, ylds;
  /* queue and results */
  queued_req *queue;
  int64_t qhead, qtail, qcap;
  result_rec *results;
  int64_t nresults, rcap;
  int64_t completed;
};

/* arena elements of a request's grids (64-element aligned extents) */
static int64_t request_elems(int nx, int ny) {
  return ((int64_t)nx * ny + 63) & ~(int64_t)63;
}

/* bytes of a request's grids R and Y in the arena */
static int64_t request_bytes(const struct lvc_server *s, int nx, int ny) {
  return 2 * request_elems(nx, ny) * (int64_t)s->real;
}

/* device bytes of the per-slot tables: descriptor, geometry, y factor
 * table, output, factor list and tiles */
static int64_t slot_bytes(const struct lvc_server *s, int col) {
  return 2 * (int64_t)sizeof(lvc_problem) + s->slot_geo * (int64_t)sizeof(double) +
         (col ? LVC_YFAC_LEN * (int64_t)s->real + (int64_t)sizeof(int32_t) : 0) +
         (int64_t)sizeof(double) + SLOT_MAX_TILES * (int64_t)sizeof(lvc_tile) +
         2 * (int64_t)sizeof(int32_t);
}

/* First-fit extent of `len` elements; -1 when none is free. */
static int64_t arena_take(struct lvc_server *s, int64_t len) {
  for (int i = 0; i < s->nfree; i++) {
    if (s->free_ext[i].len < len)
      continue;
    const int64_t off = s->free_ext[i].off;
    s->free_ext[i].off += len;
    s->free_ext[i].len -= len;
    if (s->free_ext[i].len == 0) {
      memmove(s->free_ext + i, s->free_ext + i + 1, (size_t)(s->nfree - i - 1) * sizeof(extent));
      s->nfree--;
    }
    return off;
  }
  return -1;
}

/* Returns an extent, merging it with free neighbours. */
static void arena_give(struct lvc_server *s, int64_t off, int64_t len) {
  int i = 0;
  while (i < s->nfree && s->free_ext[i].off < off)
    i++;
  const int prev = i > 0 && s->free_ext[i - 1].off + s->free_ext[i - 1].len == off;
  const int next = i < s->nfree && off + len == s->free_ext[i].off;
  if (prev && next) {
    s->free_ext[i - 1].len += len + s->free_ext[i].len;
    memmove(s->free_ext + i, s->free_ext + i + 1, (size_t)(s->nfree - i - 1) * sizeof(extent));
    s->nfree--;
  } else if (prev) {
    s->free_ext[i - 1].len += len;
  } else if (next) {
    s->free_ext[i].off = off;
    s->free_ext[i].len += len;
  } else {
    memmove(s->free_ext + i + 1, s->free_ext + i, (size_t)(s->nfree - i) * sizeof(extent));
    s->free_ext[i].off = off;
    s->free_ext[i].len = len;
    s->nfree++;
  }
}

struct lvc_server *lvc_server_new(struct futhark_context *ctx, int32_t max_slots,
                                  int64_t mem_budget_bytes, int32_t precision) {
  if (!lvc_ready(ctx))
    return NULL;
  if (max_slots <= 0 || max_slots > (1 << 20) || precision < LVC_FP32 ||
      precision > LVC_FP64 || mem_budget_bytes < 0) {
    lvc_set_error(ctx, "lvc_server_new: invalid arguments");
    return NULL;
  }
  struct lvc_server *s = calloc(1, sizeof(*s));
  if (!s)
    return NULL;
  s->ctx = ctx;
  s->max_slots = max_slots;
  s->precision = precision;
  s->budget = mem_budget_bytes;
  s->real = precision == LVC_FP64 ? sizeof(double) : sizeof(float);
  s->slot_geo = (LVC_GEO_LEN(SLOT_MAX_N, SLOT_MAX_N) + 63) & ~(int64_t)63;
  pthread_mutex_lock(&ctx->lock);
  /* segment lengths fixed for the server's life (one kernel variant for the
   * whole resident set); tuned for the largest slot grid */
  lvc_choose_segs(ctx, precision, 256, 256, &s->seg_x, &s->seg_y);
  if (lvc_lanes_for(SLOT_MAX_N, s->seg_x) > 64)
    s->seg_x = 8;
  if (!lvc_yseg_fits(SLOT_MAX_N, s->seg_y))
    s->seg_y = lvc_seg_is_col(s->seg_y) ? 64 : 8;
  /* the budget: the slot tables first, the grid arena from the rest, at most
   * what max_slots grids of the largest size take */
  const int col = lvc_seg_is_col(s->seg_y);
  const int64_t fixed = (int64_t)max_slots * slot_bytes(s, col);
  int64_t arena = (int64_t)max_slots * request_elems(SLOT_MAX_N, SLOT_MAX_N);
  if (s->budget > 0) {
    const int64_t fit = (s->budget - fixed) / (2 * (int64_t)s->real) & ~(int64_t)63;
    if (fit < request_elems(3, 3)) {
      pthread_mutex_unlock(&ctx->lock);
      lvc_set_error(ctx, "lvc_server_new: a budget of %lld bytes does not cover the %lld "
                         "bytes of the slot tables (%d slots) and a grid",
                    (long long)s->budget, (long long)fixed, max_slots);
      free(s);
      return NULL;
    }
    if (fit < arena)
      arena = fit;
  }
  s->arena_elems = arena;
  s->device_bytes = fixed + 2 * arena * (int64_t)s->real;
  s->h_slots = calloc((size_t)max_slots, sizeof(lvc_problem));
  s->slot_id = malloc((size_t)max_slots * sizeof(int64_t));
  s->steps_left = calloc((size_t)max_slots, sizeof(int));
  s->h_fprobs = calloc((size_t)max_slots, sizeof(int32_t));
  s->slot_len = calloc((size_t)max_slots, sizeof(int64_t));
  s->free_ext = malloc(((size_t)max_slots + 1) * sizeof(extent));
  s->tiles_cap = max_slots * SLOT_MAX_TILES;
  s->h_tiles = malloc((size_t)s->tiles_cap * sizeof(lvc_tile));
  s->h_adm = malloc((size_t)max_slots * sizeof(lvc_problem));
  s->h_admslot = malloc((size_t)max_slots * sizeof(int32_t));
  s->h_done = malloc((size_t)max_slots * sizeof(int32_t));
  int rc = !s->h_slots || !s->slot_id || !s->steps_left || !s->h_fprobs || !s->slot_len ||
           !s->free_ext || !s->h_tiles || !s->h_adm || !s->h_admslot || !s->h_done;
  for (int b = 0; !rc && b < LVC_SRV_PENDING; b++) {
    s->pend[b].ids = malloc((size_t)max_slots * sizeof(int64_t));
    rc = !s->pend[b].ids;
  }
  if (rc) {
    lvc_set_error(ctx, "host allocation failed");
  } else {
    s->free_ext[0].off = 0;
    s->free_ext[0].len = arena;
    s->nfree = 1;
    for (int i = 0; i < max_slots; i++) {
      s->slot_id[i] = -1;
      s->h_slots[i].k = -1;
    }
  }
  for (int b = 0; !rc && b < LVC_SRV_PENDING; b++)
    rc = lvc_check_hip(ctx, hipHostMalloc((void **)&s->pend[b].host,
                                          (size_t)max_slots * sizeof(double),
                                          hipHostMallocDefault), "hipHostMalloc") ||
         lvc_check_hip(ctx, hipEventCreateWithFlags(&s->pend[b].done, hipEventDisableTiming),
                       "hipEventCreate");
  if (!rc)
    rc = lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(lvc_problem), (void **)&s->d_slots) ||
         lvc_dev_alloc(ctx, (size_t)max_slots * s->slot_geo * sizeof(double), (void **)&s->d_geo) ||
         lvc_dev_alloc(ctx, (size_t)arena * s->real, &s->d_R) ||
         lvc_dev_alloc(ctx, (size_t)arena * s->real, &s->d_Y) ||
         lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(double), (void **)&s->d_out) ||
         lvc_dev_alloc(ctx, (size_t)s->tiles_cap * sizeof(lvc_tile), (void **)&s->d_tiles) ||
         lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(lvc_problem), (void **)&s->d_adm) ||
         lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(int32_t), (void **)&s->d_admslot) ||
         lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(int32_t), (void **)&s->d_done) ||
         (lvc_seg_is_col(s->seg_y) &&
          (lvc_dev_alloc(ctx, (size_t)max_slots * LVC_YFAC_LEN * s->real, &s->d_yfac) ||
           lvc_dev_alloc(ctx, (size_t)max_slots * sizeof(int32_t), (void **)&s->d_fprobs)));
  if (!rc)
    rc = lvc_upload(ctx, s->d_slots, s->h_slots, (size_t)max_slots * sizeof(lvc_problem));
  pthread_mutex_unlock(&ctx->lock);
  if (rc) {
    lvc_server_free(s);
    return NULL;
  }
  return s;
}

void lvc_server_free(struct lvc_server *s) {
  if (!s)
    return;
  struct futhark_context *ctx = s->ctx;
  pthread_mutex_lock(&ctx->lock);
  (void)hipStreamSynchronize(ctx->stream);
  lvc_dev_free(ctx, s->d_slots, (size_t)s->max_slots * sizeof(lvc_problem));
  lvc_dev_free(ctx, s->d_geo, (size_t)s->max_slots * s->slot_geo * sizeof(double));
  lvc_dev_free(ctx, s->d_R, (size_t)s->arena_elems * s->real);
  lvc_dev_free(ctx, s->d_Y, (size_t)s->arena_elems * s->real);
  lvc_dev_free(ctx, s->d_out, (size_t)s->max_slots * sizeof(double));
  if (s->d_yfac)
    lvc_dev_free(ctx, s->d_yfac, (size_t)s->max_slots * LVC_YFAC_LEN * s->real);
  if (s->d_fprobs)
    lvc_dev_free(ctx, s->d_fprobs, (size_t)s->max_slots * sizeof(int32_t));
  if (s->d_tiles)
    lvc_dev_free(ctx, s->d_tiles, (size_t)s->tiles_cap * sizeof(lvc_tile));
  lvc_dev_free(ctx, s->d_adm, (size_t)s->max_slots * sizeof(lvc_problem));
  lvc_dev_free(ctx, s->d_admslot, (size_t)s->max_slots * sizeof(int32_t));
  lvc_dev_free(ctx, s->d_done, (size_t)s->max_slots * sizeof(int32_t));
  for (int b = 0; b < LVC_SRV_PENDING; b++) {
    if (s->pend[b].host)
      (void)hipHostFree(s->pend[b].host);
    if (s->pend[b].done)
      (void)hipEventDestroy(s->pend[b].done);
    free(s->pend[b].ids);
  }
  pthread_mutex_unlock(&ctx->lock);
  free(s->h_slots);
  free(s->slot_id);
  free(s->steps_left);
  free(s->h_adm);
  free(s->h_admslot);
  free(s->h_done);
  free(s->h_fprobs);
  free(s->slot_len);
  free(s->free_ext);
  free(s->h_tiles);
  free(s->queue);
  free(s->results);
  free(s);
}

int64_t lvc_server_submit(struct lvc_server *s, const struct lvc_request *r) {
  if (!s || !r)
    return -1;
  struct futhark_context *ctx = s->ctx;
  if (lvc_validate_sizes(ctx, r->numX, r->numY, r->numT))
    return -1;
  if (r->numX > SLOT_MAX_N || r->numY > SLOT_MAX_N) {
    lvc_set_error(ctx, "server slots hold grids up to %d x %d", SLOT_MAX_N,
                  SLOT_MAX_N);
    return -1;
  }
  if (lvc_lanes_for(r->numX, s->seg_x) > 64 || !lvc_yseg_fits(r->numY, s->seg_y)) {
    lvc_set_error(ctx, "grid too large for the server's segment lengths");
    return -1;
  }
  if (request_elems(r->numX, r->numY) > s->arena_elems) {
    lvc_set_error(ctx, "request needs %lld bytes of grids, more than the %lld bytes the "
                       "budget of %lld leaves for them",
                  (long long)request_bytes(s, r->numX, r->numY),
                  (long long)(2 * s->arena_elems * (int64_t)s->real), (long long)s->budget);
    return -1;
  }
  if (s->qtail == s->qcap) {
    /* compact, then grow */
    if (s->qhead > 0) {
      memmove(s->queue, s->queue + s->qhead,
              (size_t)(s->qtail - s->qhead) * sizeof(queued_req));
      s->qtail -= s->qhead;
      s->qhead = 0;
    }
    if (s->qtail == s->qcap) {
      int64_t nc = s->qcap ? 2 * s->qcap : 256;
      queued_req *nq = realloc(s->queue, (size_t)nc * sizeof(queued_req));
      if (!nq)
        return -1;
      s->queue = nq;
      s->qcap = nc;
    }
  }
  if (s->nresults == s->rcap) {
    int64_t nc = s->rcap ? 2 * s->rcap : 1024;
    result_rec *nr = realloc(s->results, (size_t)nc * sizeof(result_rec));
    if (!nr)
      return -1;
    s->results = nr;
    s->rcap = nc;
  }
  const int64_t id = s->nresults++;
  s->results[id].state = 0;
  s->results[id].price = 0;
  s->queue[s->qtail].id = id;
  s->queue[s->qtail].req = *r;
  s->qtail++;
  return id;
}

static int rebuild_tiles(struct lvc_server *s) {
  struct futhark_context *ctx = s->ctx;
  int nx_t = 0, ny_t = 0;
  for (int i = 0; i < s->max_slots; i++) {
    if (s->slot_id[i] < 0)
      continue;
    const lvc_problem *q = &s->h_slots[i];
    const int rows = LVC_BLOCK / lvc_lanes_for(q->nx, s->seg_x);
    nx_t += (q->ny + rows - 1) / rows;
    ny_t += q->ytiles;
  }
  const int need = nx_t + ny_t; /* <= SLOT_MAX_TILES per resident request */
  if (need > s->tiles_cap) {
    lvc_set_error(ctx, "server tiles exceed their preallocated count");
    return 1;
  }
  int tx = 0, ty = nx_t;
  s->ylds = 0;
  s->nfac = 0;
  for (int i = 0; i < s->max_slots; i++) {
    if (s->slot_id[i] < 0)
      continue;
    const lvc_problem *q = &s->h_slots[i];
    s->h_fprobs[s->nfac++] = i;
    const int gx = lvc_lanes_for(q->nx, s->seg_x);
    const int rows = LVC_BLOCK / gx;
    int gy, cpl, cols;
    lvc_ytile_shape(q->ny, s->seg_y, (int)s->real, &gy, &cpl, &cols);
    const int lds = lvc_ytile_lds(q->ny, s->seg_y, (int)s->real);
    if (lds > s->ylds)
      s->ylds = lds;
    for (int r0 = 0; r0 < q->ny; r0 += rows) {
      lvc_tile t = {i, r0, q->ny - r0 < rows ? q->ny - r0 : rows, gx, 1, 0};
      s->h_tiles[tx++] = t;
    }
    for (int c0 = 0; c0 < q->nx; c0 += cols) {
      lvc_tile t = {i, c0, q->nx - c0 < cols ? q->nx - c0 : cols, gy, cpl, 0};
      s->h_tiles[ty++] = t;
    }
  }
  s->nxt = nx_t;
  s->nyt = ny_t;
  s->tiles_dirty = 0;
  if (s->d_fprobs && s->nfac &&
      lvc_upload(ctx, s->d_fprobs, s->h_fprobs, (size_t)s->nfac * sizeof(int32_t)))
    return 1;
  return need ? lvc_upload(ctx, s->d_tiles, s->h_tiles, (size_t)need * sizeof(lvc_tile)) : 0;
}

static int admit(struct lvc_server *s) {
  struct futhark_context *ctx = s->ctx;
  int slot = 0;
  while (s->qhead < s->qtail && s->nresident < s->max_slots) {
    const queued_req *qr = &s->queue[s->qhead];
    const int64_t len = request_elems(qr->req.numX, qr->req.numY);
    const int64_t off = arena_take(s, len);
    if (off < 0)
      break; /* in-order admission: wait for memory */
    while (s->slot_id[slot] >= 0)
      slot++;
    lvc_problem q;
    memset(&q, 0, sizeof(q));
    q.nx = qr->req.numX;
    q.ny = qr->req.numY;
    q.nt = qr->req.numT;
    q.s0 = qr->req.s0;
    q.strike = qr->req.strike;
    q.t = qr->req.t;
    q.alpha = qr->req.alpha;
    q.nu = qr->req.nu;
    q.beta = qr->req.beta;
    q.k = -1;
    lvc_grid_scalars(&q, 0);
    q.state_off = off;
    q.geo_off = (int64_t)slot * s->slot_geo;
    s->slot_len[slot] = len;
    q.fac_src = slot;
    {
      int gy, cpl, cols;
      lvc_ytile_shape(q.ny, s->seg_y, (int)s->real, &gy, &cpl, &cols);
      q.ytiles = (q.nx + cols - 1) / cols;
    }
    s->h_slots[slot] = q;
    s->slot_id[slot] = qr->id;
    s->steps_left[slot] = q.nt - 1;
    s->results[qr->id].state = 1;
    s->resident_bytes += request_bytes(s, q.nx, q.ny);
    s->nresident++;
    s->qhead++;
    s->tiles_dirty = 1;
    s->h_adm[s->nadm] = q;
    s->h_admslot[s->nadm] = slot;
    s->nadm++;
  }
  /* the tick's admissions: their descriptors and slots in one upload each,
   * set up by one launch (which also stores each descriptor to its slot) */
  if (s->nadm > 0) {
    const int n = s->nadm;
    s->nadm = 0;
    if (lvc_upload(ctx, s->d_adm, s->h_adm, (size_t)n * sizeof(lvc_problem)) ||
        lvc_upload(ctx, s->d_admslot, s->h_admslot, (size_t)n * sizeof(int32_t)))
      return 1;
    const int e = lvc_launch_setup_list((void *)ctx->stream, s->precision, s->d_slots, s->d_adm,
                                        s->d_admslot, n, s->d_geo, s->d_R);
    if (lvc_check_hip(ctx, e, "setup kernel"))
      return 1;
  }
  return 0;
}

/* Publishes the completion batches whose price copies have landed, in
 * order; with `upto` >= 0, first waits for every batch up to that sequence
 * number. */
static int publish(struct lvc_server *s, int64_t upto) {
  struct futhark_context *ctx = s->ctx;
  while (s->pend_head < s->pend_tail) {
    completion_batch *b = &s->pend[s->pend_head % LVC_SRV_PENDING];
    if (b->seq > upto) {
      const hipError_t q = hipEventQuery(b->done);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError(); /* not an error: the copy is still in flight */
        return 0;
      }
      if (lvc_check_hip(ctx, q, "hipEventQuery"))
        return 1;
    } else if (lvc_check_hip(ctx, hipEventSynchronize(b->done), "hipEventSynchronize")) {
      return 1;
    }
    for (int j = 0; j < b->count; j++) {
      result_rec *r = &s->results[b->ids[j]];
      r->price = b->host[j];
      r->state = 2;
    }
    s->pend_head++;
  }
  return 0;
}

int32_t lvc_server_step(struct lvc_server *s) {
  if (!s)
    return -1;
  struct futhark_context *ctx = s->ctx;
  pthread_mutex_lock(&ctx->lock);
  int rc = publish(s, -1);
  if (!rc)
    rc = admit(s);
  if (!rc && s->tiles_dirty)
    rc = rebuild_tiles(s);
  int32_t ndone = 0;
  if (!rc && s->nresident > 0) {
    void *st = (void *)ctx->stream;
    int tok = lvc_prof_begin(ctx, "lvc_xsweep");
    int e = lvc_launch_xsweep(st, s->precision, s->seg_x, s->d_slots, s->d_tiles,
                              s->nxt, s->d_geo, s->d_R, s->d_Y, s->d_yfac,
                              lvc_seg_is_col(s->seg_y) ? s->seg_y : 0, s->d_fprobs,
                              s->nfac, 0);
    lvc_prof_end(ctx, tok);
    rc = lvc_check_hip(ctx, e, "x sweep kernel");
    if (!rc) {
      tok = lvc_prof_begin(ctx, "lvc_ysweep");
      e = lvc_launch_ysweep(st, s->precision, s->seg_y, s->d_slots,
                            s->d_tiles + s->nxt, s->nyt, s->d_geo, s->d_Y, s->d_R,
                            s->ylds, s->d_yfac);
      lvc_prof_end(ctx, tok);
      rc = lvc_check_hip(ctx, e, "y sweep kernel");
    }
    /* completions: one gather over the finished slots, its prices copied
     * to a pinned buffer behind the launches and published once landed
     * (poll waits for them); the slots and arena extents are free at once,
     * since every later use of them runs after the gather in stream order */
    for (int i = 0; !rc && i < s->max_slots; i++) {
      if (s->slot_id[i] < 0)
        continue;
      if (--s->steps_left[i] > 0)
        continue;
      s->h_done[ndone++] = i;
    }
    completion_batch *cb = NULL;
    if (!rc && ndone > 0) {
      if (s->pend_tail - s->pend_head == LVC_SRV_PENDING)
        rc = publish(s, s->pend_head);            /* the ring is full: the oldest */
      cb = &s->pend[s->pend_tail % LVC_SRV_PENDING];
      if (!rc)
        rc = lvc_upload(ctx, s->d_done, s->h_done, (size_t)ndone * sizeof(int32_t));
      if (!rc)
        rc = lvc_check_hip(ctx, lvc_launch_gather_list(st, s->precision, s->d_slots, s->d_done,
                                                       ndone, s->d_R, s->d_out, 1),
                           "gather kernel") ||
             lvc_check_hip(ctx, hipMemcpyAsync(cb->host, s->d_out, (size_t)ndone * sizeof(double),
                                               hipMemcpyDeviceToHost, ctx->stream),
                           "hipMemcpyAsync") ||
             lvc_check_hip(ctx, hipEventRecord(cb->done, ctx->stream), "hipEventRecord");
      if (!rc) {
        cb->count = ndone;
        cb->seq = s->pend_tail++;
      }
    }
    for (int j = 0; !rc && j < ndone; j++) {
      const int i = s->h_done[j];
      {
        const int64_t id = s->slot_id[i];
        cb->ids[j] = id;
        s->results[id].state = 3;
        s->results[id].batch = cb->seq;
        s->resident_bytes -= request_bytes(s, s->h_slots[i].nx, s->h_slots[i].ny);
        arena_give(s, s->h_slots[i].state_off, s->slot_len[i]);
        s->slot_len[i] = 0;
        s->slot_id[i] = -1;
        s->h_slots[i].k = -1;
        s->nresident--;
        s->completed++;
        s->tiles_dirty = 1;
      }
    }
  }
  pthread_mutex_unlock(&ctx->lock);
  return rc ? -1 : ndone;
}

int32_t lvc_server_poll(struct lvc_server *s, int64_t id, double *price) {
  if (!s || id < 0 || id >= s->nresults)
    return -1;
  struct futhark_context *ctx = s->ctx;
  pthread_mutex_lock(&ctx->lock);
  /* finished on the device: wait for its price copy (a tick at most) */
  int rc = s->results[id].state == 3 ? publish(s, s->results[id].batch) : publish(s, -1);
  const int done = s->results[id].state == 2;
  const double v = s->results[id].price;
  pthread_mutex_unlock(&ctx->lock);
  if (rc)
    return -1;
  if (!done)
    return 0;
  if (price)
    *price = v;
  return 1;
}

int32_t lvc_server_drain(struct lvc_server *s) {
  if (!s)
    return -1;
  while (s->qhead < s->qtail || s->nresident > 0) {
    const int64_t before_q = s->qhead;
    const int was_empty = s->nresident == 0;
    int32_t r = lvc_server_step(s);
    if (r < 0)
      return -1;
    /* with nothing resident at the start of the tick, admission cannot wait
     * for memory: submit() guaranteed every queued request fits alone */
    if (was_empty && s->qhead == before_q && s->qhead < s->qtail) {
      lvc_set_error(s->ctx, "queued request cannot be admitted");
      return -1;
    }
  }
  pthread_mutex_lock(&s->ctx->lock);
  const int rc = publish(s, s->pend_tail - 1);
  pthread_mutex_unlock(&s->ctx->lock);
  return rc ? -1 : 0;
}

int64_t lvc_server_device_bytes(struct lvc_server *s) {
  return s ? s->device_bytes : -1;
}

void lvc_server_stats(struct lvc_server *s, int64_t *queued, int64_t *resident,
                      int64_t *completed, int64_t *resident_bytes) {
  if (!s)
    return;
  if (queued)
    *queued = s->qtail - s->qhead;
  if (resident)
    *resident = s->nresident;
  if (completed)
    *completed = s->completed;
  if (resident_bytes)
    *resident_bytes = s->resident_bytes;
}
