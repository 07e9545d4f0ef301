/*
 * locvolcalib.h -- C ABI of the MI355X LocVolCalib library.
 *
 * Drop-in for the C API that `futhark <backend> --library` generates for
 * finpar/LocVolCalib.fut.  The generated header is not part of the reference
 * repository; the names, argument order and ownership rules below follow the
 * generator's convention as exercised by the reference's own C callers (the
 * vendored lys frontends, e.g. accelerate/julia/lib/github.com/diku-dk/lys/
 * shared.c:15-49, console/liblys.c:243-251, console/main.c:176-180,
 * shared.h:20-28).  Each declaration cites the caller line it serves or the
 * LocVolCalib.fut line that fixes its signature.
 *
 * The configuration, context and 1-D array functions, and the conventions
 * they follow, are in futhark_context.h (shared with the other programs of
 * this build); this header adds LocVolCalib's entry points and the serving
 * API.  Everything is plain C: pointers, sizes and scalars, no HIP or torch
 * types.
 */
#ifndef LOCVOLCALIB_H
#define LOCVOLCALIB_H

#include "futhark_context.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- entry points ------------------------------------------------------- */

/* LocVolCalib.fut:238-252 (`main`): prices of `outer` options with strikes
 * 0.001*i, i = 0..outer-1, all sharing the grid sizes and parameters.
 * initGrid's scalars (stdX, dx, myXindex = i32.f32(s0/dx): :15-20) are
 * formed in f32 from the f32 arguments, as the reference does, so the price
 * index is the reference's; the rest is computed in the precision set by
 * "lvc.precision" (default FP32 with one FP64-residual refinement per
 * implicit solve, on an FP64 grid).  Errors: numX or numY < 3 or numT < 2
 * (the reference's range `1...n-2` at :37 and its timeline), outer < 0
 * (`iota` at :247), a price index outside the grid (the bounds check of
 * `myResult[myYindex, myXindex]` at :236, e.g. 20*alpha*sqrt(t) <= 1),
 * device failure.  outer == 0 returns an empty array. */
int futhark_entry_main(struct futhark_context *ctx, struct futhark_f32_1d **out0,
                       const int32_t outer_loop_count, const int32_t numX,
                       const int32_t numY, const int32_t numT, const float s0,
                       const float t, const float alpha, const float nu,
                       const float beta);

/* The same program with every real f64 (the FP64 plain form): the f32 ->
 * f64 type substitution of LocVolCalib.fut:238-252, initGrid's scalars
 * included.  Errors as futhark_entry_main.  No counterpart in the
 * reference, which is f32-only; named as the generator names an extra
 * `entry main_f64`. */
int futhark_entry_main_f64(struct futhark_context *ctx,
                           struct futhark_f64_1d **out0,
                           const int32_t outer_loop_count, const int32_t numX,
                           const int32_t numY, const int32_t numT,
                           const double s0, const double t, const double alpha,
                           const double nu, const double beta);

/* futhark_entry_main without device memory allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same options, precision ("lvc.precision"), segment lengths and kernels as
 * futhark_entry_main, so the same prices bit for bit; the prices go to
 * `out`, a device array of `outer` floats owned by the caller (for example
 * from futhark_values_raw_f32_1d), enqueued on the context's stream
 * (futhark_context_sync waits for them).  Every buffer comes from one static
 * device array of 72 MiB per GPU (held until the process ends): main(128,
 * 256, 256, ...) in the FP32 forms, main(64, 256, 256, ...) in FP64; larger
 * calls fail with a message giving the bytes they need.  Calls on one GPU
 * share that array: calls of one context are ordered by its stream, but calls
 * from two contexts on the same GPU that overlap in time race.  Segment
 * lengths come from the cache file or the built-in table (the autotuner's
 * timing allocates, so it is never run here); profiling does not cover these
 * calls.  Errors as futhark_entry_main, plus NULL `out` and the size limit. */
int futhark_entry_main_noalloc(struct futhark_context *ctx, float *out,
                               const int32_t outer_loop_count, const int32_t numX,
                               const int32_t numY, const int32_t numT,
                               const float s0, const float t, const float alpha,
                               const float nu, const float beta);

/* Batched form: problem p has its own sizes (numX[p], numY[p], numT[p]) and
 * parameters and runs LocVolCalib.fut:211-236 (`value`).  All arrays have
 * the same length n.  `precision` as "lvc.precision".  The problems advance
 * together, per time step one launch of each sweep per segment-pair class
 * (each problem takes its own size class's pair, so its price is the same
 * alone and in any batch) over the problems still running (replayed from the
 * context's kept HIP graph with "lvc.graph"), until each has done its own
 * numT-1 steps.  A problem whose price index lies outside its grid
 * gives NaN (the call succeeds).  No counterpart in the reference (its
 * `main` shares one size and parameter set). */
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
                             int32_t precision);

/* ---- serving (continuous batching) -------------------------------------
 * No counterpart in the reference.  A server owns a set of resident slots on
 * the context's device: each admitted request keeps its grid in device memory
 * and advances one time step per engine tick together with every other
 * resident request; finished requests free their slot and memory, and queued
 * requests are admitted as slots and the memory budget allow.  A request's
 * result depends only on its own inputs, never on what else is resident. */
struct lvc_server;

struct lvc_request {
  int32_t numX, numY, numT;
  double s0, strike, t, alpha, nu, beta;
};

/* max_slots: resident requests at most (grids up to 512 x 512);
 * mem_budget_bytes: device memory the server may hold, all of it allocated
 * here: per-slot tables (descriptor, geometry, y factor table, tiles: ~52 KB
 * per slot in FP32) and, from the rest, one arena in which each resident
 * request's two grids take an extent of their own size (a request waits in
 * the queue until the arena has room); 0 = an arena for max_slots grids of
 * 512 x 512.  precision as "lvc.precision". */
struct lvc_server *lvc_server_new(struct futhark_context *ctx, int32_t max_slots,
                                  int64_t mem_budget_bytes, int32_t precision);
void lvc_server_free(struct lvc_server *srv);
/* Queues a request; returns its id (>= 0) or -1 (invalid request or a
 * request that can never fit the memory budget; see get_error). */
int64_t lvc_server_submit(struct lvc_server *srv, const struct lvc_request *req);
/* Runs one engine tick: admits queued requests, advances every resident
 * request by one time step, completes the ones that finished.  Enqueues its
 * work on the context's stream without waiting for it: the prices of the
 * requests a tick completes are copied back behind its launches.  Returns
 * the number of requests completed by this tick, or -1 on error. */
int32_t lvc_server_step(struct lvc_server *srv);
/* Completed result for `id`: returns 1 and stores the price, 0 when not yet
 * complete, -1 for an unknown id.  For a request a tick has completed, waits
 * for that tick's price copy if it has not landed yet. */
int32_t lvc_server_poll(struct lvc_server *srv, int64_t id, double *price);
/* Ticks until nothing is queued or resident.  Returns 0 or -1. */
int32_t lvc_server_drain(struct lvc_server *srv);
/* Device bytes the server holds (slot tables + grid arena), <= the budget
 * when one was given; -1 for NULL. */
int64_t lvc_server_device_bytes(struct lvc_server *srv);
/* Counters: queued, resident, completed, bytes of the resident grids. */
void lvc_server_stats(struct lvc_server *srv, int64_t *queued, int64_t *resident,
                      int64_t *completed, int64_t *resident_bytes);

#ifdef __cplusplus
}
#endif

#endif /* LOCVOLCALIB_H */
