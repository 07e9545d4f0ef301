/*
 * futhark_context.h -- the part of the generated C API that every program of
 * this build shares: configuration, context, errors, synchronisation,
 * profiling and the 1-D arrays.  The program headers (locvolcalib.h,
 * optionpricing.h) include it and add their entry points; each program is
 * its own shared library, as `futhark <backend> --library` makes one per
 * program.
 *
 * Conventions (as in the generated API):
 *  - every fallible call returns int: 0 = success, non-zero = failure; the
 *    message is then available from futhark_context_get_error();
 *  - arrays are opaque handles owned by the caller, holding device memory;
 *    futhark_values_* copies them to host memory;
 *  - entry points are asynchronous with respect to the host; results are
 *    valid after futhark_context_sync() (futhark_values_* synchronises);
 *  - the configuration must outlive the context and is freed after it;
 *  - one context per device; calls on one context are serialised by an
 *    internal lock.
 *
 * Everything here is plain C: pointers, sizes and scalars, no HIP or torch
 * types.  Device pointers are exposed as `void *`.
 */
#ifndef FUTHARK_CONTEXT_H
#define FUTHARK_CONTEXT_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

struct futhark_context_config;
struct futhark_context;
struct futhark_f32_1d;
struct futhark_f64_1d;
struct futhark_i32_1d;

/* ---- configuration ------------------------------------------------------ */

/* lys shared.c:20 */
struct futhark_context_config *futhark_context_config_new(void);
/* lys console/main.c:180 */
void futhark_context_config_free(struct futhark_context_config *cfg);
/* lys shared.c:25.  Accepts "hip:N", "#N" or "N" (device ordinal). */
void futhark_context_config_set_device(struct futhark_context_config *cfg,
                                       const char *s);
/* lys shared.c:44.  Here: the autotuner's cache file (read at context
 * creation when it exists, written after tuning). */
void futhark_context_config_set_cache_file(struct futhark_context_config *cfg,
                                           const char *f);
/* Generated API (set_tuning_param; the reference's knobs are in
 * finpar/LocVolCalib.fut.tuning:1-2).  Recognised names:
 *   "main.suff_intra_par_0", "main.suff_intra_par_1"  accepted, no effect on
 *       results (the reference's version-selection thresholds);
 *   "lvc.precision"  0 = FP32, 1 = FP32 + one FP64-residual refinement per
 *       implicit solve (default), 2 = FP64 (for futhark_entry_main);
 *   "lvc.seg_x"  unknowns per lane of the x sweep's row systems: 4, 8 or
 *       16 (0 = tuned per size class);
 *   "lvc.seg_y"  the y sweep's solver: lane segments 4, 8, 16 (PCR across
 *       the lanes of a column) or column segments 32, 64 (one lane per
 *       column segment, Thomas; 0 = tuned per size class);
 *   "lvc.autotune"  1 = time the candidates on first use of a size class
 *       (default 0: the cache file, else the built-in per-size table);
 *   "lvc.xshare"  options per x tile sharing one row factorisation, 1..16
 *       (0 = chosen per batch);
 *   "lvc.graph"  1 = replay a call's step launches from the context's kept
 *       HIP graph: relaunched as is when a call repeats the launches,
 *       updated (or rebuilt) otherwise (default 0: direct launches).
 * A size too large for the chosen segment length takes the next one that
 * fits.  Returns non-zero for an unknown name or an unsupported value. */
int futhark_context_config_set_tuning_param(struct futhark_context_config *cfg,
                                            const char *param_name,
                                            size_t new_value);
/* Debugging: each entry call also reads back, after its price gather, the
 * price index (myYindex, myXindex) its first option used, and
 * futhark_context_report lists it as "price_index". */
void futhark_context_config_set_debugging(struct futhark_context_config *cfg,
                                          int flag);
/* Per-kernel timing with HIP events; read back with futhark_context_report. */
void futhark_context_config_set_profiling(struct futhark_context_config *cfg,
                                          int flag);
void futhark_context_config_set_logging(struct futhark_context_config *cfg,
                                        int flag);

int futhark_get_tuning_param_count(void);
const char *futhark_get_tuning_param_name(int i);
const char *futhark_get_tuning_param_class(int i);

/* ---- context ------------------------------------------------------------ */

/* lys shared.c:47.  Never returns NULL except on host allocation failure; a
 * device that cannot be opened is reported by futhark_context_get_error()
 * and makes every later call fail. */
struct futhark_context *futhark_context_new(struct futhark_context_config *cfg);
/* lys console/main.c:179 */
void futhark_context_free(struct futhark_context *ctx);
/* lys console/liblys.c:249 */
int futhark_context_sync(struct futhark_context *ctx);
/* lys shared.h:25.  Returns a malloc'd message (caller frees) or NULL, and
 * clears the stored error. */
char *futhark_context_get_error(struct futhark_context *ctx);
void futhark_context_set_logging_file(struct futhark_context *ctx, FILE *f);
/* JSON object: per-kernel {"count", "total_us"} collected while profiling is
 * enabled.  Caller frees. */
char *futhark_context_report(struct futhark_context *ctx);
void futhark_context_pause_profiling(struct futhark_context *ctx);
void futhark_context_unpause_profiling(struct futhark_context *ctx);
/* Releases cached device memory held by the context's allocator. */
int futhark_context_clear_caches(struct futhark_context *ctx);

/* ---- arrays (cf. lys console/liblys.c:248-250, shared.c:79) ------------- */

struct futhark_f32_1d *futhark_new_f32_1d(struct futhark_context *ctx,
                                          const float *data, int64_t dim0);
/* Wraps existing device memory (not copied, not freed by the library). */
struct futhark_f32_1d *futhark_new_raw_f32_1d(struct futhark_context *ctx,
                                              void *data, int64_t dim0);
int futhark_free_f32_1d(struct futhark_context *ctx, struct futhark_f32_1d *arr);
int futhark_values_f32_1d(struct futhark_context *ctx, struct futhark_f32_1d *arr,
                          float *data);
int futhark_index_f32_1d(struct futhark_context *ctx, float *out,
                         struct futhark_f32_1d *arr, int64_t i0);
void *futhark_values_raw_f32_1d(struct futhark_context *ctx,
                                struct futhark_f32_1d *arr);
const int64_t *futhark_shape_f32_1d(struct futhark_context *ctx,
                                    struct futhark_f32_1d *arr);

struct futhark_f64_1d *futhark_new_f64_1d(struct futhark_context *ctx,
                                          const double *data, int64_t dim0);
struct futhark_f64_1d *futhark_new_raw_f64_1d(struct futhark_context *ctx,
                                              void *data, int64_t dim0);
int futhark_free_f64_1d(struct futhark_context *ctx, struct futhark_f64_1d *arr);
int futhark_values_f64_1d(struct futhark_context *ctx, struct futhark_f64_1d *arr,
                          double *data);
int futhark_index_f64_1d(struct futhark_context *ctx, double *out,
                         struct futhark_f64_1d *arr, int64_t i0);
void *futhark_values_raw_f64_1d(struct futhark_context *ctx,
                                struct futhark_f64_1d *arr);
const int64_t *futhark_shape_f64_1d(struct futhark_context *ctx,
                                    struct futhark_f64_1d *arr);

struct futhark_i32_1d *futhark_new_i32_1d(struct futhark_context *ctx,
                                          const int32_t *data, int64_t dim0);
int futhark_free_i32_1d(struct futhark_context *ctx, struct futhark_i32_1d *arr);
int futhark_values_i32_1d(struct futhark_context *ctx, struct futhark_i32_1d *arr,
                          int32_t *data);
const int64_t *futhark_shape_i32_1d(struct futhark_context *ctx,
                                    struct futhark_i32_1d *arr);

#ifdef __cplusplus
}
#endif

#endif /* FUTHARK_CONTEXT_H */
