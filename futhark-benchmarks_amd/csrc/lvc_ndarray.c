/*
 * The multi-dimensional arrays of the generated C API (futhark_arrays.h):
 * device memory from the context's pool, uploaded and read back on the
 * context's stream.
 */
#include <stdlib.h>
#include <string.h>

#include "../../include/futhark_arrays.h"
#include "lvc_ndarray.h"
#include "lvc_runtime.h"

static int dims_ok(struct futhark_context *ctx, const int64_t *dims, int rank, size_t elem) {
  size_t n = 1;
  for (int i = 0; i < rank; i++) {
    if (dims[i] < 0) {
      lvc_set_error(ctx, "negative array size %lld", (long long)dims[i]);
      return 0;
    }
    if (dims[i] != 0 && n > SIZE_MAX / elem / (size_t)dims[i]) {
      lvc_set_error(ctx, "array too large");
      return 0;
    }
    n *= (size_t)dims[i];
  }
  return 1;
}

static size_t nd_bytes(const int64_t *dims, int rank, size_t elem) {
  size_t n = elem;
  for (int i = 0; i < rank; i++) n *= (size_t)dims[i];
  return n;
}

/* Allocates (data == NULL: uninitialised) or, with raw, wraps raw_mem; an
 * empty array holds no memory. */
static void *nd_new(struct futhark_context *ctx, size_t hdr, const void *data, int raw,
                    void *raw_mem, const int64_t *dims, int rank, size_t elem) {
  if (!lvc_ready(ctx) || !dims_ok(ctx, dims, rank, elem))
    return NULL;
  lvc_ndarray *arr = calloc(1, hdr);
  if (!arr) {
    lvc_set_error(ctx, "host allocation failed");
    return NULL;
  }
  for (int i = 0; i < 3; i++) arr->shape[i] = i < rank ? dims[i] : 1;
  arr->bytes = nd_bytes(dims, rank, elem);
  if (raw || arr->bytes == 0) {
    arr->mem = raw ? raw_mem : NULL;
    arr->raw = 1;
    return arr;
  }
  pthread_mutex_lock(&ctx->lock);
  int rc = lvc_dev_alloc(ctx, arr->bytes, &arr->mem);
  if (!rc && data && arr->bytes)
    rc = lvc_check_hip(ctx, hipMemcpyAsync(arr->mem, data, arr->bytes, hipMemcpyHostToDevice,
                                           ctx->stream),
                       "hipMemcpyAsync") ||
         lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (rc && arr->mem)
    lvc_dev_free(ctx, arr->mem, arr->bytes);
  pthread_mutex_unlock(&ctx->lock);
  if (rc) {
    free(arr);
    return NULL;
  }
  return arr;
}

static int nd_free(struct futhark_context *ctx, lvc_ndarray *arr) {
  if (!ctx || !arr)
    return 0;
  pthread_mutex_lock(&ctx->lock);
  if (!arr->raw && ctx->ok)
    lvc_dev_free(ctx, arr->mem, arr->bytes);
  pthread_mutex_unlock(&ctx->lock);
  free(arr);
  return 0;
}

static int nd_values(struct futhark_context *ctx, lvc_ndarray *arr, void *data) {
  if (!lvc_ready(ctx))
    return 1;
  if (!arr) {
    lvc_set_error(ctx, "NULL array");
    return 1;
  }
  if (arr->bytes == 0)
    return 0;
  pthread_mutex_lock(&ctx->lock);
  int rc = lvc_check_hip(ctx, hipMemcpyAsync(data, arr->mem, arr->bytes, hipMemcpyDeviceToHost,
                                             ctx->stream),
                         "hipMemcpyAsync") ||
           lvc_check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (!rc)
    lvc_prof_collect(ctx);
  pthread_mutex_unlock(&ctx->lock);
  return rc;
}

#define LVC_ND_API(TY, CT, RANK, DIMS_DECL, DIMS)                                        \
  struct futhark_##TY *futhark_new_##TY(struct futhark_context *ctx, const CT *data,     \
                                        DIMS_DECL) {                                     \
    const int64_t d[3] = {DIMS};                                                         \
    return nd_new(ctx, sizeof(struct futhark_##TY), data, 0, NULL, d, RANK, sizeof(CT));    \
  }                                                                                      \
  struct futhark_##TY *futhark_new_raw_##TY(struct futhark_context *ctx, void *data,     \
                                            DIMS_DECL) {                                 \
    const int64_t d[3] = {DIMS};                                                         \
    return nd_new(ctx, sizeof(struct futhark_##TY), NULL, 1, data, d, RANK, sizeof(CT));    \
  }                                                                                      \
  int futhark_free_##TY(struct futhark_context *ctx, struct futhark_##TY *arr) {         \
    return nd_free(ctx, arr ? &arr->a : NULL);                                           \
  }                                                                                      \
  int futhark_values_##TY(struct futhark_context *ctx, struct futhark_##TY *arr,         \
                          CT *data) {                                                    \
    return nd_values(ctx, arr ? &arr->a : NULL, data);                                   \
  }                                                                                      \
  void *futhark_values_raw_##TY(struct futhark_context *ctx, struct futhark_##TY *arr) { \
    (void)ctx;                                                                           \
    return arr ? arr->a.mem : NULL;                                                      \
  }                                                                                      \
  const int64_t *futhark_shape_##TY(struct futhark_context *ctx,                         \
                                    struct futhark_##TY *arr) {                          \
    (void)ctx;                                                                           \
    return arr ? arr->a.shape : NULL;                                                    \
  }

#define D2_DECL int64_t dim0, int64_t dim1
#define D2 dim0, dim1, 1
#define D3_DECL int64_t dim0, int64_t dim1, int64_t dim2
#define D3 dim0, dim1, dim2

LVC_ND_API(f32_2d, float, 2, D2_DECL, D2)
LVC_ND_API(i32_2d, int32_t, 2, D2_DECL, D2)
LVC_ND_API(f32_3d, float, 3, D3_DECL, D3)
