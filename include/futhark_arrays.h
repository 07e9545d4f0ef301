/*
 * futhark_arrays.h -- the multi-dimensional arrays of the generated C API
 * that the programs of this build take (the ocean-sim entries' [n][m] f32
 * systems, OptionPricing's i32 and f32 tables), with the generator's names:
 * futhark_new_<type>_<rank>d, _new_raw_, _free_, _values_, _values_raw_ and
 * _shape_.  Row-major, device memory, owned by the caller; the context API
 * is in futhark_context.h.
 */
#ifndef FUTHARK_ARRAYS_H
#define FUTHARK_ARRAYS_H

#include "futhark_context.h"

#ifdef __cplusplus
extern "C" {
#endif

struct futhark_f32_2d;
struct futhark_i32_2d;
struct futhark_f32_3d;

/* [dim0][dim1] f32 */
struct futhark_f32_2d *futhark_new_f32_2d(struct futhark_context *ctx,
                                          const float *data, int64_t dim0,
                                          int64_t dim1);
struct futhark_f32_2d *futhark_new_raw_f32_2d(struct futhark_context *ctx,
                                              void *data, int64_t dim0,
                                              int64_t dim1);
int futhark_free_f32_2d(struct futhark_context *ctx, struct futhark_f32_2d *arr);
int futhark_values_f32_2d(struct futhark_context *ctx, struct futhark_f32_2d *arr,
                          float *data);
void *futhark_values_raw_f32_2d(struct futhark_context *ctx,
                                struct futhark_f32_2d *arr);
const int64_t *futhark_shape_f32_2d(struct futhark_context *ctx,
                                    struct futhark_f32_2d *arr);

/* [dim0][dim1] i32 */
struct futhark_i32_2d *futhark_new_i32_2d(struct futhark_context *ctx,
                                          const int32_t *data, int64_t dim0,
                                          int64_t dim1);
struct futhark_i32_2d *futhark_new_raw_i32_2d(struct futhark_context *ctx,
                                              void *data, int64_t dim0,
                                              int64_t dim1);
int futhark_free_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr);
int futhark_values_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr,
                          int32_t *data);
void *futhark_values_raw_i32_2d(struct futhark_context *ctx,
                                struct futhark_i32_2d *arr);
const int64_t *futhark_shape_i32_2d(struct futhark_context *ctx,
                                    struct futhark_i32_2d *arr);

/* [dim0][dim1][dim2] f32 */
struct futhark_f32_3d *futhark_new_f32_3d(struct futhark_context *ctx,
                                          const float *data, int64_t dim0,
                                          int64_t dim1, int64_t dim2);
struct futhark_f32_3d *futhark_new_raw_f32_3d(struct futhark_context *ctx,
                                              void *data, int64_t dim0,
                                              int64_t dim1, int64_t dim2);
int futhark_free_f32_3d(struct futhark_context *ctx, struct futhark_f32_3d *arr);
int futhark_values_f32_3d(struct futhark_context *ctx, struct futhark_f32_3d *arr,
                          float *data);
void *futhark_values_raw_f32_3d(struct futhark_context *ctx,
                                struct futhark_f32_3d *arr);
const int64_t *futhark_shape_f32_3d(struct futhark_context *ctx,
                                    struct futhark_f32_3d *arr);

#ifdef __cplusplus
}
#endif

#endif /* FUTHARK_ARRAYS_H */
