/*
 * Batched tridiagonal solves of the reference's ocean-sim test program
 * (misc/ocean-sim/tridiag-test.fut), as the C API that Futhark's C backend
 * generates for its four entry points, exported by the same library as
 * include/locvolcalib.h and sharing its context (config, context, error and
 * sync functions are declared there).
 *
 * Each entry solves n independent tridiagonal systems, one per row of the
 * [n][m] f32 arrays a (sub-diagonal), b (diagonal), c (super-diagonal) and
 * y (right-hand side), into a new [n][m] array owned by the caller:
 *
 *   tridagNested, tridagNestedConst      tridagPar, the scan form
 *                                        (misc/ocean-sim/tridiag.fut:4-79)
 *   tridagNestedSeq, tridagNestedSeqConst  tridagSeq, Thomas (tridiag.fut:81-95)
 *
 * The Const entries take m = INNER_DIM = 115 only (tridiag-test.fut:4); an
 * array of another inner size is an error, as a shape mismatch is for the
 * generated API.  Every entry requires the four shapes to agree and m >= 1
 * (n = 0 returns an empty [0][m] array).  Results are those of a sequential
 * f32 evaluation of the reference's lines (one system per GPU thread, IEEE
 * division, no contraction).
 */
#ifndef OCEAN_TRIDIAG_H
#define OCEAN_TRIDIAG_H

#include <stdint.h>

#include "futhark_arrays.h"
#include "locvolcalib.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The [n][m] f32 arrays are those of futhark_arrays.h. */

/* tridiag-test.fut:14-15 */
int futhark_entry_tridagNested(struct futhark_context *ctx,
                               struct futhark_f32_2d **out0,
                               const struct futhark_f32_2d *in0,
                               const struct futhark_f32_2d *in1,
                               const struct futhark_f32_2d *in2,
                               const struct futhark_f32_2d *in3);
/* tridiag-test.fut:17-18 */
int futhark_entry_tridagNestedConst(struct futhark_context *ctx,
                                    struct futhark_f32_2d **out0,
                                    const struct futhark_f32_2d *in0,
                                    const struct futhark_f32_2d *in1,
                                    const struct futhark_f32_2d *in2,
                                    const struct futhark_f32_2d *in3);
/* tridiag-test.fut:20-22 */
int futhark_entry_tridagNestedSeq(struct futhark_context *ctx,
                                  struct futhark_f32_2d **out0,
                                  const struct futhark_f32_2d *in0,
                                  const struct futhark_f32_2d *in1,
                                  const struct futhark_f32_2d *in2,
                                  const struct futhark_f32_2d *in3);
/* tridiag-test.fut:24-26 */
int futhark_entry_tridagNestedSeqConst(struct futhark_context *ctx,
                                       struct futhark_f32_2d **out0,
                                       const struct futhark_f32_2d *in0,
                                       const struct futhark_f32_2d *in1,
                                       const struct futhark_f32_2d *in2,
                                       const struct futhark_f32_2d *in3);

#ifdef __cplusplus
}
#endif

#endif /* OCEAN_TRIDIAG_H */
