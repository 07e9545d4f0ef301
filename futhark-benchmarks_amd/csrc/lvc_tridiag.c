/*
 * Entry points of the ocean-sim tridiagonal test program
 * (misc/ocean-sim/tridiag-test.fut:14-26) and the 2-D f32 arrays they take.
 * The solves run in csrc/lvc_tridiag_kernels.hip; this file checks shapes, manages
 * device memory through the context's pool and launches them on the
 * context's stream.
 */
#include <stdlib.h>
#include <string.h>

#include "../../include/ocean_tridiag.h"
#include "lvc_ndarray.h"
#include "lvc_runtime.h"

#define INNER_DIM 115 /* tridiag-test.fut:4 */

/* form 0: tridagPar, 1: tridagSeq; inner: required m (0 = any). */
static int tridag_entry(struct futhark_context *ctx, struct futhark_f32_2d **out0,
                        const struct futhark_f32_2d *a,
                        const struct futhark_f32_2d *b,
                        const struct futhark_f32_2d *c,
                        const struct futhark_f32_2d *y, int form, int64_t inner,
                        const char *name) {
  if (!out0)
    return 1;
  *out0 = NULL;
  if (!lvc_ready(ctx))
    return 1;
  if (!a || !b || !c || !y) {
    lvc_set_error(ctx, "%s: NULL input array", name);
    return 1;
  }
  const int64_t n = a->a.shape[0], m = a->a.shape[1];
  const struct futhark_f32_2d *in[3] = {b, c, y};
  for (int k = 0; k < 3; k++)
    if (in[k]->a.shape[0] != n || in[k]->a.shape[1] != m) {
      lvc_set_error(ctx, "%s: input %d has shape [%lld][%lld], input 0 [%lld][%lld]",
                    name, k + 1, (long long)in[k]->a.shape[0],
                    (long long)in[k]->a.shape[1], (long long)n, (long long)m);
      return 1;
    }
  if (inner && m != inner) {
    lvc_set_error(ctx, "%s: inner dimension must be %lld (INNER_DIM), got %lld",
                  name, (long long)inner, (long long)m);
    return 1;
  }
  if (m < 1 && n > 0) {
    lvc_set_error(ctx, "%s: systems of size %lld (must be at least 1)", name,
                  (long long)m);
    return 1;
  }
  if (m > INT32_MAX) {
    lvc_set_error(ctx, "%s: systems of size %lld are too long", name, (long long)m);
    return 1;
  }
  struct futhark_f32_2d *out = calloc(1, sizeof(*out));
  if (!out) {
    lvc_set_error(ctx, "host allocation failed");
    return 1;
  }
  out->a.shape[0] = n;
  out->a.shape[1] = m;
  out->a.shape[2] = 1;
  out->a.bytes = (size_t)n * (size_t)m * sizeof(float);
  pthread_mutex_lock(&ctx->lock);
  void *scratch = NULL;
  const size_t sbytes = m > lvc_tridag_lds_rows() ? out->a.bytes : 0;
  int rc = lvc_dev_alloc(ctx, out->a.bytes, &out->a.mem) ||
           (sbytes && lvc_dev_alloc(ctx, sbytes, &scratch));
  if (!rc && n > 0) {
    int tok = lvc_prof_begin(ctx, "lvc_tridag");
    int e = lvc_launch_tridag((void *)ctx->stream, form, a->a.mem, b->a.mem, c->a.mem,
                              y->a.mem, out->a.mem, scratch, n, (int)m);
    lvc_prof_end(ctx, tok);
    rc = lvc_check_hip(ctx, e, "tridiagonal kernel");
  }
  if (scratch)
    lvc_dev_free(ctx, scratch, sbytes);
  if (rc && out->a.mem)
    lvc_dev_free(ctx, out->a.mem, out->a.bytes);
  pthread_mutex_unlock(&ctx->lock);
  if (rc) {
    free(out);
    return 1;
  }
  *out0 = out;
  return 0;
}

int futhark_entry_tridagNested(struct futhark_context *ctx,
                               struct futhark_f32_2d **out0,
                               const struct futhark_f32_2d *in0,
                               const struct futhark_f32_2d *in1,
                               const struct futhark_f32_2d *in2,
                               const struct futhark_f32_2d *in3) {
  return tridag_entry(ctx, out0, in0, in1, in2, in3, 0, 0, "tridagNested");
}

int futhark_entry_tridagNestedConst(struct futhark_context *ctx,
                                    struct futhark_f32_2d **out0,
                                    const struct futhark_f32_2d *in0,
                                    const struct futhark_f32_2d *in1,
                                    const struct futhark_f32_2d *in2,
                                    const struct futhark_f32_2d *in3) {
  return tridag_entry(ctx, out0, in0, in1, in2, in3, 0, INNER_DIM,
                      "tridagNestedConst");
}

int futhark_entry_tridagNestedSeq(struct futhark_context *ctx,
                                  struct futhark_f32_2d **out0,
                                  const struct futhark_f32_2d *in0,
                                  const struct futhark_f32_2d *in1,
                                  const struct futhark_f32_2d *in2,
                                  const struct futhark_f32_2d *in3) {
  return tridag_entry(ctx, out0, in0, in1, in2, in3, 1, 0, "tridagNestedSeq");
}

int futhark_entry_tridagNestedSeqConst(struct futhark_context *ctx,
                                       struct futhark_f32_2d **out0,
                                       const struct futhark_f32_2d *in0,
                                       const struct futhark_f32_2d *in1,
                                       const struct futhark_f32_2d *in2,
                                       const struct futhark_f32_2d *in3) {
  return tridag_entry(ctx, out0, in0, in1, in2, in3, 1, INNER_DIM,
                      "tridagNestedSeqConst");
}
