/*
 * Entry point of finpar/OptionPricing.fut (`main`, :295-322) for
 * liboptionpricing.so.  This file checks the shapes the reference's types
 * and size coercions impose, manages device memory through the context's
 * pool and launches csrc/op_kernels.hip on the context's stream.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/optionpricing.h"
#include "lvc_ndarray.h"
#include "lvc_runtime.h"
#include "op_internal.h"

/* The kernels take no tuning parameters. */
const lvc_program lvc_prog = {0, NULL, NULL, NULL, NULL};

/* The [num_dates][num_und] shape contract c's payoff coerces to
 * (:289-293) and the discounts it reads (payoff1 :244, payoff2 :265,
 * payoff3 :279-282); 0 for a contract without a payoff (:294). */
static int contract_shape(int32_t c, int64_t *D, int64_t *U, int64_t *K) {
  switch (c) {
    case 1: *D = 1; *U = 1; *K = 1; return 1;
    case 2: *D = 5; *U = 3; *K = 5; return 1;
    case 3: *D = 367; *U = 3; *K = 2; return 1;
    default: return 0;
  }
}

static int shape_is(const lvc_ndarray *a, int rank, int64_t d0, int64_t d1, int64_t d2) {
  return a->shape[0] == d0 && (rank < 2 || a->shape[1] == d1) &&
         (rank < 3 || a->shape[2] == d2);
}

int futhark_entry_main(struct futhark_context *ctx, struct futhark_f32_1d **out0,
                       const int32_t contract_number, const int32_t num_mc_it,
                       const struct futhark_i32_2d *dir_vs,
                       const struct futhark_f32_3d *md_cs,
                       const struct futhark_f32_3d *md_vols,
                       const struct futhark_f32_3d *md_drifts,
                       const struct futhark_f32_2d *md_sts,
                       const struct futhark_f32_2d *md_detvals,
                       const struct futhark_f32_2d *md_discts,
                       const struct futhark_i32_2d *bb_inds,
                       const struct futhark_f32_2d *bb_data) {
  if (!out0)
    return 1;
  *out0 = NULL;
  if (!lvc_ready(ctx))
    return 1;
  if (!dir_vs || !md_cs || !md_vols || !md_drifts || !md_sts || !md_detvals ||
      !md_discts || !bb_inds || !bb_data) {
    lvc_set_error(ctx, "main: NULL input array");
    return 1;
  }
  /* the size parameters of main (:296-306) */
  const int64_t M = md_cs->a.shape[0], U = md_cs->a.shape[1];
  const int64_t D = bb_inds->a.shape[1], B = dir_vs->a.shape[1];
  const int64_t K = md_discts->a.shape[1];
  const char *bad = NULL;
  if (!shape_is(&md_cs->a, 3, M, U, U))
    bad = "md_cs";
  else if (!shape_is(&bb_inds->a, 2, 3, D, 0))
    bad = "bb_inds";
  else if (!shape_is(&bb_data->a, 2, 3, D, 0))
    bad = "bb_data";
  else if (!shape_is(&dir_vs->a, 2, D * U, B, 0))
    bad = "dir_vs";
  else if (!shape_is(&md_vols->a, 3, M, D, U))
    bad = "md_vols";
  else if (!shape_is(&md_drifts->a, 3, M, D, U))
    bad = "md_drifts";
  else if (!shape_is(&md_sts->a, 2, M, U, 0))
    bad = "md_sts";
  else if (!shape_is(&md_detvals->a, 2, M, 1, 0))
    bad = "md_detvals";
  else if (!shape_is(&md_discts->a, 2, M, K, 0))
    bad = "md_discts";
  if (bad) {
    lvc_set_error(ctx, "main: %s does not have the shape the other arguments fix "
                  "(num_models %lld, num_und %lld, num_dates %lld)",
                  bad, (long long)M, (long long)U, (long long)D);
    return 1;
  }
  if (num_mc_it < 0) {
    lvc_set_error(ctx, "main: num_mc_it is %d: iota of a negative count (:308)",
                  num_mc_it);
    return 1;
  }
  int64_t cD, cU, cK;
  const int priced = contract_shape(contract_number, &cD, &cU, &cK);
  if (priced && (D != cD || U != cU)) {
    lvc_set_error(ctx, "main: contract %d takes [%lld][%lld] (num_dates, num_und), "
                  "got [%lld][%lld] (:289-293)", contract_number, (long long)cD,
                  (long long)cU, (long long)D, (long long)U);
    return 1;
  }
  if (priced && M > 0 && K < cK) {
    lvc_set_error(ctx, "main: contract %d reads %lld discounts per model, md_discts "
                  "has %lld", contract_number, (long long)cK, (long long)K);
    return 1;
  }
  if (B > OP_DV_STRIDE) {
    lvc_set_error(ctx, "main: %lld direction-vector bits (at most 32: testBit shifts "
                  "an i32, :17-18)", (long long)B);
    return 1;
  }
  if (M > (1 << 24)) {
    lvc_set_error(ctx, "main: %lld market models (at most 2^24)", (long long)M);
    return 1;
  }
  const int launch = priced && num_mc_it > 0 && M > 0;
  pthread_mutex_lock(&ctx->lock);
  void *out = NULL, *dv = NULL, *partial = NULL;
  /* the padded direction vectors, then the bridge check's flag */
  const size_t dv_bytes = (size_t)(D * U) * OP_DV_STRIDE * sizeof(int32_t) + sizeof(int32_t);
  const size_t part_bytes = (size_t)op_blocks(num_mc_it) * (size_t)M * sizeof(double);
  int rc = lvc_dev_alloc(ctx, (size_t)M * sizeof(float), &out);
  if (!rc && launch)
    rc = lvc_dev_alloc(ctx, dv_bytes, &dv) || lvc_dev_alloc(ctx, part_bytes, &partial);
  if (!rc && launch) {
    op_args a;
    memset(&a, 0, sizeof(a));
    a.bits = (int32_t)B;
    a.cs = md_cs->a.mem;
    a.vols = md_vols->a.mem;
    a.drifts = md_drifts->a.mem;
    a.sts = md_sts->a.mem;
    a.detvals = md_detvals->a.mem;
    a.discts = md_discts->a.mem;
    a.K = (int32_t)K;
    a.M = (int32_t)M;
    a.bbi = bb_inds->a.mem;
    a.bbd = bb_data->a.mem;
    a.N = num_mc_it;
    int tok = lvc_prof_begin(ctx, "op_paths");
    int e = op_launch((void *)ctx->stream, contract_number, &a, dir_vs->a.mem, dv,
                      partial, out);
    lvc_prof_end(ctx, tok);
    rc = lvc_check_hip(ctx, e, "option pricing kernels");
  } else if (!rc && M > 0) {
    /* no payoff to sum (:294) or no path: the sum is 0, divided by
     * num_mc_it (:322), which makes NaN for no path */
    float *host = malloc((size_t)M * sizeof(float));
    if (!host) {
      lvc_set_error(ctx, "host allocation failed");
      rc = 1;
    } else {
      for (int64_t m = 0; m < M; m++)
        host[m] = 0.0f / (float)num_mc_it;
      rc = lvc_upload(ctx, out, host, (size_t)M * sizeof(float));
      free(host);
    }
  }
  if (partial)
    lvc_dev_free(ctx, partial, part_bytes);
  if (dv)
    lvc_dev_free(ctx, dv, dv_bytes);
  struct futhark_f32_1d *arr = NULL;
  if (!rc) {
    arr = lvc_array_adopt_f32(out, M);
    if (!arr) {
      lvc_set_error(ctx, "host allocation failed");
      rc = 1;
    }
  }
  if (rc && out)
    lvc_dev_free(ctx, out, (size_t)M * sizeof(float));
  pthread_mutex_unlock(&ctx->lock);
  if (rc)
    return 1;
  *out0 = arr;
  return 0;
}
