/*
 * A C caller of the library in the shape of the reference's own callers of
 * the generated API (the vendored lys frontends: shared.c:15-49,
 * console/liblys.c:243-251, console/main.c:176-180, shared.h:20-28):
 * configuration, context, one entry call, values, sync, frees, compiled
 * and run by the GPU tests.
 *
 * Usage: lys_caller outer numX numY numT s0 t alpha nu beta
 * Prints the prices one per line; then checks that a bad size fails with a
 * message and that the context stays usable.  Exit status 0 when all of it
 * behaved as the generated API does.
 */
#include <stdio.h>
#include <stdlib.h>

#include "locvolcalib.h"

static int report(struct futhark_context *ctx, const char *what) {
  char *e = futhark_context_get_error(ctx); /* shared.h:25 */
  fprintf(stderr, "%s: %s\n", what, e ? e : "(no message)");
  free(e);
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 10) {
    fprintf(stderr, "usage: %s outer numX numY numT s0 t alpha nu beta\n", argv[0]);
    return 2;
  }
  const int outer = atoi(argv[1]), nx = atoi(argv[2]), ny = atoi(argv[3]), nt = atoi(argv[4]);
  const float s0 = strtof(argv[5], NULL), t = strtof(argv[6], NULL);
  const float alpha = strtof(argv[7], NULL), nu = strtof(argv[8], NULL);
  const float beta = strtof(argv[9], NULL);

  struct futhark_context_config *cfg = futhark_context_config_new(); /* shared.c:20 */
  futhark_context_config_set_device(cfg, "0");                       /* shared.c:25 */
  struct futhark_context *ctx = futhark_context_new(cfg);            /* shared.c:47 */
  char *err = futhark_context_get_error(ctx);
  if (err) {
    fprintf(stderr, "context: %s\n", err);
    free(err);
    futhark_context_free(ctx);
    futhark_context_config_free(cfg);
    return 1;
  }

  int rc = 0;
  struct futhark_f32_1d *prices = NULL;
  if (futhark_entry_main(ctx, &prices, outer, nx, ny, nt, s0, t, alpha, nu, beta) != 0) {
    rc = report(ctx, "main"); /* LocVolCalib.fut:238-246 */
  } else {
    const int64_t n = futhark_shape_f32_1d(ctx, prices)[0];
    float *host = malloc((size_t)(n ? n : 1) * sizeof(float));
    if (!host || n != outer || futhark_values_f32_1d(ctx, prices, host) != 0 || /* liblys.c:248 */
        futhark_context_sync(ctx) != 0) {                                       /* liblys.c:249 */
      rc = report(ctx, "values");
    } else {
      for (int64_t i = 0; i < n; i++)
        printf("%.9g\n", host[i]);
    }
    free(host);
    futhark_free_f32_1d(ctx, prices); /* liblys.c:250 */
  }

  /* numX < 3 fails with a message (the range 1...n-2 of LocVolCalib.fut:37),
   * and the context stays usable */
  struct futhark_f32_1d *bad = NULL;
  if (futhark_entry_main(ctx, &bad, 4, 2, ny, nt, s0, t, alpha, nu, beta) == 0) {
    fprintf(stderr, "numX = 2 was accepted\n");
    futhark_free_f32_1d(ctx, bad);
    rc = 1;
  } else {
    char *e = futhark_context_get_error(ctx);
    if (!e) {
      fprintf(stderr, "numX = 2 failed without a message\n");
      rc = 1;
    }
    free(e);
  }
  struct futhark_f32_1d *again = NULL;
  if (futhark_entry_main(ctx, &again, 2, nx, ny, 3, s0, t, alpha, nu, beta) != 0)
    rc = report(ctx, "main after an error");
  else
    futhark_free_f32_1d(ctx, again);

  futhark_context_free(ctx);        /* main.c:179 */
  futhark_context_config_free(cfg); /* main.c:180 */
  return rc;
}
