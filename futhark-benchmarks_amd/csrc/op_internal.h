/* OptionPricing: the launch interface between op_entry.c and op_kernels.hip
 * (not exported). */
#ifndef OP_INTERNAL_H
#define OP_INTERNAL_H

#include <stdint.h>

#define OP_MAX_MODELS 8
#define OP_MAX_BLOCKS 2048 /* workgroups of the path kernel at most */
#define OP_DV_STRIDE 32    /* ints per padded direction-vector row */

#ifdef __cplusplus
extern "C" {
#endif

/* Device pointers to the market data and the Brownian-bridge tables, as the
 * entry point received them (row-major), and the sizes. */
typedef struct {
  int32_t bits;        /* direction-vector bits */
  const float *cs, *vols, *drifts, *sts, *detvals, *discts;
  int32_t K, M;        /* discounts per model, models */
  const int32_t *bbi;  /* [3][dates] */
  const float *bbd;    /* [3][dates] */
  int64_t N;           /* paths */
} op_args;

/* Workgroups of the path kernel for N paths (OP_MAX_BLOCKS at most): the
 * f64 partial sums it writes are [op_blocks(N)][M]. */
int op_blocks(int64_t paths);

/* Pads dir_vs ([dims][bits]) into dv ([dims][OP_DV_STRIDE] ints, then one
 * more: whether the bridge tables let the paths skip zeroing their rows),
 * runs the paths and writes the mean payoff per model to out[M]; every launch on
 * `stream` (a hipStream_t).  The contract fixes dims (1, 15 or 1101).
 * Returns 0 or a hipError_t value. */
int op_launch(void *stream, int contract, const op_args *A, const int32_t *dir_vs,
              int32_t *dv, double *partial, float *out);

#ifdef __cplusplus
}
#endif

#endif /* OP_INTERNAL_H */
