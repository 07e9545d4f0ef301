/*
 * Internal interface between the C host code and the HIP kernels: device
 * data layouts and the launcher functions (plain C linkage, HIP types only
 * as opaque stream handles).
 *
 * Device data of one set of problems (a "batch"):
 *   - lvc_problem[n]   descriptors (sizes, parameters, offsets, step index);
 *   - geometry         per problem, at geo_off (even), LVC_GEO_LEN(nx, ny)
 *     doubles: an FP64 section of LVC_GEO_F64(nx, ny) doubles
 *       x[nx] | ax[nx] | cx[nx] | y[ny] | ay[ny] | cy[ny]
 *     followed by an FP32 section (16-byte aligned) holding the same taps
 *     rounded to float, each table padded to a multiple of 4:
 *       axf[P4(nx)] | cxf[P4(nx)] | ayf[P4(ny)] | cyf[P4(ny)]
 *     where ax/cx are -0.25 * x^(2 beta) * dxx0/dxx2 (the x operator's off-
 *     diagonal taps without the time factor exp(2y - nu^2 t)) and ay/cy are
 *     -0.25 * nu^2 * dyy0/dyy2 (the timeline t*k/(nt-1) is computed where it
 *     is used);
 *   - state R (real)   per problem, at state_off, [ny][nx] row-major;
 *   - state Y (real)   same layout, the x-sweep's output.
 * "real" is float or double according to the precision of the batch.
 */
#ifndef LVC_INTERNAL_H
#define LVC_INTERNAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum lvc_precision {
  LVC_FP32 = 0,        /* stencils and sweeps in FP32 */
  LVC_FP32_REFINE = 1, /* FP32 + one FP64-residual refinement per solve */
  LVC_FP64 = 2         /* everything FP64 */
};

typedef struct lvc_problem {
  int32_t nx, ny, nt;
  int32_t xindex, yindex;  /* initGrid's price indexes (lvc_grid_scalars) */
  int32_t k;               /* next time index to process; < 0 when done */
  int32_t status;          /* 0 = ok, 3 = price index out of range */
  int32_t ytiles;          /* y-sweep workgroups of this problem per step */
  int32_t ydone;           /* y-sweep workgroups finished in this step */
  int32_t fac_src;         /* problem whose y factor table this one uses */
  int32_t grid32;          /* initGrid's scalars in f32 (the reference's f32 entry) */
  int32_t out_index;       /* where the gather writes this problem's price */
  int32_t pad_;
  int64_t state_off;       /* element offset of R and Y */
  int64_t geo_off;         /* element offset into the f64 geometry buffer */
  double s0, strike, t, alpha, nu, beta;
  /* initGrid's scalars (LocVolCalib.fut:15-20), set on the host by
   * lvc_grid_scalars: the spacings and the value of x[0] whose log VarX
   * takes (:63).  With grid32 they are the reference's f32 values, widened. */
  double dx, dy, x0log;
} lvc_problem;

/* One workgroup's share of a sweep: `count` rows (x sweep) or columns (y
 * sweep) of problem `problem`, starting at `first`, with `lanes` lanes (a
 * power of two) per row/column (column y sweep: segments per column). */
typedef struct lvc_tile {
  int32_t problem;
  int32_t first;
  int32_t count;
  int32_t lanes;
  /* x sweep: problems problem .. problem+cpl-1, identical but for the strike,
   * sharing the row factorisations; lane-segment y sweep: columns per lane
   * (sharing the factorisation); column y sweep: 64-column groups */
  int32_t cpl;
  int32_t pad_;
} lvc_tile;

#define LVC_P4(n) (((int64_t)(n) + 3) & ~(int64_t)3)
#define LVC_GEO_F64(nx, ny) ((3 * (int64_t)(nx) + 3 * (int64_t)(ny) + 1) & ~(int64_t)1)
#define LVC_GEO_LEN(nx, ny) (LVC_GEO_F64(nx, ny) + LVC_P4(nx) + LVC_P4(ny))
#define LVC_BLOCK 256 /* threads per workgroup of the lane-segment sweeps */
#define LVC_COL_BLOCK 512 /* threads per workgroup of the column y sweep */
#define LVC_COL_ROWS 512  /* rows of a column system: 8 segments of <= 64 */
/* real elements of one problem's y factor table (column y sweep) */
#define LVC_YFAC_LEN (6 * LVC_COL_ROWS + 6 * (LVC_COL_BLOCK / 64))
#define LVC_MAX_N 1024 /* largest numX / numY accepted (64 lanes x 16) */

/* Layout of the reference's `main` (outer options sharing everything but the
 * strike) in one contiguous device region, as the allocation-free entry
 * builds it: descriptors, x tiles (groups of `xshare` options, `xtpg` row
 * tiles each) then y tiles (`ytpo` per option), factor-group leaders,
 * geometry, R, Y and the y factor tables.  The same tiles, in the same order,
 * as the batch engine builds for those options. */
typedef struct lvc_main_layout {
  int32_t outer, xshare, groups;
  int32_t xtpg, rows, gx;        /* x: tiles per group, rows per tile, lanes per row */
  int32_t ytpo, cols, gy, ycpl;  /* y: tiles per option, columns per tile, lanes
                                    (segments) per column, columns per lane (groups) */
  int32_t strike_f32;            /* strikes 0.001f*i in f32 (else 0.001*i in f64) */
  int32_t pad_;
  int64_t geo_stride, state_stride; /* per option, in doubles / in reals */
} lvc_main_layout;

/* Bytes of the static device scratch of the allocation-free entry: main of
 * 128 options at 256x256 in the FP32 forms (R and Y 64 MiB, geometry, factor
 * tables, descriptors and tiles), or of 64 options in FP64. */
#define LVC_SCRATCH_BYTES ((int64_t)72 << 20)

/* Device address of `device`'s copy of the static scratch array, looked up
 * once per device with that device current (hipGetSymbolAddress) and kept.
 * Returns 0 or a hipError_t value. */
int lvc_scratch_address(int device, void **ptr);

/* Writes main's descriptors (each `tmpl` with its strike, offsets and factor
 * leader), tiles and factor-group leaders for layout `lay`. */
int lvc_launch_main_init(void *stream, const lvc_problem *tmpl,
                         const lvc_main_layout *lay, lvc_problem *probs,
                         lvc_tile *tiles, int32_t *fprobs);

/* initGrid's scalars of one problem (LocVolCalib.fut:15-20) into q->dx, dy,
 * x0log, xindex, yindex and q->status (3 when myXindex is outside the grid,
 * the reference's failing bounds check at :236).  grid32 = 1 evaluates
 * stdX = 20*alpha*s0*sqrt(t), dx = stdX/numX, myXindex = i32.f32(s0/dx) and
 * x[0] in f32, as the reference's f32 program does (the parameters are the
 * f32 values the entry received); grid32 = 0 in FP64.  The grid arrays are
 * then built in FP64 from these scalars by the setup kernel. */
void lvc_grid_scalars(lvc_problem *q, int grid32);

/* Supported segment lengths (elements of one system per lane): 4, 8, 16 for
 * both sweeps (the lanes of a system are lanes of one wavefront, reduced
 * system by PCR); 32 and 64 for the y sweep only (one lane per column
 * segment, the segments of a column in different wavefronts, reduced system
 * by Thomas). */
int lvc_seg_supported(int seg, int ysweep);

/* Launchers: return 0 or a hipError_t value.  `stream` is a hipStream_t. */
int lvc_launch_setup(void *stream, int precision, lvc_problem *probs, int nprob,
                     double *geo, void *R);
/* x sweep; with yseg = a column segment length (32, 64) it also writes the
 * y factor table for the step of each of the nfac problems fprobs[] into
 * yfac (real elements, LVC_YFAC_LEN per problem, indexed by problem).
 * xshared: some tiles cover several problems (tile.cpl > 1); otherwise the
 * single-problem variant runs (fewer registers). */
int lvc_launch_xsweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *R, void *Y,
                      void *yfac, int yseg, const int32_t *fprobs, int nfac,
                      int xshared);
/* y sweep; column segments (seg >= 32) read the factor tables in yfac. */
int lvc_launch_ysweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *Y, void *R,
                      int lds_bytes, const void *yfac);
/* The price of each of the nprob problems into out[problem.out_index]. */
int lvc_launch_gather(void *stream, int precision, lvc_problem *probs,
                      int nprob, const void *R, void *out, int out_f64);
/* The prices of problems list[0 .. nprob-1] into out[0 .. nprob-1]. */
int lvc_launch_gather_list(void *stream, int precision, lvc_problem *probs,
                           const int32_t *list, int nprob, const void *R, void *out,
                           int out_f64);
/* Sets up problems list[0 .. nprob-1] from the descriptors src[0 .. nprob-1]
 * (stored to probs[list[i]]). */
int lvc_launch_setup_list(void *stream, int precision, lvc_problem *probs,
                          const lvc_problem *src, const int32_t *list, int nprob,
                          double *geo, void *R);

/* Batched f32 tridiagonal solves of [n][m] systems (the ocean-sim entries):
 * form 0 = the scan form (tridagPar), 1 = Thomas (tridagSeq).  scratch holds
 * n*m floats and is needed when m > lvc_tridag_lds_rows(); systems of up to
 * lvc_tridag_staged_rows() unknowns keep their forward values in registers. */
int lvc_launch_tridag(void *stream, int form, const float *a, const float *b,
                      const float *c, const float *y, float *out, float *scratch,
                      int64_t n, int m);
int lvc_tridag_lds_rows(void);
int lvc_tridag_staged_rows(void);

/* Lanes per row/column and rows/columns per workgroup for a system of
 * length n split in segments of `seg`. */
static inline int lvc_lanes_for(int n, int seg) {
  int P = (n + seg - 1) / seg, G = 1;
  while (G < P)
    G <<= 1;
  return G;
}

/* Column y sweep (seg >= 32): segments per column and whether ny fits. */
static inline int lvc_seg_is_col(int seg) { return seg >= 32; }
static inline int lvc_yseg_fits(int ny, int seg) {
  if (lvc_seg_is_col(seg))
    return (ny + seg - 1) / seg <= LVC_COL_BLOCK / 64;
  return lvc_lanes_for(ny, seg) <= 64;
}

/* Shape of a y-sweep tile for columns of length ny.  Column sweep (seg >= 32):
 * G = ceil(ny/seg) segments (wavefronts) per column, cpl = column groups of
 * 64 columns per workgroup, cols = 64 * cpl.  Lane-segment sweep: G lanes per column,
 * cpl columns per lane, cols = cpl * (LVC_BLOCK / G) columns per workgroup
 * (aiming at a 128-byte row segment: 32 floats or 16 doubles), with the LDS
 * panel of G*seg rows x (cols + 1) kept within 64 KiB.  cols is a power of
 * two dividing LVC_BLOCK. */
static inline void lvc_ytile_shape(int ny, int seg, int real, int *G, int *cpl,
                                   int *cols) {
  if (lvc_seg_is_col(seg)) {
    const int P = (ny + seg - 1) / seg;
    *G = P;
    *cpl = (LVC_COL_BLOCK / 64) / P;
    *cols = 64 * *cpl;
    return;
  }
  const int g = lvc_lanes_for(ny, seg);
  const int lc = LVC_BLOCK / g;
  int c = 1; /* a power of two, so that cols divides LVC_BLOCK */
  while (2 * c * lc <= (real == 4 ? 32 : 16))
    c *= 2;
  while (c > 1 && (long long)g * seg * (c * lc + 1) * real > 65536)
    c /= 2;
  *G = g;
  *cpl = c;
  *cols = c * lc;
}

/* Dynamic LDS bytes of a y-sweep tile of that shape (the column sweep's LDS
 * is static). */
static inline int lvc_ytile_lds(int ny, int seg, int real) {
  if (lvc_seg_is_col(seg))
    return 0;
  int G, cpl, cols;
  lvc_ytile_shape(ny, seg, real, &G, &cpl, &cols);
  return G * seg * (cols + 1) * real;
}

#ifdef __cplusplus
}
#endif

#endif /* LVC_INTERNAL_H */
