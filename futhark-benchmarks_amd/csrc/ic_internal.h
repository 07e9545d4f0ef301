/* InterestCalib: the launch interface between ic_entry.c and ic_kernels.hip
 * (not exported). */
#ifndef IC_INTERNAL_H
#define IC_INTERNAL_H

#include <stdint.h>

#define IC_MAX_PERIODS 512 /* schedule periods of one swaption (12 * term / freq) */
#define IC_LAUNCHES 2      /* evaluation launches: 1-2 periods a lane, 4-8 periods a lane */
#define IC_MAX_BITS 31     /* Sobol direction-vector bits (sobolInd's i32 shifts) */

#ifdef __cplusplus
extern "C" {
#endif

/* One swaption quote.  The host fills n, off, lanes and per_lane; ic_launch_setup
 * fills the rest, which does not depend on the genome. */
typedef struct {
  int32_t n;        /* schedule periods */
  int32_t off;      /* its first period in the period table */
  int32_t lanes;    /* lanes per evaluation (a power of two, 1..64) */
  int32_t per_lane; /* periods per lane (1, 2, 4 or 8); lanes * per_lane >= n */
  float tmat0;   /* date_act_365(maturity, today) */
  float zc_mat;  /* zc(maturity) */
  float quote;   /* the Black price (:509-525) */
  float pad1;
} ic_swaption;

/* One period of a swaption's schedule: (date_act_365(end, today),
 * date_act_365(end, maturity), c_i * zc(end) / zc(maturity), its log). */
typedef struct {
  float tod1, tod2, fact, log_fact;
} ic_period;

/* One wavefront's work: a swaption and the first of the 64 / lanes genomes
 * it evaluates. */
typedef struct {
  int32_t swaption, genome0;
} ic_work;

/* The wavefronts of the evaluation launches: [0] swaptions of 1 or 2
 * periods a lane, [1] of 4 or 8 (each wavefront's class is its swaption's
 * per_lane). */
typedef struct {
  const ic_work *work[IC_LAUNCHES];
  int32_t nwork[IC_LAUNCHES];
} ic_worklists;

/* The Sobol direction vector (:190-202). */
typedef struct {
  int32_t bits;
  int32_t dir[IC_MAX_BITS];
} ic_dirvec;

/* One MCMC move (:655-724), fixed by Sobol numbers at data-independent
 * offsets: type 1 all dimensions, 2 one dimension, 3 differential
 * evolution, 0 none; base is the Sobol offset of the move's first draw
 * after the type draw. */
typedef struct {
  int32_t type, base, dim, pad;
} ic_move;

/* Where ic_launch_step records the trajectory (NULL: not recorded):
 * genomes [conv+1][pop][5] and logliks [conv+1][pop] after the initial
 * evaluation and after each iteration's accept step; props [conv][pop][5],
 * fb [conv][pop] and newl [conv][pop], the proposals each iteration
 * evaluates, their ratios and their log-likelihoods. */
typedef struct {
  float *genomes, *logliks, *props, *fb, *newl;
} ic_trace;

enum { IC_OUT_TERMS = 0, IC_OUT_PRICES = 1, IC_OUT_SUMMARY = 2 };

/* Genome-independent swaption data: nswap rows of `stride` floats at
 * swaptions; sw[s].n, off, lanes and per_lane set.  Returns 0 or a hipError_t value. */
int ic_launch_setup(void *stream, const float *swaptions, int64_t stride, int32_t nswap,
                    ic_swaption *sw, ic_period *periods);

/* The initial population (:633-638): genomes and props [pop][5]. */
int ic_launch_init(void *stream, int32_t pop, const ic_dirvec *dv, float *genomes, float *props);

/* evalGenomeOnSwap (:484-602) of genomes [npop][5] on the swaptions of the
 * work lists.
 * IC_OUT_TERMS: out0 [nswap][npop] the log-likelihood terms (:606-625);
 * IC_OUT_PRICES: out0, out1 [npop][nswap] quotes and prices;
 * IC_OUT_SUMMARY (npop 1): out0 [nswap][3] makeSummary's rows (:205-213). */
int ic_launch_eval(void *stream, const ic_swaption *sw, const ic_period *periods,
                   const ic_worklists *wl, const float *genomes, int32_t npop,
                   const float *herm_x, const float *herm_w, int32_t nherm, int32_t nswap,
                   int mode, float *out0, float *out1);

/* Two launches: the log-likelihood of each genome, the sum of its terms
 * over the swaptions in a fixed order (32 threads per genome each sum the
 * swaptions t, t + 32, ... in order, and the partial sums are added
 * pairwise), then, with init, logliks := those sums, else the accept step
 * (:733-745) with draws from Sobol offset acc; then, unless next is no
 * move, the proposals of move `next` into props (ratios into fb).  it is
 * the iteration whose accept step this is (-1 for init), for the trace. */
int ic_launch_step(void *stream, int32_t pop, int32_t nswap, const float *terms, float *genomes,
                   float *props, float *logliks, float *fb, int init, int32_t acc,
                   const ic_move *next, const ic_dirvec *dv, int32_t it, int32_t conv,
                   const ic_trace *trace);

/* The winner (:752-756, the first of the largest log-likelihoods) into
 * out [8]: a, b, sigma, nu, rho, its log-likelihood (:849), and its genome
 * into winner [5]. */
int ic_launch_winner(void *stream, int32_t pop, const float *genomes, const float *logliks,
                     float *out, float *winner);

#ifdef __cplusplus
}
#endif

#endif /* IC_INTERNAL_H */
