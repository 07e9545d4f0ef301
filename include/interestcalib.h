/*
 * interestcalib.h -- C ABI of the MI355X InterestCalib library
 * (libinterestcalib.so), a drop-in for the C API that
 * `futhark <backend> --library` generates for finpar/InterestCalib.fut.
 * It is a library of its own, as the generator makes one per program: the
 * context API (futhark_context.h) and the arrays (futhark_arrays.h) have the
 * same names as in liblocvolcalib.so, and `futhark_entry_main` is this
 * program's.
 *
 * Results: every genome is evaluated on every swaption in f32 in the
 * reference's order of operations (dates, Black quote, G2++ terms, the
 * exactYhat root finds, the Gauss-Hermite sum), with the device's
 * expf/logf, so a log-likelihood differs from a host build's in its last
 * bits.  The MCMC accept step compares exp(new - old) with a Sobol number
 * (:733-745); where the two sides are within such a difference, a move can
 * go the other way than on another backend, and the trajectory, and the
 * winner with it, follow.  The reference's own datasets show this: their
 * expected winners are one of several optima of nearly equal likelihood.
 */
#ifndef INTERESTCALIB_H
#define INTERESTCALIB_H

#include <stdint.h>

#include "futhark_arrays.h"
#include "futhark_context.h"

#ifdef __cplusplus
extern "C" {
#endif

/* InterestCalib.fut:929-936 (`main`): calibrates the five G2++ parameters
 * to the swaption quotes.  pop genomes run mcmc_conv iterations of the
 * genetic / differential-evolution MCMC of interestCalibKernel (:627-849);
 * out0..out5 are the winner's a, b, sigma, nu, rho and log-likelihood
 * (:849) and out6 the [n][3] summary (10000 * model price, 10000 * Black
 * price, 100 * |relative error|) per swaption (:205-213).  Shapes:
 * swaptions [n][4] (maturity in years, payment frequency in months, term
 * in years, Black volatility; further columns are ignored), hermCoefs and
 * hermWeights [m], sobDirVct [bits].  l, ll and lll are unused, as in the
 * reference.  Errors, where the reference fails: pop < 1 (the winner
 * indexes an empty population), pop = 1 with a differential-evolution
 * move in the schedule (genomes[-1]), a swaption row shorter than 4, a
 * negative schedule length (12 * term / frequency), hermCoefs and
 * hermWeights of different lengths (zip).  Limits of this build: at most
 * 512 schedule periods per swaption, 31 direction-vector bits (sobolInd's
 * i32 shifts), 2^24 genomes or iterations. */
int futhark_entry_main(struct futhark_context *ctx, float *out0, float *out1, float *out2,
                       float *out3, float *out4, float *out5, struct futhark_f32_2d **out6,
                       const int32_t pop, const int32_t mcmc_conv, const int32_t l,
                       const struct futhark_f32_2d *swaptions, const int32_t ll,
                       const struct futhark_f32_1d *hermCoefs,
                       const struct futhark_f32_1d *hermWeights, const int32_t lll,
                       const struct futhark_i32_1d *sobDirVct);

/* No counterpart in the reference; for tests and diagnostics.  The same
 * calibration as futhark_entry_main, recording its trajectory into host
 * arrays: genomes [mcmc_conv+1][pop][5] and logliks [mcmc_conv+1][pop]
 * (the population after the initial evaluation and after each iteration's
 * accept step), props [mcmc_conv][pop][5], fb [mcmc_conv][pop] and newl
 * [mcmc_conv][pop] (the proposals each iteration evaluates, their
 * forward/backward ratios and their log-likelihoods), out [6] and summary
 * [n][3] (main's results).  Synchronous. */
int interestcalib_trace(struct futhark_context *ctx, const int32_t pop, const int32_t mcmc_conv,
                        const struct futhark_f32_2d *swaptions,
                        const struct futhark_f32_1d *hermCoefs,
                        const struct futhark_f32_1d *hermWeights,
                        const struct futhark_i32_1d *sobDirVct, float *genomes, float *logliks,
                        float *props, float *fb, float *newl, float *out, float *summary);

/* No counterpart in the reference; for tests and diagnostics.
 * evalGenomeOnSwap (:484-602) of ngenomes genomes (host [ngenomes][5],
 * each a, b, rho, nu, sigma) on every swaption: quotes and prices (host
 * [ngenomes][n]) and each genome's log-likelihood (host [ngenomes], the
 * sum of :639-645 in swaption order).  Synchronous. */
int interestcalib_eval(struct futhark_context *ctx, const int32_t ngenomes, const float *genomes,
                       const struct futhark_f32_2d *swaptions,
                       const struct futhark_f32_1d *hermCoefs,
                       const struct futhark_f32_1d *hermWeights, float *quotes, float *prices,
                       float *logliks);

#ifdef __cplusplus
}
#endif

#endif /* INTERESTCALIB_H */
