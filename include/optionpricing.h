/*
 * optionpricing.h -- C ABI of the MI355X OptionPricing library
 * (liboptionpricing.so), a drop-in for the C API that
 * `futhark <backend> --library` generates for finpar/OptionPricing.fut.
 * It is a library of its own, as the generator makes one per program: the
 * context API (futhark_context.h) and the arrays (futhark_arrays.h) have the
 * same names as in liblocvolcalib.so, and `futhark_entry_main` is this
 * program's.
 *
 * Results: the per-path payoffs are computed in f32 in the reference's
 * order of operations (Sobol numbers, inverse Gaussian, Brownian bridge,
 * Black-Scholes paths, payoff); their sum over the paths, which the
 * reference leaves to its backend's reduction (:318), is accumulated in f64
 * in a fixed order and divided by num_mc_it once, so a result does not
 * depend on the device or the launch shape.
 */
#ifndef OPTIONPRICING_H
#define OPTIONPRICING_H

#include <stdint.h>

#include "futhark_arrays.h"
#include "futhark_context.h"

#ifdef __cplusplus
extern "C" {
#endif

/* OptionPricing.fut:295-322 (`main`): the price of contract
 * `contract_number` (1, 2 or 3) under each of the num_models market models,
 * averaged over num_mc_it quasi-Monte-Carlo paths.  Shapes (the reference's
 * size parameters): dir_vs [num_dates * num_und][num_bits], md_cs
 * [num_models][num_und][num_und], md_vols and md_drifts
 * [num_models][num_dates][num_und], md_sts [num_models][num_und],
 * md_detvals [num_models][1], md_discts [num_models][num_discts], bb_inds
 * and bb_data [3][num_dates]; out0 [num_models].  Contract 1 takes
 * num_dates = num_und = 1, contract 2 [5][3], contract 3 [367][3] (the
 * reference's size coercions at :289-293); another shape for contracts
 * 1-3, num_mc_it < 0 or num_bits > 32 (testBit shifts an i32, :17-18) is an
 * error.  As the reference: a contract number other than 1-3 prices 0 for
 * every model (:294), num_mc_it = 0 gives NaN (0/0, :322), no model gives
 * an empty array.  Any number of models (taken 8 at a time on the device). */
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
                       const struct futhark_f32_2d *bb_data);

#ifdef __cplusplus
}
#endif

#endif /* OPTIONPRICING_H */
