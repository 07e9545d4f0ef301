"""The OptionPricing oracle (oracle/op_oracle.c) pinned against the
reference's datasets (finpar/OptionPricing-data, copied verbatim into
tests/golden/OptionPricing-data) and against what the mathematics fixes:

* the whole pipeline: the reference's expected outputs are the f32 sum of the
  per-path payoffs taken in path order, divided by num_mc_it
  (OptionPricing.fut:318-322 on a sequential backend), and the oracle's
  payoffs summed that way reproduce small and medium bit for bit;
* the Sobol numbers (:12-35) against the Gray-code recurrence
  x_n = x_(n-1) ^ v[ctz n] and, for the van der Corput direction vectors of
  dimension 0, against the dyadic points that each block of 2^k indices
  covers;
* the inverse Gaussian (:40-140) against scipy's ndtri;
* the Brownian bridge (:145-183) by the covariance of the path it builds:
  with W = cumsum(bridge(g)) = L g, L L^T = min(t_i, t_j).

* Black-Scholes and the three payoffs in closed form: with zero
  volatilities every path's prices are st * exp(cumulative drift), and
  drifts chosen per contract decide the payoff branch.

Five planted mistakes in the oracle (a dropped Gray code, a sign, a swapped
weight, the correlation summed over a whole row, an off-by-one date) are
each caught.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle
from locvolcalib import futhark_values as fv
from locvolcalib import optionpricing as op

DATA = os.path.join(GOLDEN, "OptionPricing-data")


def load(name):
    args = op.main_arguments(fv.load(os.path.join(DATA, f"{name}.in")))
    expected = np.asarray(fv.load(os.path.join(DATA, f"{name}.out"))[0], np.float32)
    return args, expected


def sequential_mean(payoffs, n):
    """The reference's reduction on a sequential backend: f32 sum in path
    order, then the f32 division of :322."""
    return np.cumsum(payoffs, axis=1, dtype=np.float32)[:, -1] / np.float32(n)


def check_expected_output(name):
    """small and medium bit-exact; large (contract 3, 367 dates): the
    sequential f32 sum lands within 4e-6 of the expected output (which a
    different libm or reduction order produced)."""
    args, expected = load(name)
    pay = oracle.op_payoffs(*args)
    assert pay.shape == (expected.shape[0], args[1])
    got = sequential_mean(pay, args[1])
    if name == "large":
        np.testing.assert_allclose(got, expected, rtol=1e-5)
    else:
        np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("name", ["small", "medium", "large"])
def test_payoffs_reproduce_expected_output(name):
    check_expected_output(name)


@pytest.mark.parametrize("name", ["small", "medium", "large"])
def test_f64_mean_is_the_sum_of_the_payoffs(name):
    """op_main (the f64-accumulated mean the GPU is checked against) equals
    the f64 mean of op_payoffs, and stays within the f32 summation error of
    the expected output (1.1e-3 relative on large)."""
    args, expected = load(name)
    m = oracle.op_main(*args)
    pay = oracle.op_payoffs(*args)
    np.testing.assert_allclose(m, (pay.astype(np.float64).sum(axis=1) / args[1]).astype(np.float32),
                               rtol=1e-6)
    np.testing.assert_allclose(m, expected, rtol=2e-3)


def check_sobol_recurrence():
    args, _ = load("large")
    dv = args[2]  # [1101][30]
    bits = dv.shape[1]
    x = np.zeros(dv.shape[0], np.int64)
    for n in range(1, 600):
        c = (n & -n).bit_length() - 1  # ctz(n): the Gray-code bit that flips
        x ^= dv[:, c]
        np.testing.assert_array_equal(oracle.op_sobol(n, dv),
                                      (x / 2.0 ** bits).astype(np.float32), err_msg=f"n={n}")


def test_sobol_gray_code_recurrence():
    check_sobol_recurrence()


def test_sobol_van_der_corput_blocks():
    """Dimension 0 has direction vectors 2^(bits-1-b): indices 1 .. 2^k - 1
    give every multiple of 2^-k in (0, 1) once."""
    args, _ = load("small")
    dv = args[2]
    bits = dv.shape[1]
    assert list(dv[0]) == [1 << (bits - 1 - b) for b in range(bits)]
    k = 9
    pts = sorted(float(oracle.op_sobol(n, dv[:1])[0]) for n in range(1, 1 << k))
    assert pts == [j / (1 << k) for j in range(1, 1 << k)]


def check_ugaussian():
    """ugaussianEl works on dp = p - 0.5 and, in the tails, on
    pp = dp + 0.5 or 0.5 - dp, all in f32 (:121-136), so it returns the
    quantile of the probability that survives that rounding; compared with
    ndtri of that probability, the rational approximations agree to f32
    rounding (a wrong coefficient or branch bound fails by far more)."""
    from scipy.special import ndtri
    p = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 4001), np.logspace(-7, -2, 200),
                        1 - np.logspace(-7, -2, 100)]).astype(np.float32)
    got = np.array([oracle.op_ugaussian(float(v)) for v in p], np.float64)
    dp = p - np.float32(0.5)
    pp = np.where(dp < 0, dp + np.float32(0.5), np.float32(0.5) - dp).astype(np.float64)
    seen = np.where(np.abs(dp) <= 0.425, 0.5 + dp.astype(np.float64),
                    np.where(dp < 0, pp, 1.0 - pp))
    want = ndtri(seen)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
    assert err.max() < 2e-6, (err.max(), p[err.argmax()])
    assert oracle.op_ugaussian(0.5) == 0.0
    # below half an ulp of 0.5, p - 0.5 rounds to -0.5 and the tail formula
    # takes log 0: the reference's f32 arithmetic gives NaN there
    assert np.isnan(oracle.op_ugaussian(1e-30))


def test_ugaussian_is_the_normal_quantile():
    check_ugaussian()


def check_bridge_covariance(name):
    args, _ = load(name)
    bbi, bbd = args[9], args[10]
    D = bbi.shape[1]
    eye = np.eye(D, dtype=np.float32)
    A = np.stack([oracle.op_bridge(bbi, bbd, eye[k]) for k in range(D)], axis=1)
    W = np.cumsum(A.astype(np.float64), axis=0)  # the path at each date, = L g
    C = W @ W.T
    t = np.diag(C)
    assert np.all(np.diff(t) > 0)
    np.testing.assert_allclose(C, np.minimum.outer(t, t), atol=1e-6 * t.max())
    # linear: the bridge of a combination is the combination of bridges
    g = np.random.default_rng(3).standard_normal(D).astype(np.float32)
    np.testing.assert_allclose(oracle.op_bridge(bbi, bbd, g), A @ g, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("name", ["medium", "large"])
def test_bridge_covariance(name):
    check_bridge_covariance(name)


def test_contract_shape_checked():
    args, _ = load("medium")
    with pytest.raises(ValueError):
        oracle.op_main(1, 8, *args[2:])


def deterministic_case(contract):
    """Market data with zero volatilities, so every path's prices are
    st * exp(cumulative drift) whatever its Gaussians, and the drifts steer
    each contract into a payoff branch whose value has a closed form:
    contract 1 finishes in the money (payoff1, :240-244); contract 2 first
    reaches fmin >= 1 on date 1 (payoff2: 1300 * disc[1], :246-265); in
    contract 3 underlying 2 touches its barrier and ends below its start
    after a last-date jump (payoff3's goto-40 branch on date 366's prices,
    :267-283).  Returns (arguments of main with 1000 paths, closed-form
    price in f64)."""
    name = {1: "small", 2: "medium", 3: "large"}[contract]
    args, _ = load(name)
    c, _n, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd = args
    D, U = vols.shape[1], vols.shape[2]
    vols = np.zeros_like(vols)
    drifts = np.zeros((1, D, U), np.float32)
    if contract == 1:
        sts = np.array([[4100.0]], np.float32)
        detv = np.array([[1.0]], np.float32)
        drifts[0, 0, 0] = 0.01
        price = 4100.0 * np.exp(0.01)
        want = max(0.0, (price - 4000.0) * 1.0) * disc[0, 0]
    elif contract == 2:
        drifts[0, 0, :] = -0.01
        drifts[0, 1, :] = 0.05
        want = 1300.0 * disc[0, 1]
    else:
        drifts[0, :200, 2] = -0.002
        drifts[0, 200:366, 2] = 0.001
        drifts[0, 366, 2] = 0.05
        logp = np.log(sts[0].astype(np.float64)) + np.cumsum(drifts[0].astype(np.float64), axis=0)
        p = np.exp(logp)
        cond = np.any((p[:, 0] <= 2630.635) | (p[:, 1] <= 8288.0) | (p[:, 2] <= 840.0))
        last = p[366]
        goto40 = cond and (last[0] < 3758.05 or last[1] < 11840.0 or last[2] < 1200.0)
        assert goto40
        amount = 1000.0 * min(last[0] / 3758.05, last[1] / 11840.0, last[2] / 1200.0)
        want = 100.0 * disc[0, 0] + amount * disc[0, 1]
    return (c, 1000, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd), float(want)


def check_deterministic_paths(contract):
    args, want = deterministic_case(contract)
    np.testing.assert_allclose(oracle.op_main(*args), [want], rtol=1e-4)


@pytest.mark.parametrize("contract", [1, 2, 3])
def test_deterministic_paths_closed_form(contract):
    check_deterministic_paths(contract)


# ---- planted mistakes: every check above bites ----------------------------

@pytest.mark.parametrize("mutation,check", [
    (1, check_sobol_recurrence),                  # Gray code dropped
    (1, lambda: check_expected_output("small")),
    (2, check_ugaussian),                         # sign of dp^2 in the central case
    (2, lambda: check_expected_output("medium")),
    (3, lambda: check_bridge_covariance("large")),  # left weight on the right point
    (4, lambda: check_expected_output("medium")),   # correlation over the whole row
    (5, lambda: check_deterministic_paths(3)),      # payoff3 on date 365
])
def test_planted_mistake_is_caught(mutation, check):
    oracle.op_set_mutation(mutation)
    try:
        with pytest.raises(AssertionError):
            check()
    finally:
        oracle.op_set_mutation(0)
    check()  # and the unmutated oracle passes
