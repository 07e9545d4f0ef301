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


@pytest.mark.parametrize("name", ["small", "medium"])
def test_payoffs_reproduce_expected_output(name):
    args, expected = load(name)
    pay = oracle.op_payoffs(*args)
    assert pay.shape == (expected.shape[0], args[1])
    np.testing.assert_array_equal(sequential_mean(pay, args[1]), expected)


def test_payoffs_large_expected_output():
    """large (contract 3, 367 dates): the sequential f32 sum lands within
    4e-6 of the expected output (which a different libm or reduction order
    produced)."""
    args, expected = load("large")
    got = sequential_mean(oracle.op_payoffs(*args), args[1])
    np.testing.assert_allclose(got, expected, rtol=1e-5)


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


def test_sobol_gray_code_recurrence():
    args, _ = load("large")
    dv = args[2]  # [1101][30]
    bits = dv.shape[1]
    x = np.zeros(dv.shape[0], np.int64)
    for n in range(1, 600):
        c = (n & -n).bit_length() - 1  # ctz(n): the Gray-code bit that flips
        x ^= dv[:, c]
        np.testing.assert_array_equal(oracle.op_sobol(n, dv),
                                      (x / 2.0 ** bits).astype(np.float32), err_msg=f"n={n}")


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


def test_ugaussian_is_the_normal_quantile():
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


@pytest.mark.parametrize("name", ["medium", "large"])
def test_bridge_covariance(name):
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


def test_contract_shape_checked():
    args, _ = load("medium")
    with pytest.raises(ValueError):
        oracle.op_main(1, 8, *args[2:])
