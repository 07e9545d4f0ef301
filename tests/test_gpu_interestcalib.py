"""InterestCalib on the GPU (libinterestcalib.so) against the oracle
(oracle/ic_oracle.c), through the C ABI.

* Pricing (interestcalib_eval): quotes, model prices and log-likelihoods of
  genomes on the datasets' swaptions and on synthetic ones (one-period,
  schedules above 64 periods, an empty schedule).  The quote is f32
  arithmetic on dates plus three expf and erf polynomials; the model price
  runs three root finds whose brackets and iteration counts depend on
  last-bit differences, so its tolerance is relative to the program's own
  f32 noise where a genome is ill-conditioned (t_fun's cancellation at
  small a or b).
* The calibration (interestcalib_trace), iteration by iteration: the initial
  population and every proposal are bit-exact against the oracle's
  propose step applied to the GPU's own state (f32 +-*/ min max only); the
  proposals' log-likelihoods match the oracle's within tolerance; every
  accept decision matches the oracle's on the GPU's log-likelihoods, except
  where exp(new - old) * fb lies within the expf difference of its Sobol
  draw; the winner and summary are main's reduction and evaluation of the
  GPU's final population.  The trajectory as a whole is not compared: one
  flipped decision changes every later iteration (the reference's own
  datasets hold one of several near-equal optima).
* main itself: the same bits as the traced run, the reference's datasets
  end near their expected log-likelihoods, and the reference's failures.
"""
import math

import numpy as np
import pytest

from conftest import load_ic_dataset
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ic():
    from locvolcalib import interestcalib as icl
    ctx = icl.InterestCalib(device=0)
    yield ctx
    ctx.free()


def data(name):
    return load_ic_dataset(name)


# genomes with a, b >= 0.1: no t_fun cancellation beyond the f32 noise
def well_conditioned(n, seed):
    rng = np.random.default_rng(seed)
    g = np.stack([rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n),
                  rng.uniform(-0.8, 0.8, n), rng.uniform(0.01, 0.2, n),
                  rng.uniform(0.005, 0.2, n)], axis=1)
    return g.astype(np.float32)


def oracle_eval(genomes, sw, hx, hw):
    q = np.empty((len(genomes), len(sw)), np.float32)
    p = np.empty_like(q)
    for i, g in enumerate(genomes):
        q[i], p[i] = oracle.ic_eval(g, sw, hx, hw)
    return q, p


def ulp_spread(genomes, sw, hx, hw):
    """The oracle's log-likelihoods of the genomes and how far they move
    when one gene moves by one ulp (the largest change over the ten such
    neighbours): the program's own f32 noise at that genome, which device
    and host expf/logf differences excite the same way."""
    g = np.ascontiguousarray(genomes, np.float32).reshape(-1, 5)
    base = oracle.ic_logliks(g, sw, hx, hw)
    spread = np.zeros(len(g), np.float64)
    for j in range(5):
        for d in (np.float32(np.inf), np.float32(-np.inf)):
            gp = g.copy()
            gp[:, j] = np.nextafter(gp[:, j], d)
            spread = np.maximum(spread, np.abs(oracle.ic_logliks(gp, sw, hx, hw).astype(np.float64)
                                               - base))
    return base, spread


def check_logliks(got, genomes, sw, hx, hw, what=""):
    """GPU log-likelihoods within 1e-4 relative plus 8x the one-ulp spread
    of the oracle's; and, over the well-conditioned genomes (one-ulp spread
    below 1e-6 relative) where there are several, the median within 1e-5
    relative."""
    want, spread = ulp_spread(genomes, sw, hx, hw)
    err = np.abs(np.asarray(got, np.float64) - want)
    tol = 1e-4 * np.abs(want) + 8.0 * spread
    worst = int(np.argmax(err - tol))
    assert np.all(err <= tol), (what, worst, float(got[worst]), float(want[worst]),
                                float(spread[worst]), np.asarray(genomes).reshape(-1, 5)[worst])
    wc = spread <= 1e-6 * np.abs(want)
    if wc.sum() >= 5:
        assert np.median(err[wc] / np.abs(want[wc])) <= 1e-5, what
    return float(np.max(err / np.maximum(1e-4 * np.abs(want) + spread, 1e-30)))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


@pytest.mark.parametrize("name", ["small", "medium"])
def test_eval_quotes_and_prices(ic, name):
    print(f"case: interestcalib_eval on {name}'s swaptions", flush=True)
    args, out6, summ = data(name)
    sw, hx, hw = args["swaptions"], args["herm_x"], args["herm_w"]
    g = well_conditioned(24, 1)
    q, p, ll = ic.eval_genomes(g, sw, hx, hw)
    oq, op = oracle_eval(g, sw, hx, hw)
    assert rel(q, oq).max() <= 1e-5  # P(d1) - P(-d1) cancels at small d1
    # the quote column of the dataset (it does not depend on the genome)
    assert rel(1e4 * q[0], summ[:, 1]).max() <= 1e-5
    r = rel(p, op)
    print(f"  price rel err max {r.max():.2e} median {np.median(r):.2e}", flush=True)
    assert np.median(r) <= 1e-5 and r.max() <= 2e-3
    check_logliks(ll, g, sw, hx, hw, name)


def test_eval_at_the_medium_winner(ic):
    print("case: interestcalib_eval, medium's expected winner", flush=True)
    args, out6, summ = data("medium")
    a, b, sigma, nu, rho = (float(v) for v in out6[:5])
    g = np.array([[a, b, rho, nu, sigma]], np.float32)
    q, p, ll = ic.eval_genomes(g, args["swaptions"], args["herm_x"], args["herm_w"])
    assert rel(1e4 * p[0], summ[:, 0]).max() <= 5e-4
    assert abs(float(ll[0]) - float(out6[5])) <= 2e-5 * abs(float(out6[5]))


def test_eval_synthetic_schedules(ic):
    """One-period swaptions (1 period a lane), 120 periods (2 a lane over 64
    lanes), 204 and 240 (4 a lane) and 360 (8 a lane, the long-schedule
    kernel), 64, a zero-period schedule (NaN, as the oracle), against the
    oracle."""
    print("case: interestcalib_eval, synthetic schedules", flush=True)
    args, _, _ = data("small")
    hx, hw = args["herm_x"], args["herm_w"]
    sw = np.array([[1, 12, 1, 0.3], [5, 12, 1, 0.3], [2, 1, 10, 0.2], [1, 1, 30, 0.2],
                   [3, 3, 16, 0.25], [2, 6, 32, 0.2], [1, 1, 17, 0.2], [1, 2, 40, 0.2],
                   [4, 12, 0, 0.3]], np.float32)
    g = well_conditioned(70, 2)  # more genomes than one wavefront holds at every width
    q, p, ll = ic.eval_genomes(g, sw, hx, hw)
    oq, op = oracle_eval(g, sw, hx, hw)
    assert np.array_equal(np.isnan(p), np.isnan(op)) and np.isnan(p[:, -1]).all()
    ok = ~np.isnan(op)
    assert rel(q[ok], oq[ok]).max() <= 1e-5
    r = rel(p[ok], op[ok])
    print(f"  price rel err max {r.max():.2e}", flush=True)
    assert r.max() <= 2e-3 and np.median(r) <= 1e-5


def test_eval_errors(ic):
    from locvolcalib.interestcalib import InterestCalibError
    args, _, _ = data("small")
    hx, hw = args["herm_x"], args["herm_w"]
    g = well_conditioned(2, 3)
    with pytest.raises(InterestCalibError, match="at most 512"):
        ic.eval_genomes(g, np.array([[1, 1, 43, 0.2]], np.float32), hx, hw)
    with pytest.raises(InterestCalibError, match="fields"):
        ic.eval_genomes(g, np.ones((3, 3), np.float32), hx, hw)
    with pytest.raises(InterestCalibError, match="weights"):
        ic.eval_genomes(g, args["swaptions"], hx, hw[:-1])


def trajectory_case(name, pop=None, conv=None, nswap=None):
    args, _, _ = data(name)
    sw = args["swaptions"] if nswap is None else args["swaptions"][:nswap]
    return (args["pop"] if pop is None else pop, args["mcmc_conv"] if conv is None else conv,
            sw, args["herm_x"], args["herm_w"], args["dir_vct"])


def check_trajectory(tr, pop, conv, sw, hx, hw, dv, ll_iters):
    """The per-iteration pins of the module docstring; ll_iters: the
    iterations whose proposals' log-likelihoods are recomputed by the
    oracle."""
    G, L, P, FB, NL = tr["genomes"], tr["logliks"], tr["props"], tr["fb"], tr["newl"]
    bmin = np.array([1e-3, 1e-3, np.float32(1e-3) - np.float32(1), 1e-3, 1e-3], np.float32)
    bmax = np.array([np.float32(1) - np.float32(1e-3)] * 3 + [0.2, 0.2], np.float32)
    init = np.array([[np.float32(oracle.ic_sobol(5 * i + 1 + j, dv)) * (bmax[j] - bmin[j]) + bmin[j]
                      for j in range(5)] for i in range(pop)], np.float32)
    assert np.array_equal(G[0], init)
    ratio = check_logliks(L[0], init, sw, hx, hw, "initial population")
    props = G[0].copy()
    offs = 5 * pop + 1
    flips = 0
    for it in range(conv):
        props, fb, acc = oracle.ic_propose(offs, G[it], props, dv)
        assert np.array_equal(P[it], props), it
        assert np.array_equal(FB[it], fb), it
        if it in ll_iters:
            ratio = max(ratio, check_logliks(NL[it], P[it], sw, hx, hw, f"iteration {it}"))
        g_or, l_or = oracle.ic_accept(acc, G[it], L[it], P[it], NL[it], FB[it], dv)
        differ = np.any(g_or != G[it + 1], axis=1) | (l_or != L[it + 1])
        for i in np.nonzero(differ)[0]:
            a = min(1.0, math.exp(float(NL[it][i]) - float(L[it][i])) * float(FB[it][i]))
            r = oracle.ic_sobol(acc + i, dv)
            assert abs(a - r) <= 1e-5 * max(r, 1e-30), (it, i, a, r)
            flips += 1
        offs = acc + pop
    w, wv = 0, np.float32(-np.inf)  # :752-756 as a left fold from (0, -inf)
    for i in range(pop):
        if wv < L[conv][i]:
            w, wv = i, L[conv][i]
    a, b, rho, nu, sigma = G[conv][w]
    assert np.array_equal(tr["out"], np.array([a, b, sigma, nu, rho, wv], np.float32))
    q, p = oracle.ic_eval(G[conv][w], sw, hx, hw)
    assert rel(tr["summary"][:, 1], np.float32(1e4) * q).max() <= 1e-5
    print(f"  log-likelihood error / (1e-4 rel + one-ulp spread): max {ratio:.2f}", flush=True)
    return flips


def test_trajectory_small_dataset(ic):
    print("case: interestcalib_trace, small (pop 256, 128 iterations, 196 swaptions)",
          flush=True)
    pop, conv, sw, hx, hw, dv = trajectory_case("small")
    tr = ic.trace(pop, conv, sw, hx, hw, dv)
    flips = check_trajectory(tr, pop, conv, sw, hx, hw, dv, ll_iters=set(range(0, conv, 16)))
    print(f"  decisions within the expf difference of their draw: {flips}", flush=True)


@pytest.mark.parametrize("name", ["medium", "large"])
def test_trajectory_full_datasets(ic, name):
    print(f"case: interestcalib_trace, {name} (pop 128, 424 iterations, 196 swaptions)",
          flush=True)
    pop, conv, sw, hx, hw, dv = trajectory_case(name)
    tr = ic.trace(pop, conv, sw, hx, hw, dv)
    flips = check_trajectory(tr, pop, conv, sw, hx, hw, dv, ll_iters=set(range(0, conv, 40)))
    print(f"  decisions within the expf difference of their draw: {flips}", flush=True)


@pytest.mark.parametrize("pop, conv", [(1, 3), (2, 12), (3, 9), (70, 5)])
def test_trajectory_small_populations(ic, pop, conv):
    print(f"case: interestcalib_trace pop {pop} conv {conv}", flush=True)
    _, _, sw, hx, hw, dv = trajectory_case("medium")
    sw = sw[::7]
    if pop == 1 and oracle.ic_schedule_has_de(pop, conv, dv):
        from locvolcalib.interestcalib import InterestCalibError
        with pytest.raises(InterestCalibError, match="genomes\\[-1\\]"):
            ic.trace(pop, conv, sw, hx, hw, dv)
        return
    tr = ic.trace(pop, conv, sw, hx, hw, dv)
    check_trajectory(tr, pop, conv, sw, hx, hw, dv, ll_iters=set(range(conv)))


def test_main_is_the_traced_run(ic):
    print("case: main vs interestcalib_trace, medium reduced", flush=True)
    pop, conv, sw, hx, hw, dv = trajectory_case("medium", conv=30)
    out, summ = ic.main(pop, conv, sw, hx, hw, dv)
    tr = ic.trace(pop, conv, sw, hx, hw, dv)
    assert np.array_equal(out, tr["out"]) and np.array_equal(summ, tr["summary"])
    out2, summ2 = ic.main(pop, conv, sw, hx, hw, dv)
    assert np.array_equal(out, out2) and np.array_equal(summ, summ2)


@pytest.mark.parametrize("name", ["small", "medium", "large"])
def test_main_on_the_datasets(ic, name):
    print(f"case: main on {name}", flush=True)
    args, out6, summ = data(name)
    out, s = ic.main(args["pop"], args["mcmc_conv"], args["swaptions"], args["herm_x"],
                     args["herm_w"], args["dir_vct"])
    print(f"  winner {out} (expected {out6})", flush=True)
    assert s.shape == (len(args["swaptions"]), 3)
    assert np.all(np.isfinite(out)) and np.all(np.isfinite(s))
    # the winner's summary is its evaluation, its log-likelihood the sum
    a, b, sigma, nu, rho = (float(v) for v in out[:5])
    g = np.array([a, b, rho, nu, sigma], np.float32)
    q, p = oracle.ic_eval(g, args["swaptions"], args["herm_x"], args["herm_w"])
    assert rel(s[:, 1], np.float32(1e4) * q).max() <= 1e-5
    check_logliks(out[5:6], g, args["swaptions"], args["herm_x"], args["herm_w"], name)
    # near the expected optimum's likelihood; large.out is a copy of medium's,
    # and large.in holds small's swaptions, so small's optimum is large's
    ref = load_ic_dataset("small" if name == "large" else name)[1][5]
    assert float(out[5]) >= float(ref) - 0.01 * abs(float(ref))


def test_main_errors(ic):
    from locvolcalib.interestcalib import InterestCalibError
    pop, conv, sw, hx, hw, dv = trajectory_case("medium", conv=40)
    with pytest.raises(InterestCalibError, match="empty"):
        ic.main(0, conv, sw, hx, hw, dv)
    with pytest.raises(InterestCalibError, match="genomes\\[-1\\]"):
        ic.main(1, 40, sw, hx, hw, dv)
    with pytest.raises(InterestCalibError, match="bits"):
        ic.main(pop, conv, sw, hx, hw, np.ones(32, np.int32))
    with pytest.raises(InterestCalibError, match="negative"):
        ic.main(pop, conv, np.array([[1, 6, -1, 0.2]], np.float32), hx, hw, dv)


def test_main_without_swaptions_is_bit_exact(ic):
    """With no swaption every log-likelihood is 0 (an empty sum, :641-645),
    so the accept step compares min(1, exp(0) * fb) with its draw: the
    whole trajectory is f32 +-*/ min max and Sobol numbers, and the GPU's
    winner must be the oracle's bit for bit."""
    print("case: main with no swaptions (pop 12, 40 iterations)", flush=True)
    _, _, _, hx, hw, dv = trajectory_case("medium")
    sw = np.zeros((0, 4), np.float32)
    out, summ = ic.main(12, 40, sw, hx, hw, dv)
    want, wsumm, _, _ = oracle.ic_main(12, 40, sw, hx, hw, dv)
    assert summ.shape == (0, 3) and wsumm.shape == (0, 3)
    assert np.array_equal(out, want), (out, want)


def test_main_without_hermite_points(ic):
    """No quadrature points: every model price is 0 (an empty sum, :594-620),
    so every genome has the same log-likelihood and the accept step again
    compares min(1, fb) with its draw; the winner's genome equals the
    oracle's bit for bit and its log-likelihood (a sum of Cauchy terms of
    the quotes) agrees within 1e-6 relative."""
    print("case: main with no Hermite points (pop 16, 30 iterations)", flush=True)
    _, _, sw, _, _, dv = trajectory_case("medium")
    sw = sw[::14]
    e = np.zeros(0, np.float32)
    out, summ = ic.main(16, 30, sw, e, e, dv)
    want, wsumm, _, _ = oracle.ic_main(16, 30, sw, e, e, dv)
    assert np.array_equal(out[:5], want[:5]), (out, want)
    assert abs(float(out[5]) - float(want[5])) <= 1e-6 * abs(float(want[5]))
    assert np.all(summ[:, 0] == 0.0) and np.all(wsumm[:, 0] == 0.0)


def test_trajectory_random_direction_vector(ic):
    """A random 30-bit direction vector (every Sobol number the XOR of a
    different subset) and random swaptions of 1..144 periods: every
    iteration pinned as on the datasets."""
    print("case: interestcalib_trace, random direction vector (pop 40, 30 iterations)",
          flush=True)
    rng = np.random.default_rng(11)
    dv = rng.integers(1, 1 << 30, 30).astype(np.int32)
    _, _, _, hx, hw, _ = trajectory_case("small")
    sw = np.stack([rng.integers(1, 11, 24), rng.choice([1, 3, 6, 12], 24),
                   rng.integers(1, 13, 24), rng.uniform(0.15, 0.6, 24)], axis=1).astype(np.float32)
    tr = ic.trace(40, 30, sw, hx, hw, dv)
    check_trajectory(tr, 40, 30, sw, hx, hw, dv, ll_iters=set(range(0, 30, 5)))
