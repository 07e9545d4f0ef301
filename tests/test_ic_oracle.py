"""The InterestCalib oracle (oracle/ic_oracle.c) pinned against the
reference's own tests and datasets and against what the mathematics fixes:

* the calendar (lib/github.com/diku-dk/date/date.fut) against every vector
  of the library's own tests (date_tests.fut:16-74);
* uGaussian_P (InterestCalib.fut:56-93) against the normal CDF within the
  error bound of Abramowitz & Stegun 7.1.26, and the exp-factor form
  (:95-101) against exp(f) * P(x);
* sobolInd (:190-202) against the Gray-code recurrence
  x_n = x_(n-1) ^ v[ctz n], and the van der Corput property of the
  datasets' direction vector;
* rootFinding_Brent (:124-185) against scipy's brentq;
* the swaption quote (:509-525), which does not depend on the genome,
  against the quote column of small.out and medium.out for all 196
  swaptions (large.out is a byte copy of medium.out and belongs to medium's
  swaptions, not large's: shown below);
* the G2++ model price (:528-602) against medium.out's price column at
  medium's winning genome, and, for one-period swaptions, against the
  closed form of Brigo & Mercurio (zero-coupon-bond put in G2++,
  "Interest Rate Models", 2nd ed., section 4.2): at the money the payer
  swaption is P(0,T) (2 Phi(Sigma/2) - 1);
* the Cauchy log-likelihood (:606-625) as a density in the model price;
* the MCMC moves (:655-745) by their invariants, and main as the
  composition of the propose / evaluate / accept steps the GPU parity tests
  drive one at a time.

small.out's price column is no pin: its winning genome sits at b = 0.001,
where the f32 program is ill-conditioned (t_fun's bracket is b^2 tau^3 / 3
~ 3e-4 against rounding of its 2/b ~ 2000 terms), and a one-ulp change of b
moves a price by 40%; the test below shows it.  The whole calibration is
not pinned against the datasets' winners either: accept/reject compares
exp(new - old) with a Sobol number, so one ulp of a log-likelihood can flip
a move and the trajectory with it (the oracle's medium run ends in the
mirror optimum (a, sigma) <-> (b, nu) of the expected one).

Six planted mistakes in the oracle are each caught.
"""
import math

import numpy as np
import pytest

from conftest import load_ic_dataset
from oracle import oracle

# date_tests.fut:16-74, entry by entry: (inputs, expected)
DATE_TRIPLES_INV = [((1, 2, 3), (1, 2, 3))]
CHECK_DATE = [((1, 2, 3), True), ((1996, 2, 29), True), ((1997, 2, 29), False),
              ((2000, 2, 29), True), ((1900, 2, 29), False), ((2400, 2, 29), True),
              ((1600, 2, 29), True), ((2017, 1, 31), True), ((2017, 1, 0), False)]
ADD_DAYS = [((1, 2, 3, 1), (1, 2, 4)), ((1996, 2, 28, 1), (1996, 2, 29)),
            ((1996, 2, 28, 2), (1996, 3, 1)), ((1996, 12, 31, 1), (1997, 1, 1)),
            ((1997, 1, 1, -1), (1996, 12, 31))]
SUB_DAYS = [((1, 2, 3, 1), (1, 2, 2)), ((1996, 3, 1, 1), (1996, 2, 29)),
            ((1996, 12, 31, -1), (1997, 1, 1)), ((1997, 1, 1, 1), (1996, 12, 31))]
DIFF_DATES = [((1, 2, 3, 1, 2, 4), 0.002740), ((1, 2, 3, 2, 2, 3), 1.0)]
ADD_MONTHS = [((1, 2, 3, 1), (1, 3, 3)), ((1996, 12, 31, 1), (1997, 1, 31)),
              ((1996, 12, 31, 2), (1997, 2, 28))]
MINUTES_IN_DAY = 24 * 60


def check_dates():
    for (y, m, d), want in DATE_TRIPLES_INV:
        assert oracle.ic_triple_of_date(oracle.ic_date_of_triple(y, m, d)) == want
    for (y, m, d), want in CHECK_DATE:
        assert oracle.ic_check_date(y, m, d) == want, (y, m, d)
    for (y, m, d, n), want in ADD_DAYS:  # date.fut:123-124
        assert oracle.ic_triple_of_date(oracle.ic_date_of_triple(y, m, d) + n * MINUTES_IN_DAY) == want
    for (y, m, d, n), want in SUB_DAYS:
        assert oracle.ic_triple_of_date(oracle.ic_date_of_triple(y, m, d) - n * MINUTES_IN_DAY) == want
    for (y1, m1, d1, y2, m2, d2), want in DIFF_DATES:
        got = oracle.ic_diff_dates(oracle.ic_date_of_triple(y1, m1, d1),
                                   oracle.ic_date_of_triple(y2, m2, d2))
        assert abs(got - want) <= 5e-7, (got, want)  # the test prints 6 decimals
    for (y, m, d, n), want in ADD_MONTHS:
        assert oracle.ic_triple_of_date(oracle.ic_add_months(oracle.ic_date_of_triple(y, m, d), n)) == want


def test_date_library_vectors():
    check_dates()


def test_date_round_trip_and_month_ends():
    # every day of 1980..2230 (the range min_date..max_date spans) round-trips,
    # and add_months lands on min(day, end of month)
    t0 = oracle.ic_date_of_triple(1980, 1, 1)
    prev = None
    for k in range(0, 250 * 366, 7):
        t = t0 + k * MINUTES_IN_DAY
        y, m, d = oracle.ic_triple_of_date(t)
        assert oracle.ic_check_date(y, m, d)
        assert oracle.ic_date_of_triple(y, m, d) == t
        if prev is not None:
            assert (y, m, d) > prev
        prev = (y, m, d)
    for y in (1999, 2000, 2100, 2024):
        t = oracle.ic_date_of_triple(y, 1, 31)
        feb = oracle.ic_triple_of_date(oracle.ic_add_months(t, 1))
        leap = y % 4 == 0 and (y % 400 == 0 or y % 100 != 0)
        assert feb == (y, 2, 29 if leap else 28)


def phi(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def check_ugaussian():
    # A&S 7.1.26: |erf error| <= 1.5e-7, so P within 0.75e-7, plus f32 rounding
    xs = np.linspace(-8.0, 8.0, 3201, dtype=np.float32)
    err = max(abs(oracle.ic_ugaussian_P(float(x)) - phi(float(x))) for x in xs)
    assert err <= 2.5e-7, err


def test_ugaussian_P_within_the_published_error_bound():
    check_ugaussian()
    assert oracle.ic_ugaussian_P(0.0) == pytest.approx(0.5, abs=1e-7)


def test_ugaussian_P_with_exp_factor():
    for x in np.linspace(-6.0, 6.0, 241, dtype=np.float32):
        for f in (-3.0, -0.5, 0.0, 1.5):
            want = math.exp(f) * phi(float(x))
            got = oracle.ic_ugaussian_P_exp(float(x), f)
            assert abs(got - want) <= 3e-7 * math.exp(f) + 2e-6 * want, (x, f, got, want)


def gray_sobol(n_max, dv):
    """x_n by the Gray-code recurrence, scaled as sobolInd does in f32."""
    m = len(dv)
    norm = np.float32(1.0) / (np.float32(1 << m) + np.float32(1.0))
    out, r = [np.float32(0.0)], 0
    for n in range(1, n_max):
        c = (n & -n).bit_length() - 1
        r ^= int(dv[c])
        out.append(np.float32(np.float32(r) * norm))
    return out


def check_sobol(dv):
    want = gray_sobol(2048, dv)
    for n in range(2048):
        assert np.float32(oracle.ic_sobol(n, dv)) == want[n], n


def test_sobol_gray_code_and_van_der_corput():
    args, _, _ = load_ic_dataset("small")
    dv = args["dir_vct"]
    assert list(dv) == [1 << (29 - i) for i in range(30)]  # van der Corput in 30 bits
    check_sobol(dv)
    # each block of 2^9 consecutive indices covers every multiple of 2^-9 once
    pts = sorted(round(oracle.ic_sobol(n, dv) * (1 << 30)) >> 21 for n in range(512, 1024))
    assert pts == list(range(512))


def test_brent_against_brentq():
    from scipy.optimize import brentq
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(2, 9))
        sc = rng.uniform(0.05, 1.0, n).astype(np.float32)
        bb = rng.uniform(0.1, 3.0, n).astype(np.float32)
        f = lambda y: float(np.sum(sc.astype(np.float64) * np.exp(-bb.astype(np.float64) * y)) - 1.0)
        lo, hi = -5.0, 5.0
        if f(lo) * f(hi) >= 0:
            continue
        root = brentq(f, lo, hi, xtol=1e-12)
        got, err = oracle.ic_brent(sc, bb, lo, hi)
        assert abs(got - root) <= 2e-4, (got, root)
        assert abs(err) < 1e-2
    # no sign change: +inf from a non-negative lower bound, -inf otherwise
    sc = np.array([0.5, 0.5], np.float32)
    bb = np.array([1.0, 2.0], np.float32)
    assert oracle.ic_brent(sc, bb, 1.0, 2.0)[1] == math.inf
    assert oracle.ic_brent(sc, bb, -3.0, -1.0)[1] == -math.inf


def genome_of_out(out6):
    """main's first five outputs (a, b, sigma, nu, rho; :849) as a genome
    (a, b, rho, nu, sigma; :487)."""
    a, b, sigma, nu, rho = (float(v) for v in out6[:5])
    return np.array([a, b, rho, nu, sigma], np.float32)


def quotes_and_prices(name, genome=None):
    args, out6, summ = load_ic_dataset(name)
    g = genome_of_out(out6) if genome is None else genome
    q, p = oracle.ic_eval(g, args["swaptions"], args["herm_x"], args["herm_w"])
    return args, out6, summ, q.astype(np.float64), p.astype(np.float64)


def check_quotes():
    for name in ("small", "medium"):
        _, _, summ, q, _ = quotes_and_prices(name)
        rel = np.max(np.abs(1e4 * q - summ[:, 1]) / summ[:, 1])
        assert rel <= 1e-5, (name, rel)


def test_quotes_match_the_datasets():
    check_quotes()


def test_large_out_is_medium_out():
    _, out_m, summ_m = load_ic_dataset("medium")
    args_l, out_l, summ_l = load_ic_dataset("large")
    assert np.array_equal(out_m, out_l) and np.array_equal(summ_m, summ_l)
    # its quote column (independent of the genome) is medium's swaptions', not large's
    q, _ = oracle.ic_eval(genome_of_out(out_l), args_l["swaptions"], args_l["herm_x"],
                          args_l["herm_w"])
    assert np.max(np.abs(1e4 * q - summ_l[:, 1]) / summ_l[:, 1]) > 0.5


def check_medium_prices():
    args, out6, summ, q, p = quotes_and_prices("medium")
    rel = np.abs(1e4 * p - summ[:, 0]) / np.abs(summ[:, 0])
    assert rel.max() <= 5e-4, rel.max()
    err_pct = 100.0 * np.abs((p - q) / q)
    assert np.max(np.abs(err_pct - summ[:, 2])) <= 0.05
    ll = np.float32(0.0)
    for qq, pp in zip(q, p):
        ll = np.float32(ll + np.float32(oracle.ic_log_likelihood(qq, pp)))
    assert abs(float(ll) - float(out6[5])) <= 2e-5 * abs(float(out6[5])), (ll, out6[5])


def test_prices_at_the_medium_winner():
    check_medium_prices()


def test_small_winner_is_ill_conditioned():
    args, out6, summ, q, p = quotes_and_prices("small")
    g = genome_of_out(out6)
    assert g[1] == np.float32(0.001)  # b at its lower bound eps0 (:226)
    g2 = g.copy()
    g2[1] = np.nextafter(g2[1], np.float32(1.0))
    _, p2 = oracle.ic_eval(g2, args["swaptions"][3:4], args["herm_x"], args["herm_w"])
    assert abs(float(p2[0]) - p[3]) > 0.1 * abs(p[3])


# one-period swaptions (12 * term / freq = 1): [maturity, freq 12, term 1, vol]
CF_MATS = (1, 3, 5, 10)
CF_GRID = [(a, b, rho, nu, sigma) for a in (0.02, 0.1, 0.5) for b in (0.5, 0.96)
           for rho in (-0.9, -0.3, 0.4) for nu in (0.05, 0.15) for sigma in (0.01, 0.03)]


def zbp_swaption(genome, mat):
    """P(0,T) (2 Phi(Sigma/2) - 1) with Sigma^2 of the G2++ bond option,
    over the same dates as the reference (maturity = today + mat years, one
    12-month period), in f64."""
    a, b, rho, nu, sigma = (float(v) for v in genome)
    today = oracle.ic_date_of_triple(2012, 1, 1)
    tm = oracle.ic_add_months(today, int(np.float32(mat) * np.float32(12.0)))
    T = float(np.float32(oracle.ic_diff_dates(today, tm)))
    tau = float(np.float32(oracle.ic_diff_dates(tm, oracle.ic_add_months(tm, 12))))
    ea, eb = 1.0 - math.exp(-a * tau), 1.0 - math.exp(-b * tau)
    s2 = (sigma ** 2 / (2 * a ** 3) * ea ** 2 * (1 - math.exp(-2 * a * T)) +
          nu ** 2 / (2 * b ** 3) * eb ** 2 * (1 - math.exp(-2 * b * T)) +
          2 * rho * sigma * nu / (a * b * (a + b)) * ea * eb * (1 - math.exp(-(a + b) * T)))
    return math.exp(-0.03 * T) * (2.0 * phi(math.sqrt(s2) / 2.0) - 1.0)


def closed_form_errors():
    args, _, _ = load_ic_dataset("small")
    sw = np.array([[m, 12.0, 1.0, 0.3] for m in CF_MATS], np.float32)
    errs = []
    for g in CF_GRID:
        _, p = oracle.ic_eval(np.array(g, np.float32), sw, args["herm_x"], args["herm_w"])
        for m, pp in zip(CF_MATS, p):
            want = zbp_swaption(g, m)
            errs.append((g, m, abs(float(pp) - want) / want))
    return errs


def check_closed_form():
    errs = closed_form_errors()
    assert max(e for _, _, e in errs) <= 1.5e-2
    # where neither f32 cancellation (a tau small: t_fun's bracket is
    # a^2 tau^3 / 3 against rounding of 2/a) nor a dominant x factor (the
    # 11-point quadrature in x of a near-step integrand) limits the program
    good = [e for g, _, e in errs if g[0] >= 0.1 and g[4] == 0.01]
    assert max(good) <= 5e-4, max(good)


def test_one_period_swaption_closed_form():
    check_closed_form()


def test_log_likelihood_is_a_cauchy_density():
    from scipy.integrate import quad
    for q in (0.001, 0.01, 0.02):
        gamma = q / 50.0 * 5.0 + 0.01
        total, _ = quad(lambda p: math.exp(oracle.ic_log_likelihood(q, p)), q - 200 * gamma,
                        q + 200 * gamma, points=[q], limit=400)
        assert abs(total - 1.0) <= 4e-3  # the tails beyond 200 gamma hold 3.2e-3
        assert oracle.ic_log_likelihood(q, q) >= oracle.ic_log_likelihood(q, q + 1e-4)


# genomeBounds (:225-231) in f32 arithmetic
_E0, _ONE = np.float32(1e-3), np.float32(1.0)
BOUNDS_MIN = np.array([_E0, _E0, _E0 - _ONE, _E0, _E0], np.float32)
BOUNDS_MAX = np.array([_ONE - _E0, _ONE - _E0, _ONE - _E0, 0.2, 0.2], np.float32)


def move_type(sob_offs, dv):
    r = np.float32(oracle.ic_sobol(sob_offs, dv))
    return 1 if r <= np.float32(0.2) else 2 if r <= np.float32(0.5) else 3


def test_mcmc_moves_keep_their_invariants():
    args, _, _ = load_ic_dataset("small")
    dv = args["dir_vct"]
    pop = 12
    rng = np.random.default_rng(3)
    genomes = (BOUNDS_MIN + rng.uniform(0, 1, (pop, 5)).astype(np.float32) *
               (BOUNDS_MAX - BOUNDS_MIN)).astype(np.float32)
    props = genomes.copy()
    seen = set()
    offs = 5 * pop + 1
    for _ in range(40):
        t = move_type(offs, dv)
        new, fb, acc = oracle.ic_propose(offs, genomes, props, dv)
        assert np.all(new >= BOUNDS_MIN) and np.all(new <= BOUNDS_MAX)
        if t == 2:  # DIMS_ONE: one gene moves (amplitude 0 elsewhere, :290-293)
            moved = np.any(new != genomes, axis=0)
            assert moved.sum() <= 1
        if t == 3:
            assert np.all(fb == 1.0)
        assert acc == offs + 1 + {1: 5 * pop, 2: 5 * pop + 1, 3: 8 * pop}[t]
        seen.add(t)
        ll = rng.normal(0, 1, pop).astype(np.float32)
        # certain acceptance (exp(+100) * fb >= 1 > any Sobol number) and rejection
        g_up, l_up = oracle.ic_accept(acc, genomes, ll, new, ll + 100.0, fb, dv)
        assert np.array_equal(g_up, new) and np.array_equal(l_up, ll + 100.0)
        g_no, l_no = oracle.ic_accept(acc, genomes, ll, new, np.full(pop, -np.inf, np.float32), fb, dv)
        assert np.array_equal(g_no, genomes) and np.array_equal(l_no, ll)
        genomes, props = g_up, new
        offs = acc + pop
    assert seen == {1, 2, 3}


def small_case(npop=10, conv=6, nswap=12):
    args, _, _ = load_ic_dataset("medium")
    sw = args["swaptions"][::16][:nswap]
    return npop, conv, sw, args["herm_x"], args["herm_w"], args["dir_vct"]


def test_main_is_the_composition_of_its_steps():
    pop, conv, sw, hx, hw, dv = small_case()
    out6, summ, fg, fl = oracle.ic_main(pop, conv, sw, hx, hw, dv)
    genomes = np.array([[np.float32(oracle.ic_sobol(5 * i + 1 + j, dv)) * (BOUNDS_MAX[j] - BOUNDS_MIN[j])
                         + BOUNDS_MIN[j] for j in range(5)] for i in range(pop)], np.float32)
    ll = oracle.ic_logliks(genomes, sw, hx, hw)
    props = genomes.copy()
    offs = 5 * pop + 1
    for _ in range(conv):
        props, fb, acc = oracle.ic_propose(offs, genomes, props, dv)
        newl = oracle.ic_logliks(props, sw, hx, hw)
        genomes, ll = oracle.ic_accept(acc, genomes, ll, props, newl, fb, dv)
        offs = acc + pop
    assert np.array_equal(fg, genomes) and np.array_equal(fl, ll)
    w = int(np.argmax(ll))
    assert out6[5] == ll[w]
    assert np.array_equal(genome_of_out(out6), genomes[w])
    q, p = oracle.ic_eval(genomes[w], sw, hx, hw)
    assert np.array_equal(summ[:, 0], np.float32(1e4) * p)
    assert np.array_equal(summ[:, 1], np.float32(1e4) * q)


def test_main_fails_where_the_reference_fails():
    pop, conv, sw, hx, hw, dv = small_case()
    with pytest.raises(ValueError):
        oracle.ic_main(0, conv, sw, hx, hw, dv)
    # one genome and a DE move index genomes[-1]
    assert any(move_type(o, dv) == 3 for o in range(6, 6 + 40))
    with pytest.raises(ValueError):
        oracle.ic_main(1, 40, sw, hx, hw, dv)
    with pytest.raises(ValueError):
        oracle.ic_main(pop, conv, sw, hx, hw, np.ones(32, np.int32))


@pytest.mark.parametrize("mutation, check", [
    (1, "sobol"), (2, "ugaussian"), (3, "closed_form"), (4, "closed_form"),
    (5, "closed_form"), (6, "dates")])
def test_planted_mistakes_are_caught(mutation, check):
    args, _, _ = load_ic_dataset("small")
    fn = {"sobol": lambda: check_sobol(args["dir_vct"]), "ugaussian": check_ugaussian,
          "closed_form": check_closed_form, "dates": check_dates}[check]
    fn()  # passes without the mistake
    with oracle.ic_mutation(mutation):
        with pytest.raises(AssertionError):
            fn()
