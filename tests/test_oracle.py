"""Pins the CPU oracle (oracle/lvc_oracle_impl.h) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails it; the
planted-mistake tests at the end show that for five such mistakes.

Sources of truth:
  * the reference's own fixtures: finpar/LocVolCalib-data/{small,medium,
    large}.{in,out} and misc/ocean-sim/data/tridiag32-small.{in,out}
    (copied under tests/golden/);
  * closed forms and exact identities that LocVolCalib.fut fixes;
  * a library routine (numpy.linalg.solve) and dense matrices on tiny sizes.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, GRID32_EDGE, GRID32_NX, grid32_tolerance, load_dataset
from oracle import oracle
from locvolcalib import futhark_values as fv

# The fixtures print 6 significant digits (small/medium) or 6 decimals
# (large).  Measured disagreement of the f32 restatement with them: small
# 1.9e-6, medium 3.6e-6, large (strikes 0, 17, ..., 255) 1.3e-6 absolute:
# the files come from builds whose f32 rounding differs from this one.
F32_FIXTURE_ABS = 5e-6
# medium.out agrees with the FP64 restatement to its print precision (5e-8).
F64_MEDIUM_ABS = 1e-7

LARGE_SUBSET = np.arange(0, 256, 17)


def _main_subset(inp, idx, dtype, thomas=False):
    """`main` restricted to strikes 0.001*i for i in idx (options are
    independent, LocVolCalib.fut:248-251)."""
    outer, nx, ny, nt, s0, t, alpha, nu, beta = inp
    n = len(idx)
    real = np.dtype(dtype).type
    strikes = (real(0.001) * np.asarray(idx).astype(real)).astype(real)
    return oracle.batch([nx] * n, [ny] * n, [nt] * n, [s0] * n, strikes, [t] * n,
                        [alpha] * n, [nu] * n, [beta] * n, dtype=dtype,
                        thomas=thomas, nthreads=oracle._threads())


def check_fixture_small():
    inp, exp = load_dataset("small")
    got = oracle.main(*inp, dtype=np.float32, nthreads=oracle._threads())
    assert got.shape == exp.shape
    np.testing.assert_allclose(got, exp, rtol=0, atol=F32_FIXTURE_ABS)


def check_fixture_medium_f64():
    inp, exp = load_dataset("medium")
    got = oracle.main(*inp, dtype=np.float64, nthreads=oracle._threads())
    np.testing.assert_allclose(got, exp, rtol=0, atol=F64_MEDIUM_ABS)


# ---- end to end: the reference's fixtures ---------------------------------

def test_fixture_small_f32():
    check_fixture_small()


def test_fixture_medium_f32():
    inp, exp = load_dataset("medium")
    got = oracle.main(*inp, dtype=np.float32, nthreads=oracle._threads())
    np.testing.assert_allclose(got, exp, rtol=0, atol=F32_FIXTURE_ABS)


def test_fixture_medium_f64():
    check_fixture_medium_f64()


def test_fixture_large_f32_subset():
    inp, exp = load_dataset("large")
    got = _main_subset(inp, LARGE_SUBSET, np.float32)
    np.testing.assert_allclose(got, exp[LARGE_SUBSET], rtol=0, atol=F32_FIXTURE_ABS)


def test_fixture_large_f64_subset():
    # f64 vs f32 fixture: the fixture's own f32 drift (2e-6 at K=0, where the
    # exact price is s0) bounds the agreement.
    inp, exp = load_dataset("large")
    got = _main_subset(inp, LARGE_SUBSET, np.float64)
    np.testing.assert_allclose(got, exp[LARGE_SUBSET], rtol=0, atol=F32_FIXTURE_ABS)


def test_main_subset_equals_main():
    # options are independent: main(16, ...)[i] == value(strike_i)
    inp, _ = load_dataset("small")
    full = oracle.main(*inp, dtype=np.float64, nthreads=4)
    sub = _main_subset(inp, [3, 11], np.float64)
    assert np.array_equal(full[[3, 11]], sub)


# ---- closed form and invariants -------------------------------------------

@pytest.mark.parametrize("thomas", [False, True])
def test_closed_form_below_grid(thomas):
    """For K below the grid's first point the payoff is linear in x and
    constant in y; every operator row annihilates it, so price = s0 - K."""
    nx, ny, nt, s0, t, alpha, nu, beta = 32, 64, 40, 0.03, 5.0, 0.2, 0.6, 0.5
    xi, yi, x, y, tl = oracle.init_grid(s0, alpha, nu, t, nx, ny, nt)
    assert x[0] > 0.004
    for K in (0.0, 0.001, 0.002, 0.004):
        p = oracle.value(nx, ny, nt, s0, K, t, alpha, nu, beta, dtype=np.float64, thomas=thomas)
        assert abs(p - (s0 - K)) < 1e-12, (K, p)


def test_price_bounds_monotone_convex():
    inp, _ = load_dataset("small")
    outer, nx, ny, nt, s0, t, alpha, nu, beta = inp
    p = oracle.main(64, nx, ny, 64, s0, t, alpha, nu, beta, dtype=np.float64, nthreads=8)
    K = 0.001 * np.arange(64)
    assert np.all(p <= s0 + 1e-12)
    assert np.all(p >= np.maximum(s0 - K, 0) - 1e-12)
    assert np.all(np.diff(p) <= 1e-12)
    assert np.all(np.diff(p, 2) >= -1e-9)


def test_scan_form_matches_thomas():
    inp, _ = load_dataset("medium")
    outer, nx, ny, nt, s0, t, alpha, nu, beta = inp
    a = oracle.main(8, nx, ny, nt, s0, t, alpha, nu, beta, dtype=np.float64, nthreads=8)
    b = oracle.main(8, nx, ny, nt, s0, t, alpha, nu, beta, dtype=np.float64, thomas=True,
                    nthreads=8)
    np.testing.assert_allclose(a, b, rtol=1e-12)


# ---- initGrid (LocVolCalib.fut:12-23) -------------------------------------

def check_init_grid():
    for (nx, ny, nt, s0, t, alpha, nu) in [(256, 256, 64, 0.03, 5.0, 0.2, 0.6),
                                           (32, 64, 10, 0.045, 2.5, 0.13, 0.7)]:
        xi, yi, x, y, tl = oracle.init_grid(s0, alpha, nu, t, nx, ny, nt)
        dx = 20.0 * alpha * s0 * np.sqrt(t) / nx
        dy = 10.0 * nu * np.sqrt(t) / ny
        assert xi == int(s0 / dx) and yi == ny // 2
        assert x[xi] == s0                          # exact with contraction off
        assert y[yi] == np.log(alpha)
        np.testing.assert_allclose(np.diff(x), dx, rtol=1e-9)
        np.testing.assert_allclose(np.diff(y), dy, rtol=1e-9)
        assert tl[0] == 0.0 and tl[-1] == t
        np.testing.assert_allclose(tl, t * np.arange(nt) / (nt - 1), rtol=1e-15)


def test_init_grid():
    check_init_grid()


# ---- initOperator (LocVolCalib.fut:26-44) ---------------------------------

def check_init_operator():
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.uniform(0.5, 1.5, 12))       # a non-uniform grid
    D, DD = oracle.init_operator(x)
    inner = slice(1, len(x) - 1)

    def apply(Op, f):
        out = np.zeros(len(x))
        for i in range(1, len(x) - 1):
            out[i] = Op[i, 0] * f[i - 1] + Op[i, 1] * f[i] + Op[i, 2] * f[i + 1]
        return out[inner]

    one = np.ones_like(x)
    np.testing.assert_allclose(apply(D, one), 0, atol=1e-12)
    np.testing.assert_allclose(apply(D, x), 1, rtol=1e-12)
    np.testing.assert_allclose(apply(DD, one), 0, atol=1e-12)
    np.testing.assert_allclose(apply(DD, x), 0, atol=1e-12)
    np.testing.assert_allclose(apply(DD, x * x), 2, rtol=1e-12)
    # boundary rows are literals (:28-29, :40-41)
    du, dl = x[1] - x[0], x[-1] - x[-2]
    np.testing.assert_array_equal(D[0], [0, -1 / du, 1 / du])
    np.testing.assert_array_equal(D[-1], [-1 / dl, 1 / dl, 0])
    np.testing.assert_array_equal(DD[0], 0)
    np.testing.assert_array_equal(DD[-1], 0)


def test_init_operator():
    check_init_operator()


# ---- setPayoff / updateParams ----------------------------------------------

def test_set_payoff():
    x = np.linspace(-0.01, 0.05, 9)
    res = oracle.set_payoff(0.02, x, 3)
    assert res.shape == (3, 9)
    np.testing.assert_array_equal(res, np.tile(np.maximum(x - 0.02, 0), (3, 1)))


def check_update_params():
    x = np.linspace(0.01, 0.3, 7)
    y = np.linspace(-2.0, -1.0, 5)
    var = oracle.update_params(x, y, 1.3, 0.5, 0.6)
    closed = x[None, :] ** (2 * 0.5) * np.exp(2 * y[:, None] - 0.36 * 1.3)
    np.testing.assert_allclose(var, closed, rtol=1e-13)


def test_update_params():
    check_update_params()


# ---- tridagPar / tridagSeq --------------------------------------------------

# ---- initGrid's scalars in f32 (the reference's f32 entry) ------------------
# GRID32_EDGE, GRID32_NX and grid32_tolerance: see conftest.py.

def _f32(*v):
    return [float(np.float32(x)) for x in v]


def test_grid32_index_is_the_f32_index():
    """grid32 mode's myXindex equals the f32 restatement's on the edge
    parameters (where the FP64 index is one less) and on random ones, and
    x[myXindex] == s0 exactly (:21)."""
    rng = np.random.default_rng(11)
    params = [(0.03, t, a) for t, a in GRID32_EDGE]
    params += [tuple(rng.uniform(lo, hi) for lo, hi in ((0.02, 0.05), (1, 10), (0.1, 0.3)))
               for _ in range(40)]
    nedge_diff = 0
    for s0, t, a in params:
        s0, t, a, nu = _f32(s0, t, a, 0.6)
        for nx in GRID32_NX:
            xi32 = oracle.init_grid(np.float32(s0), np.float32(a), np.float32(nu), np.float32(t),
                                    nx, 16, 8, dtype=np.float32)[0]
            xi, _, x, _, _ = oracle.init_grid(s0, a, nu, t, nx, 16, 8, dtype=np.float64,
                                              grid32=True)
            xi64 = oracle.init_grid(s0, a, nu, t, nx, 16, 8, dtype=np.float64)[0]
            assert xi == xi32
            nedge_diff += xi64 != xi32
            if 0 <= xi < nx:
                assert x[xi] == s0
    assert nedge_diff >= len(GRID32_EDGE) * len(GRID32_NX)


def test_grid32_closed_form_at_edge():
    """At the edge parameters x[0] = 0 exactly (s0/dx integral in f32), so a
    strike K = 0 lies on the grid and the price is s0 - K = s0 in exact
    arithmetic (payoff linear, every operator row annihilates it)."""
    for t, a in GRID32_EDGE[:3]:
        p = _f32(0.03, t, a, 0.6, 0.5)
        got = oracle.main(2, 32, 32, 8, p[0], p[1], p[2], p[3], p[4], dtype=np.float64,
                          thomas=True, grid32=True)
        assert abs(got[0] - p[0]) < 1e-12


@pytest.mark.parametrize("nx", GRID32_NX)
def test_grid32_vs_f32_restatement(nx):
    """The FP64-grid form of the reference's f32 program agrees with the f32
    restatement at the f32 grid's own rounding error, on the edge parameters
    and random ones; the pure FP64 restatement (index off by one at the edge)
    does not."""
    rng = np.random.default_rng(nx)
    sets = [_f32(0.03, 4.0, 0.2, 0.6, 0.5)]
    for _ in range(6):
        sets.append(_f32(rng.uniform(0.02, 0.05), rng.uniform(1, 10), rng.uniform(0.1, 0.3),
                         rng.uniform(0.4, 0.8), rng.uniform(0.3, 0.7)))
    tol = grid32_tolerance(nx)
    nth = oracle._threads()
    for i, (s0, t, a, nu, b) in enumerate(sets):
        p32 = oracle.main(6, nx, 32, 12, *map(np.float32, (s0, t, a, nu, b)),
                          dtype=np.float32, nthreads=nth)
        pg = oracle.main(6, nx, 32, 12, s0, t, a, nu, b, dtype=np.float64, thomas=True,
                         nthreads=nth, grid32=True)
        assert np.max(np.abs(p32 - pg)) < tol
        if i == 0:
            p64 = oracle.main(6, nx, 32, 12, s0, t, a, nu, b, dtype=np.float64, thomas=True,
                              nthreads=nth)
            assert np.max(np.abs(p32 - p64)) > 1.5 * tol


def _dd_system(rng, n, dtype=np.float64):
    a = -rng.uniform(0.1, 1.0, n)
    c = -rng.uniform(0.1, 1.0, n)
    b = rng.uniform(0.1, 1.0, n) + np.abs(a) + np.abs(c)
    y = rng.uniform(-1, 1, n)
    return [v.astype(dtype) for v in (a, b, c, y)]


def _dense(a, b, c):
    n = len(b)
    A = np.diag(b.astype(np.float64))
    for i in range(n):
        if i > 0:
            A[i, i - 1] = a[i]
        if i < n - 1:
            A[i, i + 1] = c[i]
    return A


def check_tridag_dense():
    rng = np.random.default_rng(2)
    for n in range(1, 17):
        a, b, c, y = _dd_system(rng, n)
        ref = np.linalg.solve(_dense(a, b, c), y)
        np.testing.assert_allclose(oracle.tridag_par(a, b, c, y), ref, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(oracle.tridag_seq(a, b, c, y), ref, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(oracle.tridag_par4(a, b, c, y), ref, rtol=1e-12, atol=1e-14)


def test_tridag_dense_solve():
    check_tridag_dense()


def test_tridag_par_vs_seq_long():
    rng = np.random.default_rng(3)
    a, b, c, y = _dd_system(rng, 256)
    np.testing.assert_allclose(oracle.tridag_par(a, b, c, y), oracle.tridag_seq(a, b, c, y),
                               rtol=1e-12)
    np.testing.assert_allclose(oracle.tridag_par4(a, b, c, y), oracle.tridag_seq(a, b, c, y),
                               rtol=1e-12)


def test_tridag_ocean_sim_golden():
    """misc/ocean-sim/tridiag-test.fut:7-10: one golden pair shared by its
    four entries.  The f32 Thomas restatement reproduces every value
    bit-exactly; the scan form (a different association) agrees loosely, as
    these random systems are not diagonally dominant (outputs reach 1e8)."""
    a, b, c, y = fv.load(os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.in"))
    out = fv.load(os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.out"))[0]
    assert a.dtype == np.float32 and a.shape == (100, 115)
    seq = np.stack([oracle.tridag_seq(a[i], b[i], c[i], y[i]) for i in range(100)])
    assert np.array_equal(seq, out)
    for fn in (oracle.tridag_par, oracle.tridag_par4):
        par = np.stack([fn(a[i], b[i], c[i], y[i]) for i in range(100)])
        rel = np.abs(par.astype(np.float64) - out) / np.abs(out)
        assert np.median(rel) < 1e-5
        assert np.percentile(rel, 99) < 1e-2


# ---- explicitMethod (LocVolCalib.fut:126-153) ------------------------------

def test_explicit_dense():
    rng = np.random.default_rng(4)
    rows, m = 3, 7
    D = rng.uniform(-1, 1, (m, 3))
    DD = rng.uniform(-1, 1, (m, 3))
    mu = rng.uniform(-1, 1, (rows, m))
    var = rng.uniform(0, 1, (rows, m))
    res = rng.uniform(-1, 1, (rows, m))
    got = oracle.explicit_method(D, DD, mu, var, res)
    for r in range(rows):
        A = np.zeros((m, m))
        for j in range(m):
            coef = mu[r, j] * D[j] + 0.5 * var[r, j] * DD[j]
            for k in range(3):
                if 0 <= j + k - 1 < m:
                    A[j, j + k - 1] = coef[k]
        np.testing.assert_allclose(got[r], A @ res[r], rtol=1e-12, atol=1e-14)


# ---- planted mistakes: every check above bites ----------------------------

@pytest.mark.parametrize("mutation,check", [
    (1, check_init_operator),        # centre tap loses a term
    (1, check_fixture_small),
    (2, check_tridag_dense),         # wrong index in recurrence 1
    (3, check_fixture_medium_f64),   # dropped 0.5 on v in the y right-hand side
    (4, check_update_params),        # sign of the nu^2 t term
    (5, check_fixture_small),        # untransposed operand in explicitY
])
def test_planted_mistake_is_caught(mutation, check):
    oracle.set_mutation(mutation)
    try:
        with pytest.raises(AssertionError):
            check()
    finally:
        oracle.set_mutation(0)
    check()  # and the unmutated oracle passes
