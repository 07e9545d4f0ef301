"""GPU parity of the ocean-sim tridiagonal entries
(misc/ocean-sim/tridiag-test.fut:14-26) through the C ABI.

Bars: tridagNestedSeq(Const) (Thomas) bit-exact against the reference's
golden output tridiag32-small.out; tridagNested(Const) (the scan form)
bit-exact against the oracle's restatement of the same lines
(oracle.tridag_par4) and within the oracle's own loose agreement with the
golden output (these random systems are not diagonally dominant).  Larger
random inputs (the reference's `random input` shape class, the edges of the
register-resident kernel's chunks and of its 128-unknown limit, and systems
longer than the LDS-resident limit) bit-exact against the oracle."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from locvolcalib import futhark_values as fv
from oracle import oracle
import locvolcalib as lvc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    with lvc.LocVolCalib() as c:
        yield c


@pytest.fixture(scope="module")
def golden():
    a, b, c, y = fv.load(os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.in"))
    out = fv.load(os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.out"))[0]
    return (a, b, c, y), out


def oracle_rows(fn, a, b, c, y):
    return np.stack([fn(a[i], b[i], c[i], y[i]) for i in range(a.shape[0])]) \
        if a.shape[0] else np.zeros(a.shape, np.float32)


@pytest.mark.parametrize("const", [False, True])
def test_seq_bit_exact_vs_golden(ctx, golden, const):
    say(f"case: tridagNestedSeq{'Const' if const else ''} on tridiag32-small")
    (a, b, c, y), out = golden
    got = ctx.tridag_nested(a, b, c, y, seq=True, const=const)
    assert got.dtype == np.float32 and got.shape == out.shape
    assert np.array_equal(got, out)


@pytest.mark.parametrize("const", [False, True])
def test_par_bit_exact_vs_oracle(ctx, golden, const):
    say(f"case: tridagNested{'Const' if const else ''} on tridiag32-small")
    (a, b, c, y), out = golden
    got = ctx.tridag_nested(a, b, c, y, seq=False, const=const)
    ref = oracle_rows(oracle.tridag_par4, a, b, c, y)
    assert np.array_equal(got, ref)
    rel = np.abs(got.astype(np.float64) - out) / np.abs(out)
    assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-2


# m around the staged kernel's chunk edges (16 and 32 columns) and its
# 128-unknown limit, n around its 64-system workgroups
@pytest.mark.parametrize("n,m", [(3000, 115), (70, 300), (65, 1), (64, 2), (129, 256), (5, 257),
                                 (1, 16), (63, 17), (130, 31), (66, 33), (200, 64),
                                 (127, 100), (64, 127), (190, 128), (65, 129)])
def test_random_bit_exact_vs_oracle(ctx, n, m):
    say(f"case: random [{n}][{m}] both forms")
    rng = np.random.default_rng(n * 1000 + m)
    a, b, c, y = (rng.random((n, m), dtype=np.float32) for _ in range(4))
    for seq, fn in ((True, oracle.tridag_seq), (False, oracle.tridag_par4)):
        got = ctx.tridag_nested(a, b, c, y, seq=seq)
        ref = oracle_rows(fn, a, b, c, y)
        assert np.array_equal(got, ref, equal_nan=True)


def test_empty_and_errors(ctx, golden):
    say("case: n = 0, shape mismatch, Const with another m, m = 0")
    z = np.zeros((0, 115), np.float32)
    assert ctx.tridag_nested(z, z, z, z).shape == (0, 115)
    (a, b, c, y), _ = golden
    with pytest.raises(lvc.LocVolCalibError, match="shape"):
        ctx.tridag_nested(a, b, c, y[:50])
    s = np.ones((4, 100), np.float32)
    with pytest.raises(lvc.LocVolCalibError, match="INNER_DIM"):
        ctx.tridag_nested(s, s, s, s, seq=True, const=True)
    e = np.zeros((3, 0), np.float32)
    with pytest.raises(lvc.LocVolCalibError, match="at least 1"):
        ctx.tridag_nested(e, e, e, e)
    assert ctx.tridag_nested(a, b, c, y, seq=True).shape == (100, 115)  # context still usable
