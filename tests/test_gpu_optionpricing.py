"""OptionPricing on the GPU (liboptionpricing.so through its C ABI) against
the oracle's f64-accumulated mean of the reference's f32 payoffs, on the
reference's three datasets and on cases cut from them.

Tolerance: the kernels compute each path's payoff in f32 in the
reference's order of operations; expf, logf and sqrtf on the device may
differ from the host's libm by an ulp, which moves a payoff by a few ulps
(and, at the rare path that sits on a payoff barrier, flips it), so the
means agree to rtol 2e-5 rather than bit for bit.  Against the expected
outputs of the datasets, which carry the f32 rounding of a sequential sum
over the paths, the tolerance is rtol 2e-3.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle
from locvolcalib import futhark_values as fv
from locvolcalib import optionpricing as op

pytestmark = pytest.mark.gpu

DATA = os.path.join(GOLDEN, "OptionPricing-data")
RTOL = 2e-5


def load(name):
    args = op.main_arguments(fv.load(os.path.join(DATA, f"{name}.in")))
    expected = np.asarray(fv.load(os.path.join(DATA, f"{name}.out"))[0], np.float32)
    return args, expected


@pytest.fixture(scope="module")
def ctx():
    c = op.OptionPricing(device=0)
    yield c
    c.free()


@pytest.fixture(scope="module")
def datasets():
    return {n: load(n) for n in ("small", "medium", "large")}


@pytest.mark.parametrize("name", ["small", "medium", "large"])
def test_dataset(ctx, datasets, name):
    args, expected = datasets[name]
    print(f"OptionPricing {name}: contract {args[0]}, {args[1]} paths", flush=True)
    got = ctx.main(*args)
    assert got.dtype == np.float32 and got.shape == expected.shape
    np.testing.assert_allclose(got, oracle.op_main(*args), rtol=RTOL)
    np.testing.assert_allclose(got, expected, rtol=2e-3)


@pytest.mark.parametrize("name", ["small", "medium", "large"])
@pytest.mark.parametrize("paths", [1, 2, 63, 64, 65, 255, 256, 257, 1000, 4099])
def test_path_counts(ctx, datasets, name, paths):
    """Paths 1 .. n, n around the wave (64) and workgroup (256) blocks the
    kernel splits the Sobol indices into."""
    args, _ = datasets[name]
    got = ctx.main(args[0], paths, *args[2:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], paths, *args[2:]), rtol=RTOL,
                               atol=1e-4)


def test_grid_stride(ctx, datasets):
    """More blocks of 256 paths than the grid has workgroups (2048): the
    grid-stride loop, on contract 2 (register-resident path)."""
    args, _ = datasets["medium"]
    n = 2048 * 256 * 2 + 77
    got = ctx.main(args[0], n, *args[2:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], n, *args[2:]), rtol=RTOL)


def _models(args, m, seed):
    """m market models from the dataset's one: vols, drifts, starts and
    discounts perturbed per model."""
    rng = np.random.default_rng(seed)
    c, n, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd = args
    scale = lambda a, s: (a * (1 + s * rng.standard_normal((m,) + a.shape[1:]))).astype(np.float32)
    return (c, n, dv, np.repeat(cs, m, 0), scale(np.repeat(vols, m, 0), 0.1),
            scale(np.repeat(drifts, m, 0), 0.1), scale(np.repeat(sts, m, 0), 0.02),
            np.repeat(detv, m, 0), scale(np.repeat(disc, m, 0), 0.01), bbi, bbd)


@pytest.mark.parametrize("name,m", [("small", 8), ("medium", 3), ("large", 2), ("small", 17),
                                    ("medium", 11), ("large", 9)])
def test_several_models(ctx, datasets, name, m):
    args = _models(datasets[name][0], m, seed=m)
    n = 20000
    got = ctx.main(args[0], n, *args[2:])
    assert got.shape == (m,)
    np.testing.assert_allclose(got, oracle.op_main(args[0], n, *args[2:]), rtol=RTOL)


@pytest.mark.parametrize("bits", [1, 5, 6, 7, 16])
def test_fewer_direction_bits(ctx, datasets, bits):
    """num_bits below and around the 6 low Gray-code bits a lane owns."""
    args, _ = datasets["medium"]
    dv = np.ascontiguousarray(args[2][:, :bits])
    n = 3000
    got = ctx.main(args[0], n, dv, *args[3:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], n, dv, *args[3:]), rtol=RTOL)


@pytest.mark.parametrize("bits", [31, 32])
def test_full_direction_bits(ctx, datasets, bits):
    """31 and 32 direction bits (the most testBit can address in an i32,
    :17-18; Gray-code bit 31 of a non-negative index is never set)."""
    args, _ = datasets["medium"]
    rng = np.random.default_rng(bits)
    dv = rng.integers(-(1 << 31), 1 << 31, size=(15, bits), dtype=np.int64).astype(np.int32)
    dv[:, 30] = 1 << 30
    n = 5000
    got = ctx.main(args[0], n, dv, *args[3:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], n, dv, *args[3:]), rtol=RTOL)


@pytest.mark.parametrize("contract", [1, 2, 3])
def test_deterministic_paths_closed_form(ctx, contract):
    """Zero volatilities: every path's prices follow the drifts, and the
    payoff has a closed form (see test_op_oracle.deterministic_case)."""
    from test_op_oracle import deterministic_case
    args, want = deterministic_case(contract)
    got = ctx.main(*args)
    np.testing.assert_allclose(got, [want], rtol=1e-4)
    np.testing.assert_allclose(got, oracle.op_main(*args), rtol=RTOL)


def _random_bridge(D, rng, permutation=True, scale=0.1, late_reads=False):
    """bb_inds / bb_data for a bridge in a random order: step i writes date
    bi[i] from an earlier-written right neighbour and, or not (li = 0, j =
    -1), an earlier-written left one.  Without `permutation` some dates are
    written twice and others never (they stay 0, :153).  With `late_reads`
    every third step's right neighbour is a date written at a later step,
    which the reference reads as the initial 0."""
    bi = rng.permutation(D) + 1 if permutation else rng.integers(1, D + 1, D)
    li = np.zeros(D, np.int64)
    ri = np.zeros(D, np.int64)
    for i in range(1, D):
        ri[i] = bi[rng.integers(0, i)]
        li[i] = 0 if rng.random() < 0.3 else bi[rng.integers(0, i)]
        if late_reads and i % 3 == 1 and i + 1 < D:
            ri[i] = bi[rng.integers(i + 1, D)]
    ri[0] = li[0] = 1
    inds = np.stack([bi, li, ri]).astype(np.int32)
    data = (rng.random((3, D)) * np.array([[scale], [0.5], [0.5]])).astype(np.float32)
    return inds, data


@pytest.mark.parametrize("name", ["medium", "large"])
@pytest.mark.parametrize("permutation,late_reads", [(True, False), (False, False), (True, True)])
def test_random_bridge_tables(ctx, datasets, name, permutation, late_reads):
    """The generic bridge indexing (:153-177) beyond the datasets' tables:
    a valid bridge (the kernel skips the rows' zero fill), dates written
    twice or never, and reads of dates written later (both keep it)."""
    args, _ = datasets[name]
    rng = np.random.default_rng(13 if late_reads else 11 if permutation else 12)
    D = args[9].shape[1]
    # large: wide enough paths that the barriers of payoff3 are reached
    bbi, bbd = _random_bridge(D, rng, permutation, scale=0.1 if name == "medium" else 4.0,
                              late_reads=late_reads)
    a = args[:9] + (bbi, bbd)
    n = 3000
    got = ctx.main(a[0], n, *a[2:])
    want = oracle.op_main(a[0], n, *a[2:])
    assert np.all(np.isfinite(want))
    np.testing.assert_allclose(got, want, rtol=RTOL)


@pytest.mark.parametrize("late_reads", [False, True])
def test_bridge_tables_two_paths_per_lane(ctx, datasets, late_reads):
    """Past 2048 x 256 paths the path kernel's lanes run a second path, whose
    bridge rows start with the first path's values: a valid bridge (the
    rows' zero fill skipped) must write every date before reading it, and a
    table that reads dates written later must keep the fill (those reads
    see 0, :153)."""
    args, _ = datasets["large"]
    rng = np.random.default_rng(21 if late_reads else 22)
    D = args[9].shape[1]
    bbi, bbd = _random_bridge(D, rng, True, scale=4.0, late_reads=late_reads)
    a = args[:9] + (bbi, bbd)
    n = 2048 * 256 + 5000
    got = ctx.main(a[0], n, *a[2:])
    want = oracle.op_main(a[0], n, *a[2:])
    assert np.all(np.isfinite(want))
    np.testing.assert_allclose(got, want, rtol=RTOL)


def test_profiling_report(datasets):
    """With profiling on, futhark_context_report counts the entry's kernel
    launches and their time (JSON, as the generated API's report)."""
    import json
    args, _ = datasets["medium"]
    with op.OptionPricing(device=0, profiling=True) as c:
        c.main(args[0], 5000, *args[2:])
        c.main(args[0], 5000, *args[2:])
        rep = json.loads(c.report())
    assert rep["op_paths"]["count"] == 2 and rep["op_paths"]["total_us"] > 0


def test_repeated_calls_identical(ctx, datasets):
    """The sum order is fixed by the path count: a result does not change
    from call to call."""
    args, _ = datasets["large"]
    a = ctx.main(args[0], 50000, *args[2:])
    b = ctx.main(args[0], 50000, *args[2:])
    np.testing.assert_array_equal(a, b)


def test_no_paths_is_nan(ctx, datasets):
    """num_mc_it = 0: the empty sum divided by 0 (:322)."""
    args, _ = datasets["medium"]
    assert np.all(np.isnan(ctx.main(args[0], 0, *args[2:])))


def test_other_contract_is_zero(ctx, datasets):
    """A contract number without a payoff prices at 0 (:294), whatever the
    shapes."""
    args, _ = datasets["medium"]
    np.testing.assert_array_equal(ctx.main(7, 100, *args[2:]), np.zeros(1, np.float32))


def test_no_models(ctx, datasets):
    args, _ = datasets["medium"]
    c, n, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd = args
    got = ctx.main(c, 100, dv, cs[:0], vols[:0], drifts[:0], sts[:0], detv[:0], disc[:0],
                   bbi, bbd)
    assert got.shape == (0,)


@pytest.mark.parametrize("case,fragment", [
    ("contract", "contract 1 takes"),
    ("negative", "negative"),
    ("bits", "at most 32"),
    ("discounts", "discounts"),
    ("shape", "md_vols"),
])
def test_errors(ctx, datasets, case, fragment):
    args, _ = datasets["medium"]
    c, n, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd = args
    if case == "contract":
        c = 1
    elif case == "negative":
        n = -1
    elif case == "bits":
        dv = np.zeros((15, 33), np.int32)
    elif case == "discounts":
        disc = disc[:, :4]
    elif case == "shape":
        vols = vols[:, :4]
    with pytest.raises(op.OptionPricingError, match=fragment):
        ctx.main(c, n, dv, cs, vols, drifts, sts, detv, disc, bbi, bbd)
    # the context stays usable after an error
    got = ctx.main(args[0], 64, *args[2:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], 64, *args[2:]), rtol=RTOL)


def test_both_libraries_in_one_process(ctx, datasets, gpu_ctx):
    """liblocvolcalib.so and liboptionpricing.so export the same futhark_*
    names; each context keeps working beside the other."""
    args, _ = datasets["small"]
    got = ctx.main(args[0], 1000, *args[2:])
    np.testing.assert_allclose(got, oracle.op_main(args[0], 1000, *args[2:]), rtol=RTOL)
    assert np.isfinite(gpu_ctx.main(2, 32, 32, 8, 0.03, 5.0, 0.2, 0.6, 0.5)).all()
