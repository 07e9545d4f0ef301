"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's fixtures.

Tolerances (written here, measured against the FP64 oracle):
  * FP64 form: 1e-10 relative (the stated bar).  The GPU evaluates the
    operator in difference form (exact zero row sums), the oracle as
    written; the two differ by ~1e-12.
  * FP32 + one FP64-residual refinement per implicit solve: 1e-6 relative
    to the FP64 oracle (measured ~1e-7).
  * FP32 without refinement: 2e-5 relative to the FP64 oracle.
  * vs the reference's f32 fixtures: 5e-6 absolute (their own rounding).
"""
import os
import sys

import numpy as np
import pytest

from conftest import GRID32_EDGE, GRID32_NX, grid32_tolerance, load_dataset
from oracle import oracle
import locvolcalib as lvc

pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
F64_REL = 1e-10
F32R_REL = 1e-6
F32_REL = 2e-5
# plain FP32 on short runs (11 steps, dt ~0.36) of wide rows: measured 2.8e-5
# at numX = 512 (its error grows as the time grid coarsens)
F32_COARSE_REL = 1e-4
FIXTURE_ABS = 5e-6
CFG = dict(s0=0.03, t=5.0, alpha=0.2, nu=0.6, beta=0.5)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_strikes(nx, ny, nt, idx, dtype=np.float64, s0=0.03, t=5.0, alpha=0.2, nu=0.6,
                   beta=0.5, strike_dtype=None, grid32=False, thomas=True):
    """Oracle prices for strikes 0.001*i, i in idx, with the strike formed in
    `strike_dtype` (float32 for the reference's f32 entry, :247); `grid32`:
    initGrid's scalars in f32 (the form of the f32 entry)."""
    n = len(idx)
    sreal = np.dtype(strike_dtype or dtype).type
    K = (sreal(0.001) * np.asarray(idx).astype(sreal)).astype(np.float64)
    return oracle.batch([nx] * n, [ny] * n, [nt] * n, [s0] * n, K, [t] * n, [alpha] * n,
                        [nu] * n, [beta] * n, dtype=dtype, thomas=thomas, nthreads=NTH,
                        grid32=grid32)


def rel_err(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref) / np.abs(ref)))


# ---- FP64 form vs the FP64 oracle ---------------------------------------

def test_fp64_cfg1_all_options(gpu_ctx):
    say("case: FP64 main(16, 32, 256, 256)")
    got = gpu_ctx.main_f64(16, 32, 256, 256, *CFG.values())
    ref = oracle.main(16, 32, 256, 256, *CFG.values(), dtype=np.float64, thomas=True,
                      nthreads=NTH)
    assert rel_err(got, ref) < F64_REL
    ref_scan = oracle.main(16, 32, 256, 256, *CFG.values(), dtype=np.float64, nthreads=NTH)
    assert rel_err(got, ref_scan) < F64_REL


def test_fp64_cfg2_subset(gpu_ctx):
    say("case: FP64 main(128, 256, 256, 64)")
    got = gpu_ctx.main_f64(128, 256, 256, 64, *CFG.values())
    idx = np.array([0, 1, 17, 64, 100, 127])
    ref = oracle_strikes(256, 256, 64, idx)
    assert rel_err(got[idx], ref) < F64_REL
    # closed form at K = 0 (below the grid): s0 exactly in exact arithmetic
    assert abs(got[0] - 0.03) < 1e-13


@pytest.mark.parametrize("name", ["medium", "large"])
def test_fp64_vs_fixtures(gpu_ctx, name):
    say(f"case: FP64 fixture {name}")
    inp, exp = load_dataset(name)
    got = gpu_ctx.main_f64(*inp)
    atol = 1e-7 if name == "medium" else FIXTURE_ABS
    np.testing.assert_allclose(got, exp, rtol=0, atol=atol)


# ---- the reference's f32 entry ----------------------------------------------

@pytest.mark.parametrize("name", ["small", "medium", "large"])
@pytest.mark.parametrize("precision", [lvc.PRECISION_FP32, lvc.PRECISION_FP32_REFINE])
def test_f32_entry_vs_fixtures(name, precision):
    say(f"case: f32 entry precision={precision} fixture {name}")
    inp, exp = load_dataset(name)
    with lvc.LocVolCalib(precision=precision) as ctx:
        got = ctx.main(*inp)
    assert got.dtype == np.float32 and got.shape == exp.shape
    np.testing.assert_allclose(got, exp, rtol=0, atol=FIXTURE_ABS)


@pytest.mark.parametrize("cfg", [(128, 256, 256, 64), (256, 256, 256, 128)])
def test_fp32_refine_vs_fp64_oracle(cfg):
    outer, nx, ny, nt = cfg
    say(f"case: FP32+refine main{cfg}")
    with lvc.LocVolCalib(precision=lvc.PRECISION_FP32_REFINE) as ctx:
        got = ctx.main(outer, nx, ny, nt, *CFG.values())
    idx = np.array([0, 5, 60, outer // 2, outer - 1])
    # the f32 entry takes f32 parameters: the oracle gets the same values
    p = {k: float(np.float32(v)) for k, v in CFG.items()}
    ref = oracle_strikes(nx, ny, nt, idx, strike_dtype=np.float32, **p)
    err = rel_err(got[idx], ref)
    say(f"   max rel err {err:.3e}")
    assert err < F32R_REL


def test_fp32_plain_vs_fp64_oracle():
    say("case: FP32 main(128, 256, 256, 64)")
    with lvc.LocVolCalib(precision=lvc.PRECISION_FP32) as ctx:
        got = ctx.main(128, 256, 256, 64, *CFG.values())
    idx = np.array([0, 60, 127])
    p = {k: float(np.float32(v)) for k, v in CFG.items()}
    ref = oracle_strikes(256, 256, 64, idx, strike_dtype=np.float32, **p)
    assert rel_err(got[idx], ref) < F32_REL


# ---- the f32 entry's initGrid index (LocVolCalib.fut:15-20) ---------------
# futhark_entry_main takes f32 parameters and forms stdX, dx and
# myXindex = i32.f32(s0/dx) in f32, as the reference does; the grid arrays
# are FP64 from those scalars.  Checked here: the index the device's gather
# read (debugging report) against the f32 restatement's, bit for bit; the
# prices against the FP64 oracle of the same form (grid32) at the precision
# bars, and against the f32 restatement at its own grid-rounding error.

def _f32(*v):
    return [float(np.float32(x)) for x in v]


def _f32_entry_checks(ctx, outer, nx, ny, nt, p, bar):
    s0, t, a, nu, b = p
    got = ctx.main(outer, nx, ny, nt, s0, t, a, nu, b).astype(np.float64)
    xi32 = oracle.init_grid(*map(np.float32, (s0, a, nu, t)), nx, ny, nt,
                            dtype=np.float32)[0]
    assert ctx.price_index() == (ny // 2, xi32)
    idx = np.arange(outer)
    ref = oracle_strikes(nx, ny, nt, idx, s0=s0, t=t, alpha=a, nu=nu, beta=b,
                         strike_dtype=np.float32, grid32=True)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < bar, err
    ref32 = oracle.main(outer, nx, ny, nt, *map(np.float32, p), dtype=np.float32,
                        nthreads=NTH).astype(np.float64)
    assert np.max(np.abs(got - ref32)) < grid32_tolerance(nx)
    return err


@pytest.mark.parametrize("nx", GRID32_NX)
@pytest.mark.parametrize("precision", [lvc.PRECISION_FP32_REFINE, lvc.PRECISION_FP32])
def test_f32_entry_index_at_edge(nx, precision):
    """t = 4, alpha = 0.2: s0/dx = numX/8 is integral in f32 (the FP64 index
    would be numX/8 - 1, a ~2% price error)."""
    say(f"case: f32 entry numX={nx} t=4 alpha=0.2 precision={precision}")
    bar = F32R_REL if precision == lvc.PRECISION_FP32_REFINE else F32_COARSE_REL
    with lvc.LocVolCalib(precision=precision, debugging=True) as ctx:
        for t, a in GRID32_EDGE[:3]:
            _f32_entry_checks(ctx, 8, nx, 32, 12, _f32(0.03, t, a, 0.6, 0.5), bar)


def test_f32_entry_index_random():
    """64 seeded parameter sets across numX 32..512 (FP32 + refinement)."""
    say("case: f32 entry, 64 random parameter sets")
    rng = np.random.default_rng(2024)
    worst = 0.0
    with lvc.LocVolCalib(debugging=True) as ctx:
        for _ in range(64):
            nx = int(rng.choice(GRID32_NX))
            ny = int(rng.choice([32, 64]))
            nt = int(rng.integers(6, 20))
            p = _f32(rng.uniform(0.02, 0.05), rng.uniform(1, 10), rng.uniform(0.1, 0.3),
                     rng.uniform(0.4, 0.8), rng.uniform(0.3, 0.7))
            xi = oracle.init_grid(*map(np.float32, (p[0], p[2], p[3], p[1])), nx, ny, nt,
                                  dtype=np.float32)[0]
            if not 0 <= xi < nx:
                continue
            worst = max(worst, _f32_entry_checks(ctx, 4, nx, ny, nt, p, F32R_REL))
    say(f"   worst rel err vs the FP64 grid32 oracle {worst:.3e}")


def test_main_price_index_out_of_bounds(gpu_ctx):
    """20*alpha*sqrt(t) = 1: s0/dx = numX, so myXindex = numX lies outside the
    grid; the reference's bounds-checked gather (:236) fails the call, and so
    do both main entries; the batch entry gives NaN for that problem."""
    say("case: price index out of bounds")
    p = (0.03, 0.25, 0.1, 0.6, 0.5)
    with pytest.raises(lvc.LocVolCalibError, match="out of bounds"):
        gpu_ctx.main(4, 32, 32, 8, *p)
    with pytest.raises(lvc.LocVolCalibError, match="out of bounds"):
        gpu_ctx.main_f64(4, 32, 32, 8, *p)
    assert gpu_ctx.main(0, 32, 32, 8, *p).shape == (0,)   # no option, no index
    n = 3
    pr = dict(numX=np.full(n, 32, np.int32), numY=np.full(n, 32, np.int32),
              numT=np.full(n, 8, np.int32), s0=np.full(n, 0.03), strike=np.full(n, 0.01),
              t=np.array([0.25, 5.0, 0.25]), alpha=np.array([0.1, 0.2, 0.1]),
              nu=np.full(n, 0.6), beta=np.full(n, 0.5))
    got = gpu_ctx.main_batch(**pr, precision=lvc.PRECISION_FP64)
    ref = oracle.batch(pr["numX"], pr["numY"], pr["numT"], pr["s0"], pr["strike"], pr["t"],
                       pr["alpha"], pr["nu"], pr["beta"], dtype=np.float64, thomas=True,
                       allow_nan=True)
    assert np.isnan(got[0]) and np.isnan(got[2]) and np.isnan(ref[0])
    assert abs(got[1] - ref[1]) / abs(ref[1]) < F64_REL


# ---- size-independent properties at full size ----------------------------

def test_full_size_properties():
    say("case: FP32+refine main(256, 256, 256, 128) properties")
    with lvc.LocVolCalib() as ctx:
        p = ctx.main(256, 256, 256, 128, *CFG.values()).astype(np.float64)
    K = 0.001 * np.arange(256)
    s0 = float(np.float32(0.03))
    assert np.all(np.isfinite(p))
    assert abs(p[0] - s0) < 1e-8                       # closed form, K below the grid
    assert np.all(p <= s0 + 1e-8)
    assert np.all(p >= np.maximum(s0 - K, 0) - 1e-8)
    assert np.all(np.diff(p) <= 1e-9)                  # non-increasing in K
    assert np.mean(np.diff(p, 2) >= -1e-9) > 0.99      # convex in K


# ---- batched form (per-problem sizes and parameters) ----------------------

def make_problems(n, seed, sizes=(32, 64), nt_range=(8, 40)):
    rng = np.random.default_rng(seed)
    return dict(
        numX=rng.choice(sizes, n).astype(np.int32),
        numY=rng.choice(sizes, n).astype(np.int32),
        numT=rng.integers(nt_range[0], nt_range[1] + 1, n).astype(np.int32),
        s0=rng.uniform(0.02, 0.05, n), strike=rng.uniform(0.0, 0.1, n),
        t=rng.uniform(1, 10, n), alpha=rng.uniform(0.1, 0.3, n),
        nu=rng.uniform(0.4, 0.8, n), beta=rng.uniform(0.3, 0.7, n))


def oracle_batch(pr):
    return oracle.batch(pr["numX"], pr["numY"], pr["numT"], pr["s0"], pr["strike"], pr["t"],
                        pr["alpha"], pr["nu"], pr["beta"], dtype=np.float64, thomas=True,
                        nthreads=NTH)


def test_batch_fp64_vs_oracle(gpu_ctx):
    say("case: batch FP64, 48 uneven problems")
    pr = make_problems(48, seed=42)
    got = gpu_ctx.main_batch(**pr, precision=lvc.PRECISION_FP64)
    ref = oracle_batch(pr)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    assert rel_err(got[fin], ref[fin]) < F64_REL


def test_batch_fp32_refine_vs_oracle(gpu_ctx):
    say("case: batch FP32+refine, 48 uneven problems incl. 128/256 grids")
    pr = make_problems(48, seed=7, sizes=(32, 64, 128, 256), nt_range=(8, 24))
    got = gpu_ctx.main_batch(**pr, precision=lvc.PRECISION_FP32_REFINE)
    ref = oracle_batch(pr)
    # relative to the largest price, as some out-of-the-money prices are tiny
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    say(f"   max err / max price {err:.3e}")
    assert err < F32R_REL


def test_batch_ragged_sizes(gpu_ctx):
    """Sizes that are not multiples of the segment length (identity padding)."""
    say("case: batch FP64 ragged sizes")
    n = 6
    pr = dict(numX=np.array([3, 5, 37, 100, 129, 33], np.int32),
              numY=np.array([3, 70, 9, 31, 4, 257], np.int32),
              numT=np.array([2, 5, 9, 3, 6, 4], np.int32),
              s0=np.full(n, 0.03), strike=np.linspace(0, 0.02, n), t=np.full(n, 5.0),
              alpha=np.full(n, 0.2), nu=np.full(n, 0.6), beta=np.full(n, 0.5))
    got = gpu_ctx.main_batch(**pr, precision=lvc.PRECISION_FP64)
    ref = oracle_batch(pr)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    assert rel_err(got[fin], ref[fin]) < F64_REL


def test_batch_log_hazard_matches_oracle(gpu_ctx):
    """x_0 = s0 - trunc(s0/dx)*dx may be 0 or negative (log at :63); the GPU
    and the oracle must agree on NaN/finite and on the value."""
    say("case: batch FP64 log(x<=0) hazard")
    # alpha*sqrt(t) = 0.4 gives s0/dx = numX/8, an integer for these sizes
    n = 4
    pr = dict(numX=np.array([32, 64, 128, 40], np.int32),
              numY=np.full(n, 16, np.int32), numT=np.full(n, 6, np.int32),
              s0=np.array([0.03, 0.025, 0.04, 0.031]), strike=np.full(n, 0.01),
              t=np.full(n, 4.0), alpha=np.full(n, 0.2), nu=np.full(n, 0.6),
              beta=np.full(n, 0.5))
    got = gpu_ctx.main_batch(**pr, precision=lvc.PRECISION_FP64)
    ref = oracle_batch(pr)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref)
    if fin.any():
        assert rel_err(got[fin], ref[fin]) < F64_REL


# Ragged shapes for the y-sweep variants: ny not a multiple of the segment,
# a single segment per column (plain Thomas), nx below and above one 64-column
# group, the 3x3 minimum.
Y_SHAPES_SMALL = [(37, 200, 5), (64, 64, 4), (100, 33, 4), (129, 130, 4), (256, 256, 3),
                  (3, 3, 3), (70, 31, 5), (200, 97, 3)]
Y_SHAPES_TALL = [(70, 257, 3), (16, 500, 3), (130, 384, 3), (40, 512, 2)]


def _shapes_problems(shapes, seed):
    rng = np.random.default_rng(seed)
    n = len(shapes)
    return dict(numX=np.array([s[0] for s in shapes], np.int32),
                numY=np.array([s[1] for s in shapes], np.int32),
                numT=np.array([s[2] for s in shapes], np.int32),
                s0=rng.uniform(0.02, 0.05, n), strike=rng.uniform(0.0, 0.05, n),
                t=rng.uniform(1, 10, n), alpha=rng.uniform(0.1, 0.3, n),
                nu=rng.uniform(0.4, 0.8, n), beta=rng.uniform(0.3, 0.7, n))


@pytest.mark.parametrize("seg_y", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("shapes", ["small", "tall"])
@pytest.mark.parametrize("precision", [lvc.PRECISION_FP64, lvc.PRECISION_FP32_REFINE])
def test_y_segments_vs_oracle(seg_y, shapes, precision):
    """Every y-sweep variant (lane segments 4/8/16 with PCR, column segments
    32/64 with Thomas) against the FP64 oracle on ragged shapes."""
    say(f"case: y segment {seg_y}, {shapes} shapes, precision {precision}")
    pr = _shapes_problems(Y_SHAPES_SMALL if shapes == "small" else Y_SHAPES_TALL,
                          seed=seg_y)
    with lvc.LocVolCalib(tuning={"lvc.seg_y": seg_y, "lvc.seg_x": 8}) as ctx:
        got = ctx.main_batch(**pr, precision=precision)
    ref = oracle_batch(pr)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    g, r = got[fin], ref[fin]
    if precision == lvc.PRECISION_FP64:
        # relative per price; a zero price (deep out of the money) relative to
        # the largest price of the batch
        den = np.where(r != 0, np.abs(r), np.max(np.abs(r)))
        assert np.max(np.abs(g - r) / den) < F64_REL
    else:
        err = np.max(np.abs(g - r)) / np.max(np.abs(r))
        say(f"   max err / max price {err:.3e}")
        assert err < F32R_REL


def test_factor_sharing_bit_identical():
    """x tiles shared by problems that differ only in the strike (the
    reference's main, and runs of such problems in a batch) give the same bits
    as one problem per tile."""
    say("case: x-factor sharing (lvc.xshare 1 / 4 / 16), main and a grouped batch")
    rng = np.random.default_rng(5)
    n = 20
    base = dict(numX=np.full(n, 64, np.int32), numY=np.full(n, 96, np.int32),
                numT=np.full(n, 7, np.int32), s0=np.full(n, 0.03), t=np.full(n, 4.0),
                alpha=np.full(n, 0.2), nu=np.full(n, 0.6), beta=np.full(n, 0.5),
                strike=rng.uniform(0.0, 0.05, n))
    # break the run in places: other parameters, other sizes, other numT
    base["nu"][5] = 0.55
    base["numX"][11] = 32
    base["numT"][15:] = 5
    out = {}
    for xs in (1, 4, 16):
        with lvc.LocVolCalib(tuning={"lvc.xshare": xs}) as ctx:
            out[xs] = [ctx.main(24, 64, 64, 9, *CFG.values()),
                       ctx.main_f64(24, 64, 64, 9, *CFG.values()),
                       ctx.main_batch(**base, precision=lvc.PRECISION_FP32_REFINE),
                       ctx.main_batch(**base, precision=lvc.PRECISION_FP64)]
    for xs in (4, 16):
        for a, b in zip(out[1], out[xs]):
            assert np.array_equal(a, b)
    ref = oracle_batch(base)
    assert rel_err(out[4][3], ref) < F64_REL


def test_empty_and_errors(gpu_ctx):
    say("case: outer=0 and invalid sizes")
    assert gpu_ctx.main_f64(0, 32, 32, 8, *CFG.values()).shape == (0,)
    with pytest.raises(lvc.LocVolCalibError, match="at least 3"):
        gpu_ctx.main_f64(4, 2, 32, 8, *CFG.values())
    # the context stays usable after an error
    assert gpu_ctx.main_f64(2, 32, 32, 8, *CFG.values()).shape == (2,)


# ---- serving: a request's result is the same alone and in a mix -------------

def test_server_alone_vs_mixed():
    say("case: server alone vs mixed")
    pr = make_problems(24, seed=11, sizes=(32, 64, 128), nt_range=(8, 30))
    keys = ("numX", "numY", "numT", "s0", "strike", "t", "alpha", "nu", "beta")
    reqs = [tuple(int(pr[k][i]) if k.startswith("num") else float(pr[k][i]) for k in keys)
            for i in range(24)]
    with lvc.LocVolCalib() as ctx:
        alone = []
        srv = lvc.Server(ctx, max_slots=4, precision=lvc.PRECISION_FP32_REFINE)
        for r in reqs[:6]:
            rid = srv.submit(*r)
            srv.drain()
            alone.append(srv.poll(rid))
        srv.free()
        srv = lvc.Server(ctx, max_slots=8, precision=lvc.PRECISION_FP32_REFINE)
        ids = []
        for i, r in enumerate(reqs):        # staggered arrivals while others run
            ids.append(srv.submit(*r))
            if i % 3 == 2:
                srv.step()
        srv.drain()
        mixed = [srv.poll(rid) for rid in ids]
        st = srv.stats()
        srv.free()
    assert st["completed"] == 24 and st["resident"] == 0 and st["queued"] == 0
    assert mixed[:6] == alone                      # bit-identical
    ref = oracle_batch(pr)
    err = np.max(np.abs(np.array(mixed) - ref)) / np.max(np.abs(ref))
    assert err < F32R_REL


def test_server_memory_budget():
    """The budget bounds the server's device memory: slot tables first, the
    grid arena from the rest; a request waits until the arena has room."""
    say("case: server memory budget")
    with lvc.LocVolCalib() as ctx:
        srv = lvc.Server(ctx, max_slots=8)             # no budget: 8 grids of 512x512
        fixed = srv.device_bytes() - 8 * 2 * 512 * 512 * 4
        srv.free()
        assert fixed > 0
        one = 2 * 64 * 64 * 4                          # R and Y of a 64x64 request
        budget = fixed + 2 * one + 10
        srv = lvc.Server(ctx, max_slots=8, mem_budget=budget)
        try:
            assert srv.device_bytes() <= budget
            ids = [srv.submit(64, 64, 10, 0.03, 0.001 * i, 5.0, 0.2, 0.6, 0.5)
                   for i in range(5)]
            srv.step()
            st = srv.stats()
            assert st["resident"] == 2 and st["queued"] == 3
            assert st["resident_bytes"] == 2 * one
            srv.drain()
            assert all(srv.poll(i) is not None for i in ids)
            with pytest.raises(lvc.LocVolCalibError, match="budget"):
                srv.submit(256, 256, 10, 0.03, 0.0, 5.0, 0.2, 0.6, 0.5)
        finally:
            srv.free()
        with pytest.raises(lvc.LocVolCalibError, match="does not cover"):
            lvc.Server(ctx, max_slots=8, mem_budget=fixed // 2)


def test_profiling_report():
    say("case: profiling report")
    import json
    with lvc.LocVolCalib(profiling=True) as ctx:
        ctx.main(8, 64, 64, 5, *CFG.values())
        rep = json.loads(ctx.report())
    assert rep["lvc_xsweep"]["count"] == 4 and rep["lvc_ysweep"]["count"] == 4
    assert rep["lvc_xsweep"]["total_us"] > 0



def test_graph_replay_matches_direct():
    """The step sequence replayed from a HIP graph gives the same prices as
    direct launches (bit-identical), for a uniform and an uneven batch."""
    say("case: graph replay vs direct launches")
    pr = make_problems(16, seed=9, sizes=(32, 64, 128), nt_range=(4, 12))
    out = {}
    for graph in (0, 1):
        with lvc.LocVolCalib(tuning={"lvc.graph": graph}) as ctx:
            out[graph] = [ctx.main_batch(**pr, precision=lvc.PRECISION_FP32_REFINE),
                          ctx.main_f64(8, 64, 64, 9, *CFG.values())]
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
