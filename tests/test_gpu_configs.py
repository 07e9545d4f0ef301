"""GPU parity at the benchmark configurations' own sizes and across the
reference's input domain, through the C ABI, against the FP64 oracle.

  * cfg3: the batched form's exact workload -- 512 uneven problems (numX,
    numY from {32, 64, 128, 256}, numT uniform in [32, 256], parameters drawn
    per problem, seed 42; the bench's generator);
  * cfg5: the served form on a recorded Poisson arrival sequence of 220
    requests (grids 32..256, numT 32..256), each result against the oracle
    and against the same request served alone (bit for bit);
  * numX = 512 (the largest size of LocVolCalib-data/large.in:2, "2^x, x in
    {5..9}") and 513..1024 (the library's limit), every x segment length
    (4, 8, 16) in all three precisions.

Bars: FP64 1e-10 relative per price (a zero price relative to the batch's
largest); FP32 + refinement 1e-6 of the batch's largest price (deep
out-of-the-money prices are tiny); plain FP32 1e-4 of the largest price: on
the coarse time grids used here (2 to 7 steps, dt up to 5) its error reaches
2.3e-5, against ~5e-6 at 63 steps.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle
import locvolcalib as lvc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

NTH = min(16, os.cpu_count() or 1)
F64_REL = 1e-10
BARS = {lvc.PRECISION_FP64: None, lvc.PRECISION_FP32_REFINE: 1e-6, lvc.PRECISION_FP32: 1e-4}
KEYS = ("numX", "numY", "numT", "s0", "strike", "t", "alpha", "nu", "beta")


def say(msg):
    print(msg, flush=True)


def uneven(n, seed, sizes=(32, 64, 128, 256), nt=(32, 256)):
    # the bench's generator for the batched and served workloads
    import bench
    return bench.uneven_problems(n, seed, sizes=sizes, nt=nt)


def oracle_batch(pr):
    return oracle.batch(*[pr[k] for k in KEYS], dtype=np.float64, thomas=True, nthreads=NTH,
                        allow_nan=True)


def check(got, ref, precision):
    """Max error against the bar of `precision`; returns (abs, rel-to-max)."""
    got = np.asarray(got, np.float64)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    g, r = got[fin], ref[fin]
    top = np.max(np.abs(r))
    err_abs = float(np.max(np.abs(g - r)))
    if precision == lvc.PRECISION_FP64:
        den = np.where(r != 0, np.abs(r), top)
        assert np.max(np.abs(g - r) / den) < F64_REL
    else:
        assert err_abs / top < BARS[precision], err_abs / top
    return err_abs, err_abs / top


# ---- cfg3: 512 uneven problems, seed 42 -------------------------------------

@pytest.fixture(scope="module")
def cfg3():
    pr = uneven(512, 42)
    say("oracle: cfg3 512 problems (FP64, Thomas)")
    return pr, oracle_batch(pr)


@pytest.mark.parametrize("precision", [lvc.PRECISION_FP64, lvc.PRECISION_FP32_REFINE])
def test_cfg3_batch_vs_oracle(cfg3, precision):
    pr, ref = cfg3
    say(f"case: cfg3 batch, 512 problems, precision {precision}")
    with lvc.LocVolCalib() as ctx:
        got = ctx.main_batch(**pr, precision=precision)
    ea, er = check(got, ref, precision)
    say(f"   max abs err {ea:.3e}, / max price {er:.3e}")


# ---- cfg5: a recorded Poisson arrival sequence through the server -----------

def _requests(pr, i):
    return tuple(int(pr[k][i]) if k.startswith("num") else float(pr[k][i]) for k in KEYS)


def test_cfg5_served_replay_vs_oracle_and_alone():
    """220 requests arriving as a Poisson stream (0.6 per server tick, seeded),
    admitted while others run, against the oracle; each one also served alone,
    bit for bit."""
    n = 220
    rng = np.random.default_rng(7)
    ticks = np.floor(np.cumsum(rng.exponential(1 / 0.6, n))).astype(int)
    pr = uneven(n, 8)
    reqs = [_requests(pr, i) for i in range(n)]
    say("oracle: cfg5 220 requests")
    ref = oracle_batch(pr)
    say("case: cfg5 served replay (FP32 + refinement)")
    with lvc.LocVolCalib() as ctx:
        srv = lvc.Server(ctx, max_slots=48, precision=lvc.PRECISION_FP32_REFINE)
        ids, tick, i = [], 0, 0
        max_resident = 0
        while i < n or srv.stats()["resident"] or srv.stats()["queued"]:
            while i < n and ticks[i] <= tick:
                ids.append(srv.submit(*reqs[i]))
                i += 1
            srv.step()
            max_resident = max(max_resident, srv.stats()["resident"])
            tick += 1
        mixed = np.array([srv.poll(r) for r in ids])
        srv.free()
        say(f"   {tick} ticks, up to {max_resident} requests resident")
        assert max_resident > 8            # the requests did overlap on the device
        ea, er = check(mixed, ref, lvc.PRECISION_FP32_REFINE)
        say(f"   max abs err {ea:.3e}, / max price {er:.3e}")
        say("case: cfg5 each request alone")
        srv = lvc.Server(ctx, max_slots=1, precision=lvc.PRECISION_FP32_REFINE)
        alone = []
        for r in reqs:
            rid = srv.submit(*r)
            srv.drain()
            alone.append(srv.poll(rid))
        srv.free()
    assert np.array_equal(np.array(alone), mixed)


def test_served_under_a_small_budget():
    """64 slots but an 8 MiB budget (64 grids of 512x512 would take 128 MiB):
    the server holds no more device memory than the budget, admits requests as
    the arena has room, and every result equals the oracle's and the same
    request's served alone (bit for bit)."""
    n, budget = 96, 8 << 20
    pr = uneven(n, 21, nt=(8, 24))
    reqs = [_requests(pr, i) for i in range(n)]
    say("oracle: 96 requests")
    ref = oracle_batch(pr)
    say("case: 96 requests, 64 slots, 8 MiB budget")
    with lvc.LocVolCalib() as ctx:
        srv = lvc.Server(ctx, max_slots=64, mem_budget=budget)
        dev = srv.device_bytes()
        assert dev <= budget
        ids = [srv.submit(*r) for r in reqs]
        peak_res, peak_bytes = 0, 0
        while srv.stats()["resident"] or srv.stats()["queued"]:
            srv.step()
            st = srv.stats()
            peak_res = max(peak_res, st["resident"])
            peak_bytes = max(peak_bytes, st["resident_bytes"])
        mixed = np.array([srv.poll(r) for r in ids])
        srv.free()
        say(f"   device bytes {dev}, peak resident {peak_res} requests / {peak_bytes} bytes")
        assert peak_bytes <= dev and 1 < peak_res < 64    # memory, not slots, bound it
        check(mixed, ref, lvc.PRECISION_FP32_REFINE)
        srv = lvc.Server(ctx, max_slots=1)
        alone = []
        for r in reqs[:24]:
            rid = srv.submit(*r)
            srv.drain()
            alone.append(srv.poll(rid))
        srv.free()
    assert np.array_equal(np.array(alone), mixed[:24])


# ---- numX up to 1024, every x segment length, all precisions ----------------

# numX sets per x segment length: 4 keeps <= 64 lanes up to 256 points, 8 up
# to 512, 16 up to 1024 (a batch takes the segment its widest row fits).
SEG_X_SHAPES = {4: [(37, 40), (100, 33), (256, 64)],
                8: [(257, 40), (300, 64), (512, 32)],
                16: [(513, 36), (700, 64), (1024, 32)]}


@pytest.mark.parametrize("seg_x", [4, 8, 16])
@pytest.mark.parametrize("precision", [lvc.PRECISION_FP64, lvc.PRECISION_FP32_REFINE,
                                       lvc.PRECISION_FP32])
def test_wide_rows_every_x_segment(seg_x, precision):
    say(f"case: x segment {seg_x}, precision {precision}, numX up to "
        f"{max(s[0] for s in SEG_X_SHAPES[seg_x])}")
    shapes = SEG_X_SHAPES[seg_x]
    rng = np.random.default_rng(seg_x * 10 + precision)
    n = 2 * len(shapes)
    pr = dict(numX=np.array([s[0] for s in shapes] * 2, np.int32),
              numY=np.array([s[1] for s in shapes] * 2, np.int32),
              numT=rng.integers(3, 9, n).astype(np.int32),
              s0=rng.uniform(0.02, 0.05, n), strike=rng.uniform(0.0, 0.05, n),
              t=rng.uniform(1, 10, n), alpha=rng.uniform(0.1, 0.3, n),
              nu=rng.uniform(0.4, 0.8, n), beta=rng.uniform(0.3, 0.7, n))
    ref = oracle_batch(pr)
    with lvc.LocVolCalib(tuning={"lvc.seg_x": seg_x}) as ctx:
        got = ctx.main_batch(**pr, precision=precision)
    ea, er = check(got, ref, precision)
    say(f"   max abs err {ea:.3e}, / max price {er:.3e}")


@pytest.mark.parametrize("precision", [lvc.PRECISION_FP64, lvc.PRECISION_FP32_REFINE])
def test_main_numx_512(precision):
    """main at the reference's largest stated size, numX = numY = 512."""
    say(f"case: main(16, 512, 512, 6), precision {precision}")
    with lvc.LocVolCalib(precision=precision) as ctx:
        got = ctx.main(16, 512, 512, 6, 0.03, 5.0, 0.2, 0.6, 0.5)
    p = [float(np.float32(v)) for v in (0.03, 5.0, 0.2, 0.6, 0.5)]
    idx = np.arange(16)
    K = (np.float32(0.001) * idx.astype(np.float32)).astype(np.float64)
    ref = oracle.batch([512] * 16, [512] * 16, [6] * 16, [p[0]] * 16, K, [p[1]] * 16,
                       [p[2]] * 16, [p[3]] * 16, [p[4]] * 16, dtype=np.float64, thomas=True,
                       nthreads=NTH, grid32=True)
    # main's f32 output rounds the FP64 form's price: its bar is f32 rounding
    bar = 1e-6 if precision == lvc.PRECISION_FP32_REFINE else 1e-7
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    say(f"   max err / max price {err:.3e}")
    assert err < bar
