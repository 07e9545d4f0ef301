"""Per-size solver choice (the analogue of LocVolCalib.fut.tuning:1-2 and the
version selection at LocVolCalib.fut:249) on the GPU, through the C ABI:

  * a problem runs with its own size class's segment pair, so its prices are
    the same bits alone and inside a batch of other sizes;
  * the built-in choice, fitted to each size, is what an unset configuration
    uses (checked by bit equality with the pair forced through the tuning
    parameters);
  * "lvc.autotune" times the candidates, writes the choice to the cache file,
    and a new context reads it back (same bits as the pair forced);
  * graph replay: the kept graph is launched again for a repeated call and
    updated for another shape, every replay equal to direct launches.
"""
import os

import numpy as np
import pytest

from oracle import oracle
import locvolcalib as lvc

pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
KEYS = ("numX", "numY", "numT", "s0", "strike", "t", "alpha", "nu", "beta")
PRECS = [lvc.PRECISION_FP32, lvc.PRECISION_FP32_REFINE, lvc.PRECISION_FP64]


def say(msg):
    print(msg, flush=True)


def problems(shapes, seed, nt=(6, 14)):
    rng = np.random.default_rng(seed)
    n = len(shapes)
    return dict(numX=np.array([s[0] for s in shapes], np.int32),
                numY=np.array([s[1] for s in shapes], np.int32),
                numT=rng.integers(nt[0], nt[1], n).astype(np.int32),
                s0=rng.uniform(0.02, 0.05, n), strike=rng.uniform(0.0, 0.05, n),
                t=rng.uniform(1, 10, n), alpha=rng.uniform(0.1, 0.3, n),
                nu=rng.uniform(0.4, 0.8, n), beta=rng.uniform(0.3, 0.7, n))


def take(pr, idx):
    return {k: np.asarray(v)[idx] for k, v in pr.items()}


# sizes whose fitted segment pairs differ (rows wider than 512 take x 16, in
# FP64 rows of 512 take x 8; columns taller than 256 take y 64), so that a
# mixed batch runs as several sub-batches
SHAPES = [(32, 32), (600, 40), (64, 48), (256, 256), (48, 300), (128, 32), (512, 64), (40, 30)]


@pytest.mark.parametrize("precision", PRECS)
def test_problem_bits_do_not_depend_on_the_batch(precision):
    say(f"case: alone vs mixed batch, precision {precision}")
    pr = problems(SHAPES, seed=3 + precision)
    with lvc.LocVolCalib() as ctx:
        mixed = ctx.main_batch(**pr, precision=precision)
        alone = np.array([ctx.main_batch(**take(pr, [i]), precision=precision)[0]
                          for i in range(len(SHAPES))])
        # and in another order, with the same problems twice
        order = [5, 0, 7, 2, 2, 1, 6, 3, 4, 0]
        again = ctx.main_batch(**take(pr, order), precision=precision)
    assert np.array_equal(mixed, alone)
    assert np.array_equal(again, mixed[order])
    ref = oracle.batch(*[pr[k] for k in KEYS], dtype=np.float64, thomas=True, nthreads=NTH)
    bar = {lvc.PRECISION_FP64: 1e-10, lvc.PRECISION_FP32_REFINE: 1e-6,
           lvc.PRECISION_FP32: 1e-4}[precision]
    assert np.max(np.abs(mixed - ref)) / np.max(np.abs(ref)) < bar


# (precision, numX, numY) -> the built-in pair: x 8 / y 32 (FP64 x 4), fitted
# to the size (rows of 512 need 8 per lane for <= 64 lanes; columns taller
# than 8 segments of 32 take 64, then lane segments of 16)
BUILTIN = [(lvc.PRECISION_FP32_REFINE, 32, 32, 8, 32), (lvc.PRECISION_FP32_REFINE, 256, 256, 8, 32),
           (lvc.PRECISION_FP32_REFINE, 1024, 300, 16, 64), (lvc.PRECISION_FP64, 128, 128, 4, 32),
           (lvc.PRECISION_FP64, 512, 64, 8, 32), (lvc.PRECISION_FP64, 64, 520, 4, 16),
           (lvc.PRECISION_FP32, 32, 32, 8, 32), (lvc.PRECISION_FP32, 64, 512, 8, 64)]


@pytest.mark.parametrize("precision,nx,ny,sx,sy", BUILTIN)
def test_builtin_table(precision, nx, ny, sx, sy):
    say(f"case: built-in pair for precision {precision} {nx}x{ny}: {sx}/{sy}")
    pr = problems([(nx, ny)] * 6, seed=nx + ny)
    with lvc.LocVolCalib() as ctx:
        got = ctx.main_batch(**pr, precision=precision)
    with lvc.LocVolCalib(tuning={"lvc.seg_x": sx, "lvc.seg_y": sy}) as ctx:
        forced = ctx.main_batch(**pr, precision=precision)
    assert np.array_equal(got, forced)


def _cache_lines(path):
    with open(path) as f:
        return [tuple(int(v) for v in ln.split()) for ln in f if ln.strip()]


def test_autotune_writes_and_reloads_the_cache(tmp_path):
    cache = str(tmp_path / "lvc.tuning")
    shapes = [(48, 40)] * 4 + [(96, 64)] * 3
    pr = problems(shapes, seed=17)
    say("case: autotune on, cache file written")
    with lvc.LocVolCalib(cache_file=cache, tuning={"lvc.autotune": 1}) as ctx:
        first = ctx.main_batch(**pr, precision=lvc.PRECISION_FP32_REFINE)
        repeat = ctx.main_batch(**pr, precision=lvc.PRECISION_FP32_REFINE)
    assert np.array_equal(first, repeat)
    lines = _cache_lines(cache)
    classes = {(p, nx, ny): (sx, sy) for p, nx, ny, sx, sy in lines}
    assert set(classes) == {(lvc.PRECISION_FP32_REFINE, 48, 40),
                            (lvc.PRECISION_FP32_REFINE, 96, 64)}
    ref = oracle.batch(*[pr[k] for k in KEYS], dtype=np.float64, thomas=True, nthreads=NTH)
    assert np.max(np.abs(first - ref)) / np.max(np.abs(ref)) < 1e-6
    say("case: a new context reads the cache (no tuning)")
    with lvc.LocVolCalib(cache_file=cache) as ctx:
        reloaded = ctx.main_batch(**pr, precision=lvc.PRECISION_FP32_REFINE)
    assert np.array_equal(reloaded, first)
    assert _cache_lines(cache) == lines               # nothing timed again
    # each class alone, with its cached pair forced, gives the same bits
    for (p, nx, ny), (sx, sy) in classes.items():
        idx = [i for i, s in enumerate(shapes) if s == (nx, ny)]
        with lvc.LocVolCalib(tuning={"lvc.seg_x": sx, "lvc.seg_y": sy}) as ctx:
            forced = ctx.main_batch(**take(pr, idx), precision=p)
        assert np.array_equal(forced, first[idx])


def test_graph_replay_kept_and_updated():
    """lvc.graph: a repeated call relaunches the kept graph, another shape
    updates it; every result equals direct launches, with the program's own
    eager calls and synchronisations between the replays."""
    say("case: graph replay, repeated and changing shapes")
    pa = problems([(64, 64)] * 4 + [(32, 32)] * 3, seed=5)
    pb = problems([(128, 64)] * 5, seed=6)
    with lvc.LocVolCalib() as ctx:
        want_a = ctx.main_batch(**pa, precision=lvc.PRECISION_FP32_REFINE)
        want_b = ctx.main_batch(**pb, precision=lvc.PRECISION_FP32_REFINE)
        want_m = ctx.main(16, 64, 64, 9, 0.03, 5.0, 0.2, 0.6, 0.5)
    with lvc.LocVolCalib(tuning={"lvc.graph": 1}) as ctx, \
            lvc.LocVolCalib() as eager:
        for rep in range(3):
            assert np.array_equal(ctx.main_batch(**pa, precision=lvc.PRECISION_FP32_REFINE),
                                  want_a), rep
            assert np.array_equal(ctx.main_batch(**pa, precision=lvc.PRECISION_FP32_REFINE),
                                  want_a), rep
            eager.main(4, 32, 32, 5, 0.03, 5.0, 0.2, 0.6, 0.5)   # eager work between
            eager.sync()
            assert np.array_equal(ctx.main_batch(**pb, precision=lvc.PRECISION_FP32_REFINE),
                                  want_b), rep
            assert np.array_equal(ctx.main(16, 64, 64, 9, 0.03, 5.0, 0.2, 0.6, 0.5), want_m), rep
