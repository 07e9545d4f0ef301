"""The allocation-free entry, futhark_entry_main_noalloc, through the C ABI:
its prices against futhark_entry_main's (the same kernels and tiles, so the
same bits) on two calls in a row, the second smaller than the first (the
static scratch keeps the first call's contents), in every precision; the
size limit and the reference's errors."""
import ctypes

import numpy as np
import pytest

import locvolcalib as lvc

pytestmark = pytest.mark.gpu

P = (0.03, 5.0, 0.2, 0.6, 0.5)


def say(msg):
    print(msg, flush=True)


# (first call, second call): the second is smaller in every size
CALLS = {
    lvc.PRECISION_FP32_REFINE: [(128, 256, 256, 12), (20, 64, 96, 9)],
    lvc.PRECISION_FP32: [(96, 128, 256, 10), (7, 33, 40, 6)],
    lvc.PRECISION_FP64: [(64, 256, 256, 8), (10, 32, 48, 5)],
}


@pytest.mark.parametrize("precision", sorted(CALLS))
def test_noalloc_matches_main_twice(precision):
    with lvc.LocVolCalib(precision=precision) as ctx:
        for outer, nx, ny, nt in CALLS[precision]:
            say(f"case: main_noalloc({outer}, {nx}, {ny}, {nt}) precision {precision}")
            got = ctx.main_noalloc(outer, nx, ny, nt, *P)
            want = ctx.main(outer, nx, ny, nt, *P)
            assert got.shape == want.shape == (outer,)
            assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_noalloc_repeat_and_other_parameters():
    """Back to back on one context with other parameters, then the first
    call again: each equals main's result for its own arguments."""
    runs = [(16, 64, 64, 20, (0.03, 5.0, 0.2, 0.6, 0.5)),
            (40, 128, 32, 7, (0.025, 2.0, 0.15, 0.5, 0.4)),
            (16, 64, 64, 20, (0.03, 5.0, 0.2, 0.6, 0.5))]
    with lvc.LocVolCalib() as ctx:
        for outer, nx, ny, nt, p in runs:
            say(f"case: main_noalloc({outer}, {nx}, {ny}, {nt}, {p})")
            assert np.array_equal(ctx.main_noalloc(outer, nx, ny, nt, *p),
                                  ctx.main(outer, nx, ny, nt, *p))


def test_noalloc_size_limit_and_errors():
    with lvc.LocVolCalib() as ctx:
        say("case: main_noalloc beyond the static scratch")
        with pytest.raises(lvc.LocVolCalibError, match="needs .* bytes of device scratch"):
            ctx.main_noalloc(256, 256, 256, 8, *P)
        say("case: main_noalloc price index out of bounds, bad sizes, NULL output")
        with pytest.raises(lvc.LocVolCalibError, match="out of bounds"):
            ctx.main_noalloc(4, 32, 32, 8, 0.03, 0.25, 0.1, 0.6, 0.5)
        with pytest.raises(lvc.LocVolCalibError, match="at least 3"):
            ctx.main_noalloc(4, 2, 32, 8, *P)
        with pytest.raises(lvc.LocVolCalibError, match="NULL output"):
            ctx.main_noalloc_raw(None, 4, 32, 32, 8, *P)
        assert ctx.main_noalloc(0, 32, 32, 8, *P).shape == (0,)
        # the context still works after the refusals
        assert np.array_equal(ctx.main_noalloc(8, 32, 32, 8, *P), ctx.main(8, 32, 32, 8, *P))


def test_noalloc_into_a_caller_array():
    """The caller's own device array (made with futhark_new_f32_1d) receives
    the prices; its values beyond `outer` are left alone."""
    with lvc.LocVolCalib() as ctx:
        L = ctx._L
        init = np.full(12, -7.0, np.float32)
        h = L.futhark_new_f32_1d(ctx.handle, init.ctypes.data, 12)
        assert h
        try:
            ptr = L.futhark_values_raw_f32_1d(ctx.handle, h)
            ctx.main_noalloc_raw(ptr, 10, 32, 64, 6, *P)
            ctx.sync()
            vals = ctx.values_f32(h)
        finally:
            L.futhark_free_f32_1d(ctx.handle, h)
        assert np.array_equal(vals[:10], ctx.main(10, 32, 64, 6, *P))
        assert np.all(vals[10:] == np.float32(-7.0))
