"""Futhark value I/O (text and binary formats) on the reference's datasets."""
import os

import numpy as np

from conftest import GOLDEN
from locvolcalib import futhark_values as fv


def test_text_dataset():
    vals = fv.load(os.path.join(GOLDEN, "LocVolCalib-data", "large.in"))
    assert [type(v) for v in vals[:4]] == [np.int32] * 4
    assert [v.item() for v in vals[:4]] == [256, 256, 256, 64]
    assert vals[4].dtype == np.float32 and vals[4] == np.float32(0.03)
    out = fv.load(os.path.join(GOLDEN, "LocVolCalib-data", "large.out"))[0]
    assert out.shape == (256,) and out.dtype == np.float32
    assert out[0] == np.float32(0.029998) and out[-1] == np.float32(0.000562)


def test_comment_after_value():
    vals = fv.load(os.path.join(GOLDEN, "LocVolCalib-data", "small.in"))
    assert vals[0].item() == 16 and vals[1].item() == 32  # "16 --128  -- OUTER"


def test_binary_dataset():
    vals = fv.load(os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.in"))
    assert len(vals) == 4
    for v in vals:
        assert v.shape == (100, 115) and v.dtype == np.float32
    assert 0.3 <= vals[0].min() and vals[0].max() <= 1.3


def test_text_roundtrip():
    vals = [np.int32(7), np.float64(0.25), np.array([1.5, -2.0], np.float32),
            np.array([[1, 2], [3, 4]], np.int64), np.bool_(True)]
    back = fv.loads(fv.dumps(vals))
    assert len(back) == len(vals)
    for a, b in zip(vals, back):
        assert np.asarray(a).dtype == np.asarray(b).dtype
        assert np.array_equal(a, b)


def test_binary_roundtrip():
    vals = [np.int32(-7), np.float64(0.25), np.array([1.5, -2.0], np.float32),
            np.array([[1, 2], [3, 4]], np.int64), np.bool_(True),
            np.zeros((0, 3), np.float32), np.arange(6, dtype=np.uint8).reshape(2, 3)]
    back = fv.loads(fv.dumps_binary(vals))
    assert len(back) == len(vals)
    for a, b in zip(vals, back):
        assert np.asarray(a).dtype == np.asarray(b).dtype
        assert np.asarray(a).shape == np.asarray(b).shape
        assert np.array_equal(a, b)


def test_binary_writer_matches_the_reference_dataset():
    """The ocean-sim dataset is in the binary format (written by Futhark's own
    tools): reading it and writing it back gives the same bytes."""
    path = os.path.join(GOLDEN, "ocean-sim", "tridiag32-small.in")
    with open(path, "rb") as f:
        raw = f.read()
    assert fv.dumps_binary(fv.loads(raw)) == raw.strip()
