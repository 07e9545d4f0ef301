"""The futhark-bench-shaped dataset runner: stanza parsing, argument
mapping and validation on the CPU; a real run over the reference's three
datasets on the GPU."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from locvolcalib import bench_runner, futhark_values

# The reference's stanza (LocVolCalib.fut:1-10), as data.
STANZA = """-- LocVolCalib
-- ==
-- compiled input @ LocVolCalib-data/small.in
-- output @ LocVolCalib-data/small.out
--
-- compiled input @ LocVolCalib-data/medium.in
-- output @ LocVolCalib-data/medium.out
--
-- compiled input @ LocVolCalib-data/large.in
-- output @ LocVolCalib-data/large.out

def initGrid (s0: f32) = s0
"""


def test_parse_stanza():
    pairs = bench_runner.parse_stanza(STANZA)
    assert pairs == [("LocVolCalib-data/%s.in" % n, "LocVolCalib-data/%s.out" % n)
                     for n in ("small", "medium", "large")]
    assert bench_runner.parse_stanza("def f = 1") == []


def test_main_arguments_from_datasets():
    args = bench_runner.main_arguments(
        futhark_values.load(os.path.join(GOLDEN, "LocVolCalib-data", "small.in")))
    assert args[:4] == (16, 32, 256, 256)
    assert np.allclose(args[4:], (0.03, 5.0, 0.2, 0.6, 0.5))
    with pytest.raises(ValueError):
        bench_runner.main_arguments([np.int32(1)])


def test_compare():
    ok, err = bench_runner.compare(np.array([1.0, 2.0]), np.array([1.0, 2.0 + 1e-6]), 5e-6)
    assert ok and err < 2e-6
    ok, _ = bench_runner.compare(np.array([1.0]), np.array([1.1]), 5e-6)
    assert not ok
    ok, err = bench_runner.compare(np.zeros(2), np.zeros(3), 5e-6)
    assert not ok and err == float("inf")


@pytest.mark.gpu
def test_runner_on_reference_datasets(tmp_path):
    print("case: bench runner on small/medium/large", flush=True)
    ds = [(os.path.join(GOLDEN, "LocVolCalib-data", n + ".in"), None)
          for n in ("small", "medium", "large")]
    res, ok = bench_runner.bench(ds, runs=2, names=["small", "medium", "large"])
    assert ok
    sets = res[bench_runner.PROGRAM_KEY]["datasets"]
    assert set(sets) == {"small", "medium", "large"}
    assert all(len(v["runtimes"]) == 2 and min(v["runtimes"]) > 0 for v in sets.values())
    out = tmp_path / "r.json"
    assert bench_runner.main([ds[0][0], "--runs", "1", "--json", str(out)]) == 0
    assert bench_runner.PROGRAM_KEY in json.loads(out.read_text())


# The ocean-sim stanza (misc/ocean-sim/tridiag-test.fut:6-12), as data.
OCEAN_STANZA = """-- ==
-- entry: tridagNested tridagNestedConst tridagNestedSeq tridagNestedSeqConst
--
-- compiled input @ data/tridiag32-small.in
-- output @ data/tridiag32-small.out
--
-- compiled random input { [57600][115]f32 [57600][115]f32 [57600][115]f32 [57600][115]f32 }

entry tridagNested [n] [m] (a: [n][m]DTYPE) = a
"""


def test_parse_entries():
    assert bench_runner.parse_entries(OCEAN_STANZA) == list(bench_runner.OCEAN_ENTRIES)
    assert bench_runner.parse_entries(STANZA) == ["main"]
    assert bench_runner.parse_stanza(OCEAN_STANZA) == [("data/tridiag32-small.in",
                                                        "data/tridiag32-small.out")]


@pytest.mark.gpu
def test_runner_on_ocean_dataset(tmp_path):
    print("case: bench runner, four ocean-sim entries on tridiag32-small", flush=True)
    base = os.path.join(GOLDEN, "ocean-sim")
    ds = [(os.path.join(base, "tridiag32-small.in"), os.path.join(base, "tridiag32-small.out"))]
    res, ok = bench_runner.bench_ocean(ds, list(bench_runner.OCEAN_ENTRIES), runs=2,
                                       names=["data/tridiag32-small.in"])
    assert ok
    assert set(res) == {f"{bench_runner.OCEAN_PROGRAM}:{e}" for e in bench_runner.OCEAN_ENTRIES}
    for v in res.values():
        rt = v["datasets"]["data/tridiag32-small.in"]["runtimes"]
        assert len(rt) == 2 and min(rt) > 0
    prog = tmp_path / "tridiag-test.fut"
    prog.write_text(OCEAN_STANZA.replace("data/", base + "/"))
    out = tmp_path / "o.json"
    assert bench_runner.main(["--program", str(prog), "--runs", "1", "--json", str(out)]) == 0
    assert len(json.loads(out.read_text())) == 4


# A stanza in OptionPricing.fut's layout (:1-10): three datasets, each with
# its expected output.
OP_STANZA = """-- Generic pricing
-- ==
-- compiled input @ OptionPricing-data/small.in
-- output @ OptionPricing-data/small.out
--
-- compiled input @ OptionPricing-data/medium.in
-- output @ OptionPricing-data/medium.out
--
-- compiled input @ OptionPricing-data/large.in
-- output @ OptionPricing-data/large.out
"""


def test_parse_option_pricing_stanza():
    pairs = bench_runner.parse_stanza(OP_STANZA)
    assert [p[0] for p in pairs] == [f"OptionPricing-data/{n}.in"
                                     for n in ("small", "medium", "large")]
    assert all(o == i[:-3] + ".out" for i, o in pairs)
    assert bench_runner.parse_entries(OP_STANZA) == ["main"]


@pytest.mark.gpu
def test_runner_on_option_pricing_datasets(tmp_path):
    print("case: bench runner, OptionPricing small/medium/large", flush=True)
    prog = tmp_path / "OptionPricing.fut"
    prog.write_text(OP_STANZA)
    os.symlink(os.path.join(GOLDEN, "OptionPricing-data"), tmp_path / "OptionPricing-data")
    out = tmp_path / "o.json"
    assert bench_runner.main(["--program", str(prog), "--runs", "2", "--json", str(out)]) == 0
    res = json.loads(out.read_text())[bench_runner.OP_PROGRAM_KEY]["datasets"]
    assert set(res) == {f"OptionPricing-data/{n}.in" for n in ("small", "medium", "large")}
    for v in res.values():
        assert len(v["runtimes"]) == 2 and min(v["runtimes"]) > 0 and v["stderr"] == ""


# InterestCalib.fut:1-13's stanza, as data.
IC_STANZA = """-- Interest rate calibration
-- ==
-- tags { nobench no_opencl }
-- compiled input @ InterestCalib-data/small.in
-- output @ InterestCalib-data/small.out
--
-- compiled input @ InterestCalib-data/medium.in
-- output @ InterestCalib-data/medium.out
--
-- compiled input @ InterestCalib-data/large.in
-- output @ InterestCalib-data/large.out
"""


def test_parse_interest_calib_stanza():
    pairs = bench_runner.parse_stanza(IC_STANZA)
    assert [p[0] for p in pairs] == [f"InterestCalib-data/{n}.in"
                                     for n in ("small", "medium", "large")]
    assert bench_runner.parse_entries(IC_STANZA) == ["main"]


@pytest.mark.gpu
def test_runner_on_interest_calib_datasets(tmp_path, capsys):
    print("case: bench runner, InterestCalib small/medium/large", flush=True)
    prog = tmp_path / "InterestCalib.fut"
    prog.write_text(IC_STANZA)
    os.symlink(os.path.join(GOLDEN, "InterestCalib-data"), tmp_path / "InterestCalib-data")
    out = tmp_path / "o.json"
    # large.out is a copy of medium.out, whose quotes are medium's swaptions':
    # the runner reports that dataset's check as failed
    assert bench_runner.main(["--program", str(prog), "--runs", "2", "--json", str(out)]) == 1
    res = json.loads(out.read_text())[bench_runner.IC_PROGRAM_KEY]["datasets"]
    assert set(res) == {f"InterestCalib-data/{n}.in" for n in ("small", "medium", "large")}
    for n, v in res.items():
        assert len(v["runtimes"]) == 2 and min(v["runtimes"]) > 0
        assert (v["stderr"] == "") == (not n.endswith("large.in")), (n, v["stderr"])
    assert "quote column" in res["InterestCalib-data/large.in"]["stderr"]
