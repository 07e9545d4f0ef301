"""A ``futhark bench``-shaped runner for the library's entry points.

Reads the datasets of a program's benchmark stanza -- LocVolCalib.fut:1-10
(``compiled input @ LocVolCalib-data/small.in`` / ``output @ ...small.out``),
OptionPricing.fut:1-10, InterestCalib.fut:4-13, or
misc/ocean-sim/tridiag-test.fut:6-12 (its four entries on
``data/tridiag32-small.in``) -- in Futhark's text or binary value format,
runs each entry on each dataset a number of times through the C ABI, checks
the result against the expected output, and reports per-run wall times in
microseconds: on stdout as one line per dataset, and with ``--json FILE`` in
the JSON layout ``futhark bench --json`` writes (README.md:36-37)::

    {"finpar/LocVolCalib.fut:main":
       {"datasets": {"LocVolCalib-data/small.in": {"runtimes": [...],
                                                    "stderr": ""}}}}

A run is the entry call plus the context sync that ends it, with the inputs
already on the device and the result left there; one warm-up run precedes the
timed ones, as ``futhark bench`` does.

Usage::

    python -m locvolcalib.bench_runner --program path/to/finpar/LocVolCalib.fut
    python -m locvolcalib.bench_runner --program path/to/misc/ocean-sim/tridiag-test.fut
    python -m locvolcalib.bench_runner --program path/to/finpar/OptionPricing.fut
    python -m locvolcalib.bench_runner --program path/to/finpar/InterestCalib.fut
    python -m locvolcalib.bench_runner path/to/small.in path/to/medium.in ...

The expected output of ``X.in`` is ``X.out`` beside it (or the stanza's
``output @`` file).  Results are compared element-wise with an absolute
tolerance (default 5e-6: the reference's f32 outputs are printed to 6
significant digits and carry their own rounding).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import futhark_values

PROGRAM_KEY = "finpar/LocVolCalib.fut:main"
OP_PROGRAM_KEY = "finpar/OptionPricing.fut:main"
IC_PROGRAM_KEY = "finpar/InterestCalib.fut:main"
OCEAN_PROGRAM = "misc/ocean-sim/tridiag-test.fut"
OCEAN_ENTRIES = ("tridagNested", "tridagNestedConst", "tridagNestedSeq", "tridagNestedSeqConst")


def parse_entries(text: str) -> List[str]:
    """Entry names of a stanza's ``entry:`` line (tridiag-test.fut:7); the
    default is ``main``."""
    for ln in text.splitlines():
        m = re.match(r"\s*--\s*entry:\s*(.+)$", ln)
        if m:
            return m.group(1).split()
    return ["main"]


def parse_stanza(text: str) -> List[Tuple[str, Optional[str]]]:
    """(input, output) dataset paths of a Futhark test stanza: the comment
    block that starts at ``-- ==`` (LocVolCalib.fut:2-10)."""
    lines = text.splitlines()
    try:
        start = next(i for i, ln in enumerate(lines) if ln.strip().startswith("-- =="))
    except StopIteration:
        return []
    pairs: List[Tuple[str, Optional[str]]] = []
    for ln in lines[start + 1:]:
        s = ln.strip()
        if not s.startswith("--"):
            break
        body = s[2:].strip()
        m = re.match(r"(?:compiled\s+)?input\s*@\s*(\S+)", body)
        if m:
            pairs.append((m.group(1), None))
            continue
        m = re.match(r"output\s*@\s*(\S+)", body)
        if m and pairs and pairs[-1][1] is None:
            pairs[-1] = (pairs[-1][0], m.group(1))
    return pairs


def main_arguments(values) -> tuple:
    """The nine arguments of ``main`` (LocVolCalib.fut:238-246) from a
    dataset's values: outer, numX, numY, numT (i32) and s0, t, alpha, nu,
    beta (f32)."""
    if len(values) != 9:
        raise ValueError(f"main takes 9 values, the dataset has {len(values)}")
    ints = tuple(int(v) for v in values[:4])
    reals = tuple(float(np.float32(v)) for v in values[4:])
    return ints + reals


def compare(got: np.ndarray, expected: np.ndarray, abs_tol: float) -> Tuple[bool, float]:
    got = np.asarray(got, np.float64)
    expected = np.asarray(expected, np.float64)
    if got.shape != expected.shape:
        return False, float("inf")
    if got.size == 0:
        return True, 0.0
    err = float(np.max(np.abs(got - expected)))
    return err <= abs_tol, err


def run_dataset(ctx, args: tuple, precision: int, runs: int) -> Tuple[List[int], np.ndarray]:
    """Timed runs of main on one dataset: (runtimes in us, the result)."""
    if precision == 2:
        raw, free, fetch = ctx.main_f64_raw, ctx.free_f64, ctx.values_f64
    else:
        raw, free, fetch = ctx.main_raw, ctx.free_f32, ctx.values_f32
    free(raw(*args))  # warm-up
    ctx.sync()
    times = []
    result = None
    for r in range(runs):
        t0 = time.perf_counter()
        h = raw(*args)
        ctx.sync()
        times.append(int(round((time.perf_counter() - t0) * 1e6)))
        if r == runs - 1:
            result = fetch(h)
        free(h)
    return times, result


def bench(datasets: List[Tuple[str, Optional[str]]], runs: int = 10, precision: int = 1,
          abs_tol: float = 5e-6, device: int = 0, names: Optional[List[str]] = None,
          out=sys.stdout) -> Tuple[Dict, bool]:
    """Runs every dataset; returns (the futhark-bench JSON object, all valid)."""
    from . import LocVolCalib

    results: Dict[str, Dict] = {}
    all_ok = True
    with LocVolCalib(device=device, precision=precision) as ctx:
        for i, (inp, outp) in enumerate(datasets):
            name = names[i] if names else inp
            args = main_arguments(futhark_values.load(inp))
            times, got = run_dataset(ctx, args, precision, runs)
            status = ""
            if outp is None:
                cand = os.path.splitext(inp)[0] + ".out"
                outp = cand if os.path.exists(cand) else None
            if outp is not None:
                expected = futhark_values.load(outp)[0]
                ok, err = compare(got, expected, abs_tol)
                status = f"max abs err {err:.2e} {'ok' if ok else 'FAILED'}"
                all_ok &= ok
            mean = float(np.mean(times))
            rsd = float(np.std(times) / mean) if mean > 0 else 0.0
            print(f"{name}: {mean:10.0f}us (RSD: {rsd:.3f}; min: {min(times)}us) {status}",
                  file=out, flush=True)
            results[name] = {"runtimes": times, "stderr": "" if not status.endswith("FAILED")
                             else f"result differs from the expected output ({status})"}
    return {PROGRAM_KEY: {"datasets": results}}, all_ok


def bench_ocean(datasets: List[Tuple[str, Optional[str]]], entries: List[str], runs: int = 10,
                device: int = 0, names: Optional[List[str]] = None, out=sys.stdout,
                abs_tol: float = 0.0) -> Tuple[Dict, bool]:
    """The ocean-sim entries on [n][m] f32 datasets (four arrays a, b, c, y).
    The Seq entries are checked for equality with the expected output (the
    bits of a sequential f32 Thomas solve); the scan-form entries, which
    associate differently, by median relative difference (< 1e-5)."""
    from . import LocVolCalib

    results: Dict[str, Dict] = {}
    all_ok = True
    with LocVolCalib(device=device) as ctx:
        for entry in entries:
            if entry not in OCEAN_ENTRIES:
                raise ValueError(f"unknown entry {entry}")
            seq, const = "Seq" in entry, entry.endswith("Const")
            sets: Dict[str, Dict] = {}
            for i, (inp, outp) in enumerate(datasets):
                name = names[i] if names else inp
                vals = futhark_values.load(inp)
                if len(vals) != 4:
                    raise ValueError(f"{entry} takes 4 arrays, the dataset has {len(vals)}")
                hs = [ctx.upload_f32_2d(v) for v in vals]
                try:
                    ctx.free_f32_2d(ctx.tridag_nested_raw(*hs, seq=seq, const=const))
                    ctx.sync()
                    times = []
                    for _ in range(runs):
                        t0 = time.perf_counter()
                        h = ctx.tridag_nested_raw(*hs, seq=seq, const=const)
                        ctx.sync()
                        times.append(int(round((time.perf_counter() - t0) * 1e6)))
                        ctx.free_f32_2d(h)
                finally:
                    for h in hs:
                        ctx.free_f32_2d(h)
                status = ""
                if outp is None:
                    cand = os.path.splitext(inp)[0] + ".out"
                    outp = cand if os.path.exists(cand) else None
                if outp is not None:
                    got = ctx.tridag_nested(*vals, seq=seq, const=const)
                    expected = futhark_values.load(outp)[0]
                    if seq:
                        ok, err = compare(got, expected, abs_tol)
                        status = f"max abs err {err:.2e} {'ok' if ok else 'FAILED'}"
                    else:
                        rel = np.abs(got.astype(np.float64) - expected) / np.abs(expected)
                        med = float(np.median(rel)) if rel.size else 0.0
                        ok = med < 1e-5
                        status = f"median rel diff {med:.2e} {'ok' if ok else 'FAILED'}"
                    all_ok &= ok
                mean = float(np.mean(times))
                print(f"{entry} {name}: {mean:10.0f}us (min: {min(times)}us) {status}",
                      file=out, flush=True)
                sets[name] = {"runtimes": times,
                              "stderr": "" if not status.endswith("FAILED") else status}
            results[f"{OCEAN_PROGRAM}:{entry}"] = {"datasets": sets}
    return results, all_ok


def bench_option_pricing(datasets: List[Tuple[str, Optional[str]]], runs: int = 10,
                         device: int = 0, names: Optional[List[str]] = None, out=sys.stdout,
                         rtol: float = 2e-3) -> Tuple[Dict, bool]:
    """finpar/OptionPricing.fut's ``main`` on its datasets (eleven values:
    OptionPricing.fut:295-306).  The expected outputs are the f32 sums of a
    sequential backend; the library sums the f32 payoffs in f64, so results
    are compared with a relative tolerance (default 2e-3, above the 1.1e-3
    f32 rounding of the large dataset's expected output)."""
    from . import optionpricing as op

    results: Dict[str, Dict] = {}
    all_ok = True
    with op.OptionPricing(device=device) as ctx:
        for i, (inp, outp) in enumerate(datasets):
            name = names[i] if names else inp
            contract, n, *arrays = op.main_arguments(futhark_values.load(inp))
            hs = ctx.upload_arguments(arrays)
            try:
                ctx.free_f32(ctx.main_raw(contract, n, hs))
                ctx.sync()
                times = []
                got = None
                for r in range(runs):
                    t0 = time.perf_counter()
                    h = ctx.main_raw(contract, n, hs)
                    ctx.sync()
                    times.append(int(round((time.perf_counter() - t0) * 1e6)))
                    if r == runs - 1:
                        got = ctx.values_f32(h)
                    ctx.free_f32(h)
            finally:
                ctx.free_arguments(hs)
            status = ""
            if outp is None:
                cand = os.path.splitext(inp)[0] + ".out"
                outp = cand if os.path.exists(cand) else None
            if outp is not None:
                expected = np.asarray(futhark_values.load(outp)[0], np.float64)
                rel = (float(np.max(np.abs(got - expected) / np.abs(expected)))
                       if expected.size else 0.0)
                ok = got.shape == expected.shape and rel <= rtol
                status = f"max rel err {rel:.2e} {'ok' if ok else 'FAILED'}"
                all_ok &= ok
            mean = float(np.mean(times))
            print(f"{name}: {mean:10.0f}us (min: {min(times)}us; {n / (min(times) * 1e-6):.3g} "
                  f"paths/s) {status}", file=out, flush=True)
            results[name] = {"runtimes": times,
                             "stderr": "" if not status.endswith("FAILED") else status}
    return {OP_PROGRAM_KEY: {"datasets": results}}, all_ok


def bench_interest_calib(datasets: List[Tuple[str, Optional[str]]], runs: int = 10,
                         device: int = 0, names: Optional[List[str]] = None, out=sys.stdout,
                         rtol: float = 1e-5) -> Tuple[Dict, bool]:
    """finpar/InterestCalib.fut's ``main`` on its datasets (nine values:
    InterestCalib.fut:929-932).  The calibration's winner depends on accept
    decisions that one ulp of a log-likelihood can flip (the reference's own
    expected winners are one of several near-equal optima), so the check
    covers what every backend reproduces: the summary's shape and its Black
    quote column, which does not depend on the genome (relative tolerance
    1e-5); the winner's log-likelihood is printed beside the expected one."""
    from . import interestcalib as icl

    results: Dict[str, Dict] = {}
    all_ok = True
    with icl.InterestCalib(device=device) as ctx:
        for i, (inp, outp) in enumerate(datasets):
            name = names[i] if names else inp
            a = icl.main_arguments(futhark_values.load(inp))
            hs = ctx.upload_inputs(a["swaptions"], a["herm_x"], a["herm_w"], a["dir_vct"])
            try:
                o, h = ctx.main_raw(a["pop"], a["mcmc_conv"], hs)
                ctx.free_summary(h)
                times = []
                for r in range(runs):
                    t0 = time.perf_counter()
                    o, h = ctx.main_raw(a["pop"], a["mcmc_conv"], hs)
                    times.append(int(round((time.perf_counter() - t0) * 1e6)))
                    if r == runs - 1:
                        shape = ctx._L.futhark_shape_f32_2d(ctx._ctx, h)
                        summ = np.empty((shape[0], shape[1]), np.float32)
                        if summ.size:
                            ctx._check(ctx._L.futhark_values_f32_2d(ctx._ctx, h,
                                                                     summ.ctypes.data),
                                       "futhark_values_f32_2d")
                    ctx.free_summary(h)
            finally:
                ctx.free_inputs(hs)
            status = f"loglik {float(o[5]):.3f}"
            if outp is None:
                cand = os.path.splitext(inp)[0] + ".out"
                outp = cand if os.path.exists(cand) else None
            if outp is not None:
                exp = futhark_values.load(outp)
                esum = np.asarray(exp[6], np.float64)
                ok = summ.shape == esum.shape
                rel = float(np.max(np.abs(summ[:, 1] - esum[:, 1]) / np.abs(esum[:, 1]))) \
                    if ok and esum.size else 0.0
                ok = ok and rel <= rtol
                status += (f" (expected {float(exp[5]):.3f}); quote column max rel err "
                           f"{rel:.2e} {'ok' if ok else 'FAILED'}")
                all_ok &= ok
            mean = float(np.mean(times))
            print(f"{name}: {mean:10.0f}us (min: {min(times)}us; "
                  f"{a['mcmc_conv'] / (min(times) * 1e-6):.3g} iterations/s) {status}",
                  file=out, flush=True)
            results[name] = {"runtimes": times,
                             "stderr": "" if not status.endswith("FAILED") else status}
    return {IC_PROGRAM_KEY: {"datasets": results}}, all_ok


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("datasets", nargs="*", help="dataset .in files")
    ap.add_argument("--program", help="the reference's LocVolCalib.fut, OptionPricing.fut, "
                    "InterestCalib.fut or tridiag-test.fut: run its stanza's datasets")
    ap.add_argument("--option-pricing", action="store_true",
                    help="the datasets given are OptionPricing.fut's")
    ap.add_argument("--interest-calib", action="store_true",
                    help="the datasets given are InterestCalib.fut's")
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--precision", choices=("fp32", "fp32r", "fp64"), default="fp32r")
    ap.add_argument("--tolerance", type=float, default=5e-6)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--json", help="write the futhark-bench JSON here")
    a = ap.parse_args(argv)
    pairs: List[Tuple[str, Optional[str]]] = []
    names: List[str] = []
    entries = ["main"]
    option_pricing = a.option_pricing
    interest_calib = a.interest_calib
    if a.program:
        option_pricing |= os.path.basename(a.program) == "OptionPricing.fut"
        interest_calib |= os.path.basename(a.program) == "InterestCalib.fut"
        base = os.path.dirname(os.path.abspath(a.program))
        with open(a.program) as f:
            text = f.read()
        entries = parse_entries(text)
        for inp, outp in parse_stanza(text):
            pairs.append((os.path.join(base, inp),
                          os.path.join(base, outp) if outp else None))
            names.append(inp)
    for d in a.datasets:
        pairs.append((d, None))
        names.append(d)
    if not pairs:
        ap.error("no datasets: give .in files or --program")
    if option_pricing:
        res, ok = bench_option_pricing(pairs, runs=a.runs, device=a.device, names=names)
    elif interest_calib:
        res, ok = bench_interest_calib(pairs, runs=a.runs, device=a.device, names=names)
    elif entries != ["main"]:
        res, ok = bench_ocean(pairs, entries, runs=a.runs, device=a.device, names=names)
    else:
        prec = {"fp32": 0, "fp32r": 1, "fp64": 2}[a.precision]
        res, ok = bench(pairs, runs=a.runs, precision=prec, abs_tol=a.tolerance,
                        device=a.device, names=names)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
