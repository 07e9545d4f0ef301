"""Compiled callers and sanitizer builds.

  * The libraries' host C (LocVolCalib's and InterestCalib's) and the
    oracle built under AddressSanitizer and UndefinedBehaviorSanitizer
    (`make asan` in futhark-benchmarks_amd/ and oracle/), loaded into a test
    process that preloads gcc's libasan: the oracle and ABI tests on the CPU,
    and a set of GPU tests (batches, the server, the allocation-free entry,
    the tuner cache, graph replay; InterestCalib's pricing, small-population
    trajectories and error paths) on the GPU, the kernels as usual.
  * tests/native/lys_caller.c, a C program calling the library the way the
    reference's lys frontends call the generated API (shared.c:15-49,
    console/liblys.c:243-251, console/main.c:176-180), built against the
    public header: on the reference's small dataset (LocVolCalib-data/
    small.in) against small.out, plainly and under the sanitizers.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, load_dataset

NATIVE = os.path.join(ROOT, "tests", "native", "_build")
ASAN_LIB = os.path.join(ROOT, "futhark-benchmarks_amd", "_build", "asan", "liblocvolcalib.so")
ASAN_ORACLE = os.path.join(ROOT, "oracle", "_build", "liblvc_oracle_asan.so")
ASAN_IC = os.path.join(ROOT, "futhark-benchmarks_amd", "_build", "asan", "libinterestcalib.so")


def _libasan():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                           text=True, timeout=60).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return p if p and os.path.isabs(p) and os.path.exists(p) else None


def _sanitized_env():
    asan = _libasan()
    if not asan or not all(os.path.exists(p) for p in (ASAN_LIB, ASAN_ORACLE, ASAN_IC)):
        pytest.skip("sanitizer builds or gcc's libasan missing (build with __graft_entry__.build())")
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               LVC_LIBRARY=ASAN_LIB, LVC_ORACLE_LIB=ASAN_ORACLE, IC_LIBRARY=ASAN_IC)
    return env


def _pytest(args, env, timeout):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] +
                       args, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    return r


def test_oracle_and_abi_under_sanitizers():
    env = _sanitized_env()
    r = _pytest(["tests/test_oracle.py", "tests/test_op_oracle.py", "tests/test_ic_oracle.py",
                 "tests/test_abi.py", "-m", "not gpu"], env, timeout=900)
    assert r.returncode == 0
    assert " passed" in r.stdout and "AddressSanitizer" not in r.stderr


@pytest.mark.gpu
def test_gpu_paths_under_sanitizers():
    print("case: GPU tests with the host code under ASan + UBSan", flush=True)
    env = _sanitized_env()
    r = _pytest(["tests/test_gpu_parity.py", "tests/test_gpu_configs.py", "tests/test_gpu_tuning.py",
                 "tests/test_zz_no_alloc.py", "tests/test_gpu_interestcalib.py", "-m", "gpu", "-k",
                 "server or batch or budget or noalloc or autotune or graph or alone or "
                 "eval or trajectory_small_populations or main_errors"],
                env, timeout=900)
    assert r.returncode == 0
    assert " passed" in r.stdout and "AddressSanitizer" not in r.stderr


def _run_caller(exe, args, env=None):
    path = os.path.join(NATIVE, exe)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (build with __graft_entry__.build())")
    r = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stderr[-2000:], flush=True)
    return r


def _small_args():
    inp, out = load_dataset("small")
    return list(inp), np.asarray(out, np.float64)


@pytest.mark.gpu
def test_c_caller_on_small_dataset():
    print("case: lys-style C caller, small dataset", flush=True)
    args, want = _small_args()
    r = _run_caller("lys_caller", args)
    assert r.returncode == 0, r.stderr
    got = np.array([float(v) for v in r.stdout.split()])
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 5e-6


@pytest.mark.gpu
def test_c_caller_under_sanitizers():
    print("case: lys-style C caller under ASan + UBSan", flush=True)
    args, _ = _small_args()
    plain = _run_caller("lys_caller", args)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    san = _run_caller("lys_caller_asan", args, env=env)
    assert san.returncode == 0, san.stderr
    assert san.stdout == plain.stdout
