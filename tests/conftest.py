"""Shared test setup: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "futhark-benchmarks_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


# The reference's datasets (finpar/LocVolCalib-data, copied verbatim).
DATASETS = ("small", "medium", "large")


def load_dataset(name):
    from locvolcalib import futhark_values as fv
    inp = [v.item() for v in fv.load(os.path.join(GOLDEN, "LocVolCalib-data", f"{name}.in"))]
    out = fv.load(os.path.join(GOLDEN, "LocVolCalib-data", f"{name}.out"))[0]
    return inp, np.asarray(out)


def load_ic_dataset(name):
    """finpar/InterestCalib-data/<name>.{in,out} (copied verbatim): main's
    arguments (pop, mcmc_conv, swaptions [n][4], hermCoefs, hermWeights,
    sobDirVct) and the expected outputs (six f32 scalars, summary [n][3])."""
    from locvolcalib import futhark_values as fv
    d = os.path.join(GOLDEN, "InterestCalib-data")
    v = fv.load(os.path.join(d, f"{name}.in"))
    o = fv.load(os.path.join(d, f"{name}.out"))
    args = dict(pop=int(v[0]), mcmc_conv=int(v[1]),
                swaptions=np.ascontiguousarray(v[3], np.float32),
                herm_x=np.ascontiguousarray(v[5], np.float32),
                herm_w=np.ascontiguousarray(v[6], np.float32),
                dir_vct=np.ascontiguousarray(v[8], np.int32))
    return args, np.array([float(x) for x in o[:6]], np.float32), np.asarray(o[6], np.float64)


# initGrid's scalars in f32 (the reference's f32 entry).
# (t, alpha) for which 20*alpha*sqrt(t) rounds to a power of two in f32, so
# that s0/dx = numX/(20*alpha*sqrt(t)) is an integer in f32 (LocVolCalib.fut:
# 15-20) but just below one in f64 from the same f32 inputs: there the f64
# index is one less than the reference's.
GRID32_EDGE = [(4.0, 0.2), (1.0, 0.2), (16.0, 0.1), (1.0, 0.4), (0.25, 0.4), (4.0, 0.1)]
GRID32_NX = (32, 64, 128, 256, 512)


def grid32_tolerance(nx):
    """Absolute tolerance between the FP64-grid form and the f32 restatement
    on strikes of main: the f32 restatement's own error from rounding its grid
    to f32 grows ~4x per doubling of numX (measured worst over 160 random
    parameter sets: 2e-7, 7.6e-7, 2.9e-6, 1.3e-5, 6.9e-5 for numX 32..512),
    while a grid index off by one moves the price by at least 1.9x this
    tolerance at every numX (4e-3 ... 2.2e-4 at the edge parameters)."""
    return 2e-6 + 1.1e-4 * (nx / 512.0) ** 2


@pytest.fixture(scope="session")
def gpu_ctx():
    import locvolcalib as lvc
    ctx = lvc.LocVolCalib(device=0)
    yield ctx
    ctx.free()
