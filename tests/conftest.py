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


@pytest.fixture(scope="session")
def gpu_ctx():
    import locvolcalib as lvc
    ctx = lvc.LocVolCalib(device=0)
    yield ctx
    ctx.free()
