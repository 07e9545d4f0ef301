"""The C ABI library: it loads, exports every function include/*.h
declares, and the parts that need no GPU (configuration, argument
validation, error reporting) behave like the generated Futhark API."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT
import locvolcalib as lvc
from locvolcalib import interestcalib as icl
from locvolcalib import optionpricing as op

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
# Each generated library declares the context and array API beside its own
# entry points; optionpricing.h belongs to liboptionpricing.so and
# interestcalib.h to libinterestcalib.so.
OWN_HEADERS = {"optionpricing.h", "interestcalib.h"}
SHARED = ("futhark_context.h", "futhark_arrays.h")


def header_functions(library="locvolcalib"):
    names = set()
    if library == "locvolcalib":
        hs = [h for h in HEADERS if os.path.basename(h) not in OWN_HEADERS]
    else:
        hs = [os.path.join(ROOT, "include", n) for n in SHARED + (library + ".h",)]
    for h in hs:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b((?:futhark|lvc|interestcalib)_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    L = lvc.lib()
    names = header_functions()
    assert len(names) >= 45
    for n in names:
        assert hasattr(L, n), f"{n} declared in the header but not exported"


def test_binding_covers_header():
    assert set(header_functions()) == set(lvc.exported_functions())


def test_only_public_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", lvc.library_path],
                         capture_output=True, text=True, check=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if l.strip()]
    assert syms and all(s.startswith(("futhark_", "lvc_server_")) for s in syms), syms


def test_config_and_tuning_params():
    L = lvc.lib()
    cfg = L.futhark_context_config_new()
    assert cfg
    n = L.futhark_get_tuning_param_count()
    names = [L.futhark_get_tuning_param_name(i).decode() for i in range(n)]
    assert "main.suff_intra_par_0" in names and "main.suff_intra_par_1" in names
    assert L.futhark_get_tuning_param_name(n) is None
    for name in names:
        assert L.futhark_context_config_set_tuning_param(cfg, name.encode(), 0) == 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"no.such.param", 1) != 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.precision", 7) != 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.seg_x", 3) != 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.seg_x", 8) == 0
    # column segments (one lane per segment) exist for the y sweep only
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.seg_x", 32) != 0
    for seg in (4, 8, 16, 32, 64):
        assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.seg_y", seg) == 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.seg_y", 128) != 0
    for share in (1, 4, 16):
        assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.xshare", share) == 0
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.xshare", 17) != 0
    L.futhark_context_config_set_device(cfg, b"hip:0")
    L.futhark_context_config_set_cache_file(cfg, b"/tmp/lvc-tune-cache.txt")
    L.futhark_context_config_free(cfg)


def _no_device_context():
    L = lvc.lib()
    cfg = L.futhark_context_config_new()
    ctx = L.futhark_context_new(cfg)
    err = lvc._take_string(L.futhark_context_get_error(ctx))
    return L, cfg, ctx, err


def test_context_without_device_reports_error():
    L, cfg, ctx, err = _no_device_context()
    try:
        if err is None:
            pytest.skip("a GPU is present; this checks the no-device path")
        assert "device" in err.lower()
        out = ctypes.c_void_p()
        rc = L.futhark_entry_main(ctx, ctypes.byref(out), 4, 32, 32, 8,
                                  0.03, 5.0, 0.2, 0.6, 0.5)
        assert rc != 0 and not out.value
        msg = lvc._take_string(L.futhark_context_get_error(ctx))
        assert msg and "device" in msg.lower()
        assert L.futhark_context_sync(ctx) != 0
        with pytest.raises(lvc.LocVolCalibError):
            lvc.LocVolCalib()
    finally:
        L.futhark_context_free(ctx)
        L.futhark_context_config_free(cfg)


@pytest.mark.parametrize("args,fragment", [
    ((4, 2, 32, 8), "at least 3"),       # LocVolCalib.fut:37 needs n >= 3
    ((4, 32, 1, 8), "at least 3"),
    ((4, 32, 32, 1), "numT"),
    ((-1, 32, 32, 8), "non-negative"),   # iota of a negative count (:247)
    ((4, 2048, 32, 8), "at most"),
])
def test_argument_validation(args, fragment):
    """Checked before any device work, so it holds with or without a GPU."""
    L, cfg, ctx, _ = _no_device_context()
    try:
        out = ctypes.c_void_p()
        rc = L.futhark_entry_main(ctx, ctypes.byref(out), *args, 0.03, 5.0, 0.2, 0.6, 0.5)
        assert rc != 0
        msg = lvc._take_string(L.futhark_context_get_error(ctx))
        assert fragment in msg
        out64 = ctypes.c_void_p()
        rc = L.futhark_entry_main_f64(ctx, ctypes.byref(out64), *args, 0.03, 5.0, 0.2, 0.6, 0.5)
        assert rc != 0
        assert fragment in lvc._take_string(L.futhark_context_get_error(ctx))
    finally:
        L.futhark_context_free(ctx)
        L.futhark_context_config_free(cfg)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(lvc, "_lib", None)
    monkeypatch.setattr(lvc, "library_path", str(tmp_path / "missing.so"))
    with pytest.raises(lvc.LocVolCalibError, match="missing"):
        lvc.lib()


# ---- liboptionpricing.so -------------------------------------------------

def test_optionpricing_exports_every_declared_function():
    L = op.lib()
    names = header_functions("optionpricing")
    assert "futhark_entry_main" in names and "futhark_new_f32_3d" in names
    for n in names:
        assert hasattr(L, n), f"{n} declared in the header but not exported"


def test_optionpricing_binding_covers_header():
    assert set(header_functions("optionpricing")) == set(op.exported_functions())


def test_optionpricing_only_public_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", op.library_path],
                         capture_output=True, text=True, check=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if l.strip()]
    assert syms and all(s.startswith("futhark_") for s in syms), syms


def test_optionpricing_has_no_tuning_params():
    L = op.lib()
    assert L.futhark_get_tuning_param_count() == 0
    assert L.futhark_get_tuning_param_name(0) is None
    cfg = L.futhark_context_config_new()
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.precision", 1) != 0
    L.futhark_context_config_free(cfg)


def test_optionpricing_without_device_reports_error():
    L = op.lib()
    cfg = L.futhark_context_config_new()
    ctx = L.futhark_context_new(cfg)
    try:
        err = lvc._take_string(L.futhark_context_get_error(ctx))
        if err is None:
            pytest.skip("a GPU is present; this checks the no-device path")
        out = ctypes.c_void_p()
        rc = L.futhark_entry_main(ctx, ctypes.byref(out), 1, 8, *([None] * 9))
        assert rc != 0 and not out.value
        assert "device" in lvc._take_string(L.futhark_context_get_error(ctx)).lower()
        with pytest.raises(op.OptionPricingError):
            op.OptionPricing()
    finally:
        L.futhark_context_free(ctx)
        L.futhark_context_config_free(cfg)


# ---- libinterestcalib.so ---------------------------------------------------

def test_interestcalib_exports_every_declared_function():
    L = icl.lib()
    names = header_functions("interestcalib")
    assert {"futhark_entry_main", "interestcalib_trace", "interestcalib_eval",
            "futhark_new_f32_2d", "futhark_new_i32_1d"} <= set(names)
    for n in names:
        assert hasattr(L, n), f"{n} declared in the header but not exported"


def test_interestcalib_binding_covers_header():
    assert set(header_functions("interestcalib")) == set(icl.exported_functions())


def test_interestcalib_only_public_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", icl.library_path],
                         capture_output=True, text=True, check=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if l.strip()]
    assert syms and all(s.startswith(("futhark_", "interestcalib_")) for s in syms), syms


def test_interestcalib_has_no_tuning_params():
    L = icl.lib()
    assert L.futhark_get_tuning_param_count() == 0
    cfg = L.futhark_context_config_new()
    assert L.futhark_context_config_set_tuning_param(cfg, b"lvc.precision", 1) != 0
    L.futhark_context_config_free(cfg)


def test_interestcalib_without_device_reports_error():
    L = icl.lib()
    cfg = L.futhark_context_config_new()
    ctx = L.futhark_context_new(cfg)
    try:
        err = lvc._take_string(L.futhark_context_get_error(ctx))
        if err is None:
            pytest.skip("a GPU is present; this checks the no-device path")
        outs = [ctypes.c_float() for _ in range(6)]
        arr = ctypes.c_void_p()
        rc = L.futhark_entry_main(ctx, *[ctypes.byref(o) for o in outs], ctypes.byref(arr), 8,
                                  4, 0, None, 0, None, None, 0, None)
        assert rc != 0 and not arr.value
        assert "device" in lvc._take_string(L.futhark_context_get_error(ctx)).lower()
        with pytest.raises(icl.InterestCalibError):
            icl.InterestCalib()
    finally:
        L.futhark_context_free(ctx)
        L.futhark_context_config_free(cfg)
