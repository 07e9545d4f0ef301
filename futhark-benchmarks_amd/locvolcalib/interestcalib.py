"""Python binding of the MI355X InterestCalib library (``libinterestcalib.so``).

The library exports the C API that ``futhark <backend> --library`` generates
for ``finpar/InterestCalib.fut`` (declared in ``include/interestcalib.h``):
the same context and array functions as ``liblocvolcalib.so`` and this
program's ``futhark_entry_main``, plus two calls with no counterpart in the
reference, ``interestcalib_trace`` (the calibration with its trajectory
recorded) and ``interestcalib_eval`` (genomes priced on the swaptions).
:class:`InterestCalib` is a context whose :meth:`~InterestCalib.main` takes
the reference's arguments (InterestCalib.fut:929-936) as numpy arrays.
Errors raise :class:`InterestCalibError` with the message of
``futhark_context_get_error``.  There is no CPU path: a missing library or
device raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from . import LocVolCalibError, _take_string

__all__ = ["InterestCalib", "InterestCalibError", "lib", "library_path", "exported_functions",
           "main_arguments"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# IC_LIBRARY names another build of the same sources (the tests' sanitizer
# build, futhark-benchmarks_amd/_build/asan/libinterestcalib.so)
library_path = os.environ.get("IC_LIBRARY") or os.path.join(_HERE, "libinterestcalib.so")
_lib: Optional[ctypes.CDLL] = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_fp = ctypes.POINTER(ctypes.c_float)


class InterestCalibError(LocVolCalibError):
    """A call of the InterestCalib C API returned an error."""


def _array_api(name: str, rank: int) -> dict:
    dims = [_i64] * rank
    return {
        f"futhark_new_{name}": (_vp, [_vp, _vp] + dims),
        f"futhark_new_raw_{name}": (_vp, [_vp, _vp] + dims),
        f"futhark_free_{name}": (ctypes.c_int, [_vp, _vp]),
        f"futhark_values_{name}": (ctypes.c_int, [_vp, _vp, _vp]),
        f"futhark_values_raw_{name}": (_vp, [_vp, _vp]),
        f"futhark_shape_{name}": (ctypes.POINTER(_i64), [_vp, _vp]),
    }


_SIGNATURES = {
    "futhark_context_config_new": (_vp, []),
    "futhark_context_config_free": (None, [_vp]),
    "futhark_context_config_set_device": (None, [_vp, ctypes.c_char_p]),
    "futhark_context_config_set_cache_file": (None, [_vp, ctypes.c_char_p]),
    "futhark_context_config_set_tuning_param": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    "futhark_context_config_set_debugging": (None, [_vp, ctypes.c_int]),
    "futhark_context_config_set_profiling": (None, [_vp, ctypes.c_int]),
    "futhark_context_config_set_logging": (None, [_vp, ctypes.c_int]),
    "futhark_get_tuning_param_count": (ctypes.c_int, []),
    "futhark_get_tuning_param_name": (ctypes.c_char_p, [ctypes.c_int]),
    "futhark_get_tuning_param_class": (ctypes.c_char_p, [ctypes.c_int]),
    "futhark_context_new": (_vp, [_vp]),
    "futhark_context_free": (None, [_vp]),
    "futhark_context_sync": (ctypes.c_int, [_vp]),
    "futhark_context_get_error": (_vp, [_vp]),
    "futhark_context_set_logging_file": (None, [_vp, _vp]),
    "futhark_context_report": (_vp, [_vp]),
    "futhark_context_pause_profiling": (None, [_vp]),
    "futhark_context_unpause_profiling": (None, [_vp]),
    "futhark_context_clear_caches": (ctypes.c_int, [_vp]),
    "futhark_new_f32_1d": (_vp, [_vp, _vp, _i64]),
    "futhark_new_raw_f32_1d": (_vp, [_vp, _vp, _i64]),
    "futhark_free_f32_1d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_f32_1d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_index_f32_1d": (ctypes.c_int, [_vp, _vp, _vp, _i64]),
    "futhark_values_raw_f32_1d": (_vp, [_vp, _vp]),
    "futhark_shape_f32_1d": (ctypes.POINTER(_i64), [_vp, _vp]),
    "futhark_new_f64_1d": (_vp, [_vp, _vp, _i64]),
    "futhark_new_raw_f64_1d": (_vp, [_vp, _vp, _i64]),
    "futhark_free_f64_1d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_f64_1d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_index_f64_1d": (ctypes.c_int, [_vp, _vp, _vp, _i64]),
    "futhark_values_raw_f64_1d": (_vp, [_vp, _vp]),
    "futhark_shape_f64_1d": (ctypes.POINTER(_i64), [_vp, _vp]),
    "futhark_new_i32_1d": (_vp, [_vp, _vp, _i64]),
    "futhark_free_i32_1d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_i32_1d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_shape_i32_1d": (ctypes.POINTER(_i64), [_vp, _vp]),
    **_array_api("f32_2d", 2),
    **_array_api("i32_2d", 2),
    **_array_api("f32_3d", 3),
    "futhark_entry_main": (ctypes.c_int, [_vp] + [_fp] * 6 + [ctypes.POINTER(_vp), _i32, _i32,
                                                                 _i32, _vp, _i32, _vp, _vp,
                                                                 _i32, _vp]),
    "interestcalib_trace": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp, _vp] + [_vp] * 7),
    "interestcalib_eval": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}


def exported_functions() -> list:
    """Names of every function the binding declares (the C ABI)."""
    return sorted(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise InterestCalibError(
                f"{library_path} is missing: build it with `make -C futhark-benchmarks_amd` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(library_path)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def main_arguments(values: Sequence) -> dict:
    """main's arguments (InterestCalib.fut:929-932) from a dataset's nine
    values: pop, mcmc_conv, (l), swaptions [n][4], (ll), hermCoefs,
    hermWeights, (lll), sobDirVct."""
    if len(values) != 9:
        raise ValueError(f"main takes 9 values, the dataset has {len(values)}")
    return dict(pop=int(values[0]), mcmc_conv=int(values[1]),
                swaptions=np.ascontiguousarray(values[3], np.float32),
                herm_x=np.ascontiguousarray(values[5], np.float32),
                herm_w=np.ascontiguousarray(values[6], np.float32),
                dir_vct=np.ascontiguousarray(values[8], np.int32))


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class InterestCalib:
    """A configuration + context on one GPU (``futhark_context_new``)."""

    def __init__(self, device: int = 0, profiling: bool = False, logging: bool = False):
        L = lib()
        self._L = L
        self._ctx = None
        self._cfg = L.futhark_context_config_new()
        if not self._cfg:
            raise InterestCalibError("futhark_context_config_new failed")
        L.futhark_context_config_set_device(self._cfg, str(device).encode())
        L.futhark_context_config_set_profiling(self._cfg, int(profiling))
        L.futhark_context_config_set_logging(self._cfg, int(logging))
        self._ctx = L.futhark_context_new(self._cfg)
        if not self._ctx:
            raise InterestCalibError("futhark_context_new failed")
        err = self._error()
        if err:
            self.free()
            raise InterestCalibError(err)

    def free(self) -> None:
        if getattr(self, "_ctx", None):
            self._L.futhark_context_free(self._ctx)
            self._ctx = None
        if getattr(self, "_cfg", None):
            self._L.futhark_context_config_free(self._cfg)
            self._cfg = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _error(self) -> Optional[str]:
        return _take_string(self._L.futhark_context_get_error(self._ctx))

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise InterestCalibError(f"{what}: {self._error() or 'unknown error'}")

    def sync(self) -> None:
        self._check(self._L.futhark_context_sync(self._ctx), "futhark_context_sync")

    def report(self) -> str:
        """``futhark_context_report``: with profiling on, the GPU time of
        each kind of launch since the last report (ic_eval, ic_step,
        ic_other), as JSON."""
        return _take_string(self._L.futhark_context_report(self._ctx)) or "{}"

    def pause_profiling(self) -> None:
        self._L.futhark_context_pause_profiling(self._ctx)

    def unpause_profiling(self) -> None:
        self._L.futhark_context_unpause_profiling(self._ctx)

    # -- arrays -------------------------------------------------------------
    def _inputs(self, swaptions, herm_x, herm_w, dir_vct=None):
        L, c = self._L, self._ctx
        sw = np.ascontiguousarray(swaptions, np.float32)
        if sw.ndim != 2:
            raise ValueError("swaptions must be [n][4]")
        hx = np.ascontiguousarray(herm_x, np.float32)
        hw = np.ascontiguousarray(herm_w, np.float32)
        hs = []
        try:
            hs.append(("f32_2d", L.futhark_new_f32_2d(c, _p(sw), *sw.shape)))
            hs.append(("f32_1d", L.futhark_new_f32_1d(c, _p(hx), hx.shape[0])))
            hs.append(("f32_1d", L.futhark_new_f32_1d(c, _p(hw), hw.shape[0])))
            if dir_vct is not None:
                dv = np.ascontiguousarray(dir_vct, np.int32)
                hs.append(("i32_1d", L.futhark_new_i32_1d(c, _p(dv), dv.shape[0])))
            if not all(h for _, h in hs):
                self._check(1, "futhark_new_*")
        except Exception:
            self._free_inputs(hs)
            raise
        return hs

    def _free_inputs(self, hs):
        for ty, h in hs:
            if h:
                getattr(self._L, f"futhark_free_{ty}")(self._ctx, h)

    # -- the entry points ---------------------------------------------------
    def upload_inputs(self, swaptions, herm_x, herm_w, dir_vct) -> list:
        """Device handles of main's four arrays, for :meth:`main_raw`; free
        them with :meth:`free_inputs`."""
        return self._inputs(swaptions, herm_x, herm_w, dir_vct)

    def free_inputs(self, handles) -> None:
        self._free_inputs(handles)

    def main_raw(self, pop: int, mcmc_conv: int, handles) -> tuple:
        """``futhark_entry_main`` on uploaded inputs: (the six scalars, the
        summary's device handle; free it with :meth:`free_summary`)."""
        outs = [ctypes.c_float() for _ in range(6)]
        arr = _vp()
        n = int(self._L.futhark_shape_f32_2d(self._ctx, handles[0][1])[0])
        m = int(self._L.futhark_shape_f32_1d(self._ctx, handles[1][1])[0])
        b = int(self._L.futhark_shape_i32_1d(self._ctx, handles[3][1])[0])
        self._check(self._L.futhark_entry_main(
            self._ctx, *[ctypes.byref(o) for o in outs], ctypes.byref(arr), int(pop),
            int(mcmc_conv), n, handles[0][1], m, handles[1][1], handles[2][1], b,
            handles[3][1]), "futhark_entry_main")
        return np.array([o.value for o in outs], np.float32), arr.value

    def free_summary(self, h) -> None:
        self._L.futhark_free_f32_2d(self._ctx, h)

    def main(self, pop: int, mcmc_conv: int, swaptions, herm_x, herm_w, dir_vct) -> tuple:
        """InterestCalib.fut:929-936: (the six f32 scalars a, b, sigma, nu,
        rho, log-likelihood of the winner; the summary [n][3])."""
        hs = self._inputs(swaptions, herm_x, herm_w, dir_vct)
        try:
            outs = [ctypes.c_float() for _ in range(6)]
            arr = _vp()
            n = int(np.shape(swaptions)[0])
            self._check(self._L.futhark_entry_main(
                self._ctx, *[ctypes.byref(o) for o in outs], ctypes.byref(arr), int(pop),
                int(mcmc_conv), n, hs[0][1], len(herm_x), hs[1][1], hs[2][1], len(dir_vct),
                hs[3][1]), "futhark_entry_main")
            try:
                shape = self._L.futhark_shape_f32_2d(self._ctx, arr.value)
                summ = np.empty((shape[0], shape[1]), np.float32)
                if summ.size:
                    self._check(self._L.futhark_values_f32_2d(self._ctx, arr.value, _p(summ)),
                                "futhark_values_f32_2d")
            finally:
                self._L.futhark_free_f32_2d(self._ctx, arr.value)
            return np.array([o.value for o in outs], np.float32), summ
        finally:
            self._free_inputs(hs)

    def trace(self, pop: int, mcmc_conv: int, swaptions, herm_x, herm_w, dir_vct) -> dict:
        """The calibration with its trajectory (``interestcalib_trace``):
        genomes [conv+1][pop][5], logliks [conv+1][pop], props [conv][pop][5],
        fb and newl [conv][pop], out [6], summary [n][3]."""
        conv, p = max(int(mcmc_conv), 0), max(int(pop), 0)
        n = int(np.shape(swaptions)[0])
        r = dict(genomes=np.empty((conv + 1, p, 5), np.float32),
                 logliks=np.empty((conv + 1, p), np.float32),
                 props=np.empty((conv, p, 5), np.float32), fb=np.empty((conv, p), np.float32),
                 newl=np.empty((conv, p), np.float32), out=np.empty(6, np.float32),
                 summary=np.empty((n, 3), np.float32))
        hs = self._inputs(swaptions, herm_x, herm_w, dir_vct)
        try:
            self._check(self._L.interestcalib_trace(
                self._ctx, int(pop), int(mcmc_conv), hs[0][1], hs[1][1], hs[2][1], hs[3][1],
                *[_p(r[k]) for k in ("genomes", "logliks", "props", "fb", "newl", "out",
                                     "summary")]), "interestcalib_trace")
        finally:
            self._free_inputs(hs)
        return r

    def eval_genomes(self, genomes, swaptions, herm_x, herm_w) -> tuple:
        """evalGenomeOnSwap of each genome [g][5] (a, b, rho, nu, sigma) on
        each swaption: (quotes [g][n], prices [g][n], log-likelihoods [g])."""
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1, 5)
        n = int(np.shape(swaptions)[0])
        q = np.empty((len(g), n), np.float32)
        pr = np.empty((len(g), n), np.float32)
        ll = np.empty(len(g), np.float32)
        hs = self._inputs(swaptions, herm_x, herm_w)
        try:
            self._check(self._L.interestcalib_eval(self._ctx, len(g), _p(g), hs[0][1], hs[1][1],
                                                   hs[2][1], _p(q), _p(pr), _p(ll)),
                        "interestcalib_eval")
        finally:
            self._free_inputs(hs)
        return q, pr, ll
