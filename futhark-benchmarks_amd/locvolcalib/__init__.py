"""Python binding of the MI355X LocVolCalib library (``liblocvolcalib.so``).

The library exports the C API that ``futhark <backend> --library`` generates
for ``finpar/LocVolCalib.fut`` (declared in ``include/locvolcalib.h``).  This
module is a thin ctypes layer over that C ABI, shaped like the Python class
that ``futhark pyopencl --library`` generates for a program: a context object
whose methods are the entry points (``main``, ``main_f64``, ``main_batch``)
and return numpy arrays.  Errors raise :class:`LocVolCalibError` with the
message of ``futhark_context_get_error``.

There is no CPU path: every computation runs the HIP kernels on the GPU, and
a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from typing import Optional, Sequence

import numpy as np

from . import futhark_values  # noqa: F401  (re-exported value I/O)

__all__ = ["LocVolCalib", "Server", "LocVolCalibError", "lib", "library_path",
           "PRECISION_FP32", "PRECISION_FP32_REFINE", "PRECISION_FP64",
           "exported_functions"]

PRECISION_FP32 = 0
PRECISION_FP32_REFINE = 1
PRECISION_FP64 = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
# LVC_LIBRARY names another build of the same sources (the tests' sanitizer
# build of the host code, _build/asan/liblocvolcalib.so)
library_path = os.environ.get("LVC_LIBRARY") or os.path.join(_HERE, "liblocvolcalib.so")
_lib: Optional[ctypes.CDLL] = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class LocVolCalibError(RuntimeError):
    """A call of the C API returned an error."""


class _Request(ctypes.Structure):
    _fields_ = [("numX", _i32), ("numY", _i32), ("numT", _i32),
                ("s0", _f64), ("strike", _f64), ("t", _f64),
                ("alpha", _f64), ("nu", _f64), ("beta", _f64)]


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
    "futhark_entry_main": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _i32, _i32, _i32, _i32,
                                          _f32, _f32, _f32, _f32, _f32]),
    "futhark_entry_main_f64": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _i32, _i32, _i32, _i32,
                                              _f64, _f64, _f64, _f64, _f64]),
    "futhark_entry_main_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)] + [_vp] * 9 + [_i32]),
    "futhark_entry_main_noalloc": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32,
                                                  _f32, _f32, _f32, _f32, _f32]),
    "futhark_new_f32_2d": (_vp, [_vp, _vp, _i64, _i64]),
    "futhark_new_raw_f32_2d": (_vp, [_vp, _vp, _i64, _i64]),
    "futhark_free_f32_2d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_f32_2d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_values_raw_f32_2d": (_vp, [_vp, _vp]),
    "futhark_shape_f32_2d": (ctypes.POINTER(_i64), [_vp, _vp]),
    "futhark_new_i32_2d": (_vp, [_vp, _vp, _i64, _i64]),
    "futhark_new_raw_i32_2d": (_vp, [_vp, _vp, _i64, _i64]),
    "futhark_free_i32_2d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_i32_2d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_values_raw_i32_2d": (_vp, [_vp, _vp]),
    "futhark_shape_i32_2d": (ctypes.POINTER(_i64), [_vp, _vp]),
    "futhark_new_f32_3d": (_vp, [_vp, _vp, _i64, _i64, _i64]),
    "futhark_new_raw_f32_3d": (_vp, [_vp, _vp, _i64, _i64, _i64]),
    "futhark_free_f32_3d": (ctypes.c_int, [_vp, _vp]),
    "futhark_values_f32_3d": (ctypes.c_int, [_vp, _vp, _vp]),
    "futhark_values_raw_f32_3d": (_vp, [_vp, _vp]),
    "futhark_shape_f32_3d": (ctypes.POINTER(_i64), [_vp, _vp]),
    "futhark_entry_tridagNested": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)] + [_vp] * 4),
    "futhark_entry_tridagNestedConst": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)] + [_vp] * 4),
    "futhark_entry_tridagNestedSeq": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)] + [_vp] * 4),
    "futhark_entry_tridagNestedSeqConst": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)] + [_vp] * 4),
    "lvc_server_new": (_vp, [_vp, _i32, _i64, _i32]),
    "lvc_server_free": (None, [_vp]),
    "lvc_server_submit": (_i64, [_vp, ctypes.POINTER(_Request)]),
    "lvc_server_step": (_i32, [_vp]),
    "lvc_server_poll": (_i32, [_vp, _i64, ctypes.POINTER(_f64)]),
    "lvc_server_drain": (_i32, [_vp]),
    "lvc_server_stats": (None, [_vp] + [ctypes.POINTER(_i64)] * 4),
    "lvc_server_device_bytes": (_i64, [_vp]),
}


def exported_functions() -> list:
    """Names of every function the binding declares (the C ABI)."""
    return sorted(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise LocVolCalibError(
                f"{library_path} is missing: build it with `make -C futhark-benchmarks_amd` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(library_path)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_libc = ctypes.CDLL(None)
_libc.free.argtypes = [_vp]
_libc.free.restype = None


def _take_string(ptr) -> Optional[str]:
    if not ptr:
        return None
    s = ctypes.string_at(ptr).decode("utf-8", "replace")
    _libc.free(ptr)
    return s


class LocVolCalib:
    """A configuration + context on one GPU (``futhark_context_new``).

    ``precision`` selects the arithmetic of :meth:`main` (the reference's f32
    entry): ``PRECISION_FP32``, ``PRECISION_FP32_REFINE`` (default) or
    ``PRECISION_FP64``.  ``tuning`` maps tuning-parameter names to values.
    """

    def __init__(self, device: int = 0, precision: Optional[int] = None,
                 profiling: bool = False, cache_file: Optional[str] = None,
                 tuning: Optional[dict] = None, logging: bool = False,
                 debugging: bool = False):
        L = lib()
        self._L = L
        self._cfg = L.futhark_context_config_new()
        if not self._cfg:
            raise LocVolCalibError("futhark_context_config_new failed")
        self._ctx = None
        self._servers = weakref.WeakSet()
        L.futhark_context_config_set_device(self._cfg, str(device).encode())
        L.futhark_context_config_set_profiling(self._cfg, int(profiling))
        L.futhark_context_config_set_logging(self._cfg, int(logging))
        L.futhark_context_config_set_debugging(self._cfg, int(debugging))
        if cache_file:
            L.futhark_context_config_set_cache_file(self._cfg, cache_file.encode())
        params = dict(tuning or {})
        if precision is not None:
            params["lvc.precision"] = precision
        for k, v in params.items():
            if L.futhark_context_config_set_tuning_param(self._cfg, k.encode(), int(v)):
                raise LocVolCalibError(f"unknown tuning parameter or value: {k}={v}")
        self._ctx = L.futhark_context_new(self._cfg)
        if not self._ctx:
            raise LocVolCalibError("futhark_context_new failed")
        err = self._error()
        if err:
            self.free()
            raise LocVolCalibError(err)

    # -- lifecycle --------------------------------------------------------
    def free(self) -> None:
        # servers live on the context: release them first
        for srv in list(getattr(self, "_servers", ())):
            srv.free()
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
            raise LocVolCalibError(f"{what}: {self._error() or 'unknown error'}")

    @property
    def handle(self) -> int:
        return self._ctx

    def sync(self) -> None:
        self._check(self._L.futhark_context_sync(self._ctx), "futhark_context_sync")

    def report(self) -> str:
        return _take_string(self._L.futhark_context_report(self._ctx)) or "{}"

    def price_index(self) -> Optional[tuple]:
        """(myYindex, myXindex) that the last entry call's price gather read
        from the device (LocVolCalib.fut:20, :236); needs ``debugging``."""
        import json
        idx = json.loads(self.report()).get("price_index")
        return tuple(idx) if idx is not None else None

    def pause_profiling(self) -> None:
        self._L.futhark_context_pause_profiling(self._ctx)

    def unpause_profiling(self) -> None:
        self._L.futhark_context_unpause_profiling(self._ctx)

    def clear_caches(self) -> None:
        self._check(self._L.futhark_context_clear_caches(self._ctx), "clear_caches")

    # -- arrays -----------------------------------------------------------
    def _new(self, kind: str, arr: np.ndarray) -> int:
        dt = {"f32": np.float32, "f64": np.float64, "i32": np.int32}[kind]
        a = np.ascontiguousarray(arr, dtype=dt)
        h = getattr(self._L, f"futhark_new_{kind}_1d")(self._ctx, a.ctypes.data if a.size else None, a.size)
        if not h:
            raise LocVolCalibError(f"futhark_new_{kind}_1d: {self._error()}")
        return h

    def _values(self, kind: str, h: int) -> np.ndarray:
        dt = {"f32": np.float32, "f64": np.float64, "i32": np.int32}[kind]
        n = getattr(self._L, f"futhark_shape_{kind}_1d")(self._ctx, h)[0]
        out = np.empty(n, dtype=dt)
        if n:
            self._check(getattr(self._L, f"futhark_values_{kind}_1d")(self._ctx, h, out.ctypes.data),
                        f"futhark_values_{kind}_1d")
        return out

    def _free(self, kind: str, h: int) -> None:
        getattr(self._L, f"futhark_free_{kind}_1d")(self._ctx, h)

    # -- entry points -----------------------------------------------------
    def main_raw(self, outer, numX, numY, numT, s0, t, alpha, nu, beta) -> int:
        """Enqueue ``futhark_entry_main``; returns the f32 array handle
        (device-resident, asynchronous).  Free it with :meth:`free_f32`."""
        out = _vp()
        self._check(self._L.futhark_entry_main(self._ctx, ctypes.byref(out), outer, numX, numY,
                                               numT, s0, t, alpha, nu, beta),
                    "futhark_entry_main")
        return out.value

    def free_f32(self, h: int) -> None:
        self._free("f32", h)

    def free_f64(self, h: int) -> None:
        self._free("f64", h)

    def values_f32(self, h: int) -> np.ndarray:
        """Copy of a device f32 array handle's values (the handle stays)."""
        return self._values("f32", h)

    def values_f64(self, h: int) -> np.ndarray:
        return self._values("f64", h)

    def main(self, outer, numX, numY, numT, s0, t, alpha, nu, beta) -> np.ndarray:
        """LocVolCalib.fut:238-252: ``outer`` prices (float32)."""
        h = self.main_raw(outer, numX, numY, numT, s0, t, alpha, nu, beta)
        try:
            return self._values("f32", h)
        finally:
            self._free("f32", h)

    def main_noalloc_raw(self, out_ptr: int, outer, numX, numY, numT, s0, t, alpha, nu,
                         beta) -> None:
        """Enqueue ``futhark_entry_main_noalloc``: prices into the device
        array at ``out_ptr`` (``outer`` floats), no device allocation."""
        self._check(self._L.futhark_entry_main_noalloc(self._ctx, out_ptr, outer, numX, numY,
                                                       numT, s0, t, alpha, nu, beta),
                    "futhark_entry_main_noalloc")

    def main_noalloc(self, outer, numX, numY, numT, s0, t, alpha, nu, beta) -> np.ndarray:
        """``main`` through the allocation-free entry: the output is a
        library array made here beforehand, its device memory passed raw."""
        h = self._new("f32", np.zeros(max(int(outer), 1), np.float32))
        try:
            ptr = self._L.futhark_values_raw_f32_1d(self._ctx, h)
            self.main_noalloc_raw(ptr, outer, numX, numY, numT, s0, t, alpha, nu, beta)
            return self._values("f32", h)[:max(int(outer), 0)]
        finally:
            self._free("f32", h)

    def main_f64_raw(self, outer, numX, numY, numT, s0, t, alpha, nu, beta) -> int:
        out = _vp()
        self._check(self._L.futhark_entry_main_f64(self._ctx, ctypes.byref(out), outer, numX, numY,
                                                   numT, s0, t, alpha, nu, beta),
                    "futhark_entry_main_f64")
        return out.value

    def main_f64(self, outer, numX, numY, numT, s0, t, alpha, nu, beta) -> np.ndarray:
        """The FP64 form of ``main`` (float64 prices)."""
        h = self.main_f64_raw(outer, numX, numY, numT, s0, t, alpha, nu, beta)
        try:
            return self._values("f64", h)
        finally:
            self._free("f64", h)

    def upload_f32_2d(self, v) -> int:
        """A device [n][m] f32 array (handle) holding v; free with free_f32_2d."""
        v = np.ascontiguousarray(v, dtype=np.float32)
        h = self._L.futhark_new_f32_2d(self._ctx, v.ctypes.data, v.shape[0], v.shape[1])
        if not h:
            self._check(1, "futhark_new_f32_2d")
        return h

    def free_f32_2d(self, h: int) -> None:
        self._L.futhark_free_f32_2d(self._ctx, h)

    def tridag_nested_raw(self, ha: int, hb: int, hc: int, hy: int, seq: bool = False,
                          const: bool = False) -> int:
        """Enqueue a tridagNested* entry on device arrays; returns the result
        handle (asynchronous)."""
        name = "futhark_entry_tridagNested" + ("Seq" if seq else "") + ("Const" if const else "")
        out = _vp()
        self._check(getattr(self._L, name)(self._ctx, ctypes.byref(out), ha, hb, hc, hy), name)
        return out.value

    def tridag_nested(self, a, b, c, y, seq: bool = False, const: bool = False) -> np.ndarray:
        """misc/ocean-sim/tridiag-test.fut:14-26: solves the n tridiagonal
        systems in the rows of the [n][m] f32 arrays a, b, c, y (entry
        tridagNested, or ...Seq for Thomas, ...Const for m = 115)."""
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (a, b, c, y)]
        if any(v.ndim != 2 for v in arrs):
            raise ValueError("tridag_nested takes [n][m] arrays")
        name = "futhark_entry_tridagNested" + ("Seq" if seq else "") + ("Const" if const else "")
        L = self._L
        handles = []
        try:
            for v in arrs:
                h = L.futhark_new_f32_2d(self._ctx, v.ctypes.data, v.shape[0], v.shape[1])
                if not h:
                    self._check(1, "futhark_new_f32_2d")
                handles.append(h)
            out = _vp()
            self._check(getattr(L, name)(self._ctx, ctypes.byref(out), *handles), name)
            try:
                shape = L.futhark_shape_f32_2d(self._ctx, out)
                res = np.empty((shape[0], shape[1]), np.float32)
                if res.size:
                    self._check(L.futhark_values_f32_2d(self._ctx, out, res.ctypes.data),
                                "futhark_values_f32_2d")
                return res
            finally:
                L.futhark_free_f32_2d(self._ctx, out)
        finally:
            for h in handles:
                L.futhark_free_f32_2d(self._ctx, h)

    def main_batch(self, numX: Sequence[int], numY: Sequence[int], numT: Sequence[int],
                   s0, strike, t, alpha, nu, beta,
                   precision: int = PRECISION_FP32_REFINE) -> np.ndarray:
        """Per-problem sizes and parameters; float64 prices."""
        handles = []
        try:
            for kind, v in (("i32", numX), ("i32", numY), ("i32", numT), ("f64", s0),
                            ("f64", strike), ("f64", t), ("f64", alpha), ("f64", nu),
                            ("f64", beta)):
                handles.append((kind, self._new(kind, np.asarray(v))))
            out = _vp()
            self._check(self._L.futhark_entry_main_batch(self._ctx, ctypes.byref(out),
                                                         *[h for _, h in handles], int(precision)),
                        "futhark_entry_main_batch")
            try:
                return self._values("f64", out.value)
            finally:
                self._free("f64", out.value)
        finally:
            for kind, h in handles:
                self._free(kind, h)


class Server:
    """Continuous-batching server on a context (``lvc_server_*``)."""

    def __init__(self, ctx: LocVolCalib, max_slots: int = 256, mem_budget: int = 0,
                 precision: int = PRECISION_FP32_REFINE):
        self._ctx = ctx
        self._L = ctx._L
        self._s = None
        if not ctx.handle:
            raise LocVolCalibError("context already freed")
        self._s = self._L.lvc_server_new(ctx.handle, max_slots, mem_budget, precision)
        if not self._s:
            raise LocVolCalibError(f"lvc_server_new: {ctx._error()}")
        ctx._servers.add(self)

    def free(self) -> None:
        if getattr(self, "_s", None):
            self._L.lvc_server_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def submit(self, numX, numY, numT, s0, strike, t, alpha, nu, beta) -> int:
        req = _Request(numX, numY, numT, s0, strike, t, alpha, nu, beta)
        rid = self._L.lvc_server_submit(self._s, ctypes.byref(req))
        if rid < 0:
            raise LocVolCalibError(f"lvc_server_submit: {self._ctx._error()}")
        return rid

    def step(self) -> int:
        n = self._L.lvc_server_step(self._s)
        if n < 0:
            raise LocVolCalibError(f"lvc_server_step: {self._ctx._error()}")
        return n

    def poll(self, rid: int) -> Optional[float]:
        v = _f64()
        r = self._L.lvc_server_poll(self._s, rid, ctypes.byref(v))
        if r < 0:
            raise LocVolCalibError(f"unknown request id {rid}")
        return v.value if r == 1 else None

    def drain(self) -> None:
        if self._L.lvc_server_drain(self._s) != 0:
            raise LocVolCalibError(f"lvc_server_drain: {self._ctx._error()}")

    def device_bytes(self) -> int:
        """Device memory the server holds (slot tables + grid arena)."""
        return int(self._L.lvc_server_device_bytes(self._s))

    def stats(self) -> dict:
        vals = [_i64() for _ in range(4)]
        self._L.lvc_server_stats(self._s, *[ctypes.byref(v) for v in vals])
        return dict(zip(("queued", "resident", "completed", "resident_bytes"),
                        (v.value for v in vals)))
