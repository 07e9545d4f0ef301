"""Python binding of the MI355X OptionPricing library (``liboptionpricing.so``).

The library exports the C API that ``futhark <backend> --library`` generates
for ``finpar/OptionPricing.fut`` (declared in ``include/optionpricing.h``):
the same context and array functions as ``liblocvolcalib.so`` and this
program's ``futhark_entry_main``.  :class:`OptionPricing` is a context whose
:meth:`~OptionPricing.main` takes the reference's eleven arguments
(OptionPricing.fut:295-306) as numpy arrays and returns the price per market
model.  Errors raise :class:`OptionPricingError` with the message of
``futhark_context_get_error``.  There is no CPU path: a missing library or
device raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

from . import LocVolCalibError, _take_string

__all__ = ["OptionPricing", "OptionPricingError", "lib", "library_path", "exported_functions",
           "main_arguments"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "liboptionpricing.so")
_lib: Optional[ctypes.CDLL] = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class OptionPricingError(LocVolCalibError):
    """A call of the OptionPricing C API returned an error."""


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
    "futhark_entry_main": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _i32, _i32] + [_vp] * 9),
}

# main's array arguments (OptionPricing.fut:297-306): (name, type, rank)
ARRAYS = (("dir_vs", "i32", 2), ("md_cs", "f32", 3), ("md_vols", "f32", 3),
          ("md_drifts", "f32", 3), ("md_sts", "f32", 2), ("md_detvals", "f32", 2),
          ("md_discts", "f32", 2), ("bb_inds", "i32", 2), ("bb_data", "f32", 2))


def exported_functions() -> list:
    """Names of every function the binding declares (the C ABI)."""
    return sorted(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise OptionPricingError(
                f"{library_path} is missing: build it with `make -C futhark-benchmarks_amd` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(library_path)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def main_arguments(values: Sequence) -> tuple:
    """The eleven arguments of ``main`` (OptionPricing.fut:295-306) from a
    dataset's values: contract_number and num_mc_it (i32), then the nine
    arrays with the element types the reference declares."""
    if len(values) != 11:
        raise ValueError(f"main takes 11 values, the dataset has {len(values)}")
    arrays = []
    for v, (name, ty, rank) in zip(values[2:], ARRAYS):
        a = np.ascontiguousarray(v, dtype=np.int32 if ty == "i32" else np.float32)
        if a.ndim != rank:
            raise ValueError(f"{name} has rank {a.ndim}, main takes rank {rank}")
        arrays.append(a)
    return (int(values[0]), int(values[1])) + tuple(arrays)


class OptionPricing:
    """A configuration + context on one GPU (``futhark_context_new``)."""

    def __init__(self, device: int = 0, profiling: bool = False, logging: bool = False):
        L = lib()
        self._L = L
        self._ctx = None
        self._cfg = L.futhark_context_config_new()
        if not self._cfg:
            raise OptionPricingError("futhark_context_config_new failed")
        L.futhark_context_config_set_device(self._cfg, str(device).encode())
        L.futhark_context_config_set_profiling(self._cfg, int(profiling))
        L.futhark_context_config_set_logging(self._cfg, int(logging))
        self._ctx = L.futhark_context_new(self._cfg)
        if not self._ctx:
            raise OptionPricingError("futhark_context_new failed")
        err = self._error()
        if err:
            self.free()
            raise OptionPricingError(err)

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
            raise OptionPricingError(f"{what}: {self._error() or 'unknown error'}")

    def sync(self) -> None:
        self._check(self._L.futhark_context_sync(self._ctx), "futhark_context_sync")

    def report(self) -> str:
        return _take_string(self._L.futhark_context_report(self._ctx)) or "{}"

    # -- arrays -------------------------------------------------------------
    def upload(self, v: np.ndarray, ty: str) -> int:
        """A device array handle holding v (``i32`` or ``f32``, rank 2 or 3);
        free it with :meth:`free_array`."""
        v = np.ascontiguousarray(v, dtype=np.int32 if ty == "i32" else np.float32)
        name = f"{ty}_{v.ndim}d"
        h = getattr(self._L, f"futhark_new_{name}")(self._ctx, v.ctypes.data, *v.shape)
        if not h:
            self._check(1, f"futhark_new_{name}")
        return h

    def free_array(self, h: int, ty: str, rank: int) -> None:
        getattr(self._L, f"futhark_free_{ty}_{rank}d")(self._ctx, h)

    def upload_arguments(self, arrays: Sequence[np.ndarray]) -> List[int]:
        """Handles of main's nine arrays (in :data:`ARRAYS` order)."""
        hs: List[int] = []
        try:
            for v, (name, ty, rank) in zip(arrays, ARRAYS):
                if np.ndim(v) != rank:
                    raise ValueError(f"{name} must have rank {rank}")
                hs.append(self.upload(v, ty))
        except Exception:
            self.free_arguments(hs)
            raise
        return hs

    def free_arguments(self, handles: Sequence[int]) -> None:
        for h, (_, ty, rank) in zip(handles, ARRAYS):
            self.free_array(h, ty, rank)

    def values_f32(self, h: int) -> np.ndarray:
        n = self._L.futhark_shape_f32_1d(self._ctx, h)[0]
        out = np.empty(n, np.float32)
        if n:
            self._check(self._L.futhark_values_f32_1d(self._ctx, h, out.ctypes.data),
                        "futhark_values_f32_1d")
        return out

    def free_f32(self, h: int) -> None:
        self._L.futhark_free_f32_1d(self._ctx, h)

    # -- the entry point ----------------------------------------------------
    def main_raw(self, contract_number: int, num_mc_it: int, handles: Sequence[int]) -> int:
        """Enqueue ``main`` on device arrays; returns the result handle
        (asynchronous)."""
        out = _vp()
        self._check(self._L.futhark_entry_main(self._ctx, ctypes.byref(out), contract_number,
                                               num_mc_it, *handles), "futhark_entry_main")
        return out.value

    def main(self, contract_number: int, num_mc_it: int, *arrays: np.ndarray) -> np.ndarray:
        """OptionPricing.fut:295-322: the price of the contract under each
        market model (float32), averaged over num_mc_it Sobol paths."""
        if len(arrays) != len(ARRAYS):
            raise ValueError(f"main takes {len(ARRAYS)} arrays after the two scalars")
        hs = self.upload_arguments(arrays)
        try:
            h = self.main_raw(contract_number, num_mc_it, hs)
            try:
                return self.values_f32(h)
            finally:
                self.free_f32(h)
        finally:
            self.free_arguments(hs)
