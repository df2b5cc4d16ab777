"""ctypes binding of the C ABI (include/ceos_amd.h) for device-pointer calls.

Used by the benchmark and serving drivers to pass HBM-resident tensors (torch
``data_ptr()`` values) straight to the library. Every call checks the status code
and raises with ``ceos_last_error()``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

from . import LIB_PATH

_lib = None

_HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ceos_amd.h")


class CeosError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
        P = C.c_void_p
        sig = {
            "ceos_last_error": (C.c_char_p, []),
            "ceos_create": (i32, [i32, i32, C.POINTER(C.c_void_p)]),
            "ceos_destroy": (None, [vp]),
            "ceos_set_index_param": (i32, [vp, i32, i32, i32, i32, i32]),
            "ceos_set_query_param": (i32, [vp, i32, i32, i32]),
            "ceos_get_query_param": (i32, [vp, P, P, P]),
            "ceos_get_dims": (i32, [vp, P, P, P, P, P]),
            "ceos_build_coceos": (i32, [vp, P, i64, i32]),
            "ceos_build_coceos_device": (i32, [vp, P, i64, i32, vp]),
            "ceos_search_coceos": (i32, [vp, P, i32, i32, i32, P, P]),
            "ceos_search_coceos_device": (i32, [vp, P, i32, i32, i32, P, P, vp]),
            "ceos_search_coceos_device_noalloc": (i32, [vp, P, i32, i32, i32, P, P, vp]),
            "ceos_search_batched": (i32, [vp, P, i32, i32, P, P, P, P, P]),
            "ceos_search_batched_device": (i32, [vp, P, i32, i32, P, P, P, P, i32, i32, i32, P, P, vp]),
            "ceos_set_search_config": (i32, [vp, i32, i32]),
            "ceos_get_search_config": (i32, [vp, P, P, P]),
            "ceos_set_rerank_bounds": (i32, [vp, i32]),
            "ceos_get_rerank_bounds": (i32, [vp, P]),
            "ceos_get_stream": (vp, [vp]),
            "ceos_plan_create": (i32, [vp, i32, i32, i32, i32, i32, C.POINTER(C.c_void_p)]),
            "ceos_stream_plan_create": (i32, [vp, i32, i32, i32, i32, i32, C.POINTER(C.c_void_p)]),
            "ceos_plan_destroy": (None, [vp]),
            "ceos_plan_buffers": (i32, [vp, P, P, P, P, P, P]),
            "ceos_plan_run": (i32, [vp, i32, vp]),
            "ceos_plan_search": (i32, [vp, P, i32, i32, P, P, P, P, P]),
            "ceos_plan_config": (i32, [vp, P, P, P]),
            "ceos_plan_tuning": (i32, [vp, i32, P, P, P]),
            "ceos_set_kernel_timing": (i32, [vp, i32]),
            "ceos_get_kernel_timing": (i32, [vp, P, P, P]),
            "ceos_read_bandwidth": (i32, [P, i64, P, i64, i32, vp, P]),
            "ceos_stream_create": (i32, [i32, C.POINTER(C.c_void_p)]),
            "ceos_stream_destroy": (None, [vp]),
            "ceos_stream_set_index_param": (i32, [vp, i32, i32, i32, i32, i32]),
            "ceos_stream_build": (i32, [vp, P, i64, i32]),
            "ceos_stream_update": (i32, [vp, P, i64, i32, i32]),
            "ceos_stream_get_build_times": (i32, [vp, P]),
            "ceos_get_build_times": (i32, [vp, P]),
            "ceos_set_build_budget": (i32, [vp, i64]),
            "ceos_get_build_rescans": (i32, [vp, P, P, P]),
            "ceos_stream_set_build_budget": (i32, [vp, i64]),
            "ceos_get_search_times": (i32, [vp, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise CeosError(f"ceos error {rc}: {lib().ceos_last_error().decode()}")


def header_symbols(path: str = _HEADER):
    """Function names declared in include/ceos_amd.h."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(ceos_\w+)\s*\(", text, re.M)))


class DeviceIndex:
    """Thin owner of a `ceos_index*` for device-pointer calls (torch tensors)."""

    def __init__(self, n: int, d: int):
        h = C.c_void_p()
        check(lib().ceos_create(n, d, C.byref(h)))
        self.h = h
        self.d = d

    def __del__(self):
        if getattr(self, "h", None):
            lib().ceos_destroy(self.h)
            self.h = None

    def set_index_param(self, D, R, m, threads=-1, seed=42):
        check(lib().ceos_set_index_param(self.h, D, R, m, threads, seed))

    def set_query_param(self, s, mp, c):
        check(lib().ceos_set_query_param(self.h, s, mp, c))

    def build_device(self, X_ptr: int, n: int, stream: int = 0):
        check(lib().ceos_build_coceos_device(self.h, X_ptr, n, self.d, stream or None))

    def search_device(self, Q_ptr, nq, k, ids_ptr, ips_ptr, stream: int = 0):
        check(lib().ceos_search_coceos_device(self.h, Q_ptr, nq, self.d, k, ids_ptr, ips_ptr, stream or None))

    def search_batched_device(self, Q_ptr, nq, k_ptr, s_ptr, c_ptr, off_ptr, k_max, s_max, c_max, ids_ptr,
                              ips_ptr, stream: int = 0):
        check(lib().ceos_search_batched_device(self.h, Q_ptr, nq, self.d, k_ptr, s_ptr, c_ptr, off_ptr, k_max,
                                               s_max, c_max, ids_ptr, ips_ptr, stream or None))

    def set_search_config(self, wg_per_cu=0, table_log2=0):
        check(lib().ceos_set_search_config(self.h, wg_per_cu, table_log2))

    def search_config(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().ceos_get_search_config(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_rerank_bounds(self, on: bool):
        check(lib().ceos_set_rerank_bounds(self.h, 1 if on else 0))

    def rerank_bounds(self) -> bool:
        v = C.c_int()
        check(lib().ceos_get_rerank_bounds(self.h, C.byref(v)))
        return bool(v.value)

    def stream(self) -> int:
        return lib().ceos_get_stream(self.h) or 0
