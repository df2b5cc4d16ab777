"""ctypes wrapper over the CPU oracle (libceos_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the reference answer. The product never imports it.

The classes mirror the reference's Python surface (python/python_wrapper.cpp:13-75):
`OracleCEOs` for class `CEOs`, `OracleStream` for class `streamCEOs`.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libceos_oracle.so")
_FFHT_PATH = os.path.join(_HERE, "_ref", "libffht.so")
# the same restatement built with the reference's FFHT, tsl::robin_map and an FP32 SIMD
# dot (see the Makefile): timed as the CPU baseline, never used as the parity checker
TIMING_PATH = os.path.join(_HERE, "_ref", "libceos_ref_timing.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

_lib = None
_timing_lib = None


def build() -> None:
    """Compile the oracle (and _ref/libffht.so when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(timing: bool = False):
    """The oracle library; timing=True: the timing build under _ref/ (None if absent)."""
    global _lib, _timing_lib
    if timing:
        if _timing_lib is None and os.path.exists(TIMING_PATH):
            _timing_lib = _load(TIMING_PATH)
        return _timing_lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


def _load(path):
    L = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
    sig = {
        "oracle_sign_bits": (None, [i32, i32, _u8p]),
        "oracle_fwht_f32": (None, [_f32p, i32]),
        "oracle_fwht_f64": (None, [_f64p, i32]),
        "oracle_fht_dim": (i32, [i32, i32]),
        "oracle_ceos_new": (vp, [i32, i32]),
        "oracle_ceos_free": (None, [vp]),
        "oracle_ceos_set_index_param": (None, [vp, i32, i32, i32, i32, i32]),
        "oracle_ceos_fht_dim": (i32, [vp]),
        "oracle_ceos_set_threads": (None, [vp, i32]),
        "oracle_ceos_build_coceos": (None, [vp, _f32p]),
        "oracle_ceos_get_buckets": (None, [vp, _i32p, _f32p, _i32p, _f32p]),
        "oracle_ceos_get_bits": (None, [vp, _u8p]),
        "oracle_ceos_load_index": (None, [vp, _f32p, _i32p, _f32p, _i32p, _f32p]),
        "oracle_ceos_project": (None, [vp, _f32p, i64, _f32p]),
        "oracle_ceos_set_sizes": (None, [vp, _i32p, _i32p]),
        "oracle_ceos_bucket_columns": (None, [vp, _f32p, i64, _i32p, i32, _i32p, _f32p, _i32p, _f32p]),
        "oracle_ceos_search_coceos": (i32, [vp, _f32p, i32, _i32p, _i32p, i32, _i32p, _i64p,
                                             _i32p, _f32p, _i32p, _f32p]),
        "oracle_ceos_build_est": (None, [vp, _f32p]),
        "oracle_ceos_search_est": (None, [vp, _f32p, i32, i32, i32, i32, _i32p, _f32p]),
        "oracle_exact_mips": (None, [_f32p, i64, i32, _f32p, i32, i32, i32, _i32p, _f32p]),
        "oracle_stream_new": (vp, [i32]),
        "oracle_stream_free": (None, [vp]),
        "oracle_stream_set_index_param": (None, [vp, i32, i32, i32, i32, i32]),
        "oracle_stream_build": (None, [vp, _f32p, i64]),
        "oracle_stream_update": (i32, [vp, _f32p, i64, i32]),
        "oracle_stream_get_buckets": (None, [vp, _i32p, _f32p, _i32p, _f32p, _i32p, _i32p]),
        "oracle_stream_size": (i32, [vp]),
        "oracle_column_mean": (None, [_f32p, i64, i32, _f32p]),
        "oracle_stream_set_centering": (None, [vp, i32, C.c_void_p]),
        "oracle_stream_get_center": (None, [vp, _f32p]),
        "oracle_stream_search": (i32, [vp, _f32p, i32, i32, i32, i32, i32, _i32p, _f32p]),
        "oracle_ceos_search_hash": (i32, [vp, _f32p, i32, i32, i32, i32, _i32p, _f32p, _i32p]),
        "oracle_stream_hash_search": (i32, [vp, _f32p, i32, i32, i32, i32, _i32p, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def sign_bits(nbits: int, seed: int) -> np.ndarray:
    out = np.zeros(nbits, np.uint8)
    lib().oracle_sign_bits(nbits, seed, out)
    return out


def fwht_f32(x: np.ndarray) -> np.ndarray:
    buf = _f32(x).copy()
    lib().oracle_fwht_f32(buf, int(np.log2(buf.size)))
    return buf


def fwht_f64(x: np.ndarray) -> np.ndarray:
    buf = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().oracle_fwht_f64(buf, int(np.log2(buf.size)))
    return buf


def fht_dim(n_proj: int, d: int) -> int:
    return lib().oracle_fht_dim(n_proj, d)


def ffht_ref():
    """The reference's own FFHT (src/fht.c) compiled under _ref/, or None."""
    if not os.path.exists(_FFHT_PATH):
        return None
    L = C.CDLL(_FFHT_PATH)
    L.fht_float.restype = C.c_int
    L.fht_float.argtypes = [_f32p, C.c_int]
    return L


def exact_mips(X, Q, k, threads=0):
    X, Q = _f32(X), _f32(Q)
    nq = Q.shape[0]
    ids = np.zeros((nq, k), np.int32)
    ips = np.zeros((nq, k), np.float32)
    lib().oracle_exact_mips(X, X.shape[0], X.shape[1], Q, nq, k, threads, ids, ips)
    return ids, ips


def column_mean(X):
    """FP32 column mean in the order of Eigen 3.4's colwise().mean() (src/streamCEOs.cpp:24-25)."""
    X = _f32(X)
    out = np.zeros(X.shape[1], np.float32)
    lib().oracle_column_mean(X, X.shape[0], X.shape[1], out)
    return out


class OracleCEOs:
    """Restatement of class CEOs (src/CEOs.h:18-114)."""

    def __init__(self, n_points: int, n_features: int, timing: bool = False):
        self._L = lib(timing)
        if self._L is None:
            raise FileNotFoundError(TIMING_PATH)
        self.n, self.d = n_points, n_features
        self._h = self._L.oracle_ceos_new(n_points, n_features)
        self.n_probed_vectors = 10
        self.n_probed_points = 10
        self.n_cand = 10
        self.top_m = 100

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_ceos_free(self._h)
            self._h = None

    def setIndexParam(self, n_proj, n_repeats=1, top_m=10, n_threads=-1, random_seed=-1):
        self.D, self.R, self.top_m = n_proj, n_repeats, top_m
        self.n_probed_points = top_m  # src/CEOs.h:69
        self._L.oracle_ceos_set_index_param(self._h, n_proj, n_repeats, top_m, n_threads, random_seed)
        self.F = self._L.oracle_ceos_fht_dim(self._h)

    def set_threads(self, n_threads):
        self._L.oracle_ceos_set_threads(self._h, n_threads)

    def build_coCEOs_Est(self, X):
        self._L.oracle_ceos_build_coceos(self._h, _f32(X))

    def buckets(self):
        DR, m = self.D * self.R, self.top_m
        cid = np.zeros((DR, m), np.int32); fid = np.zeros((DR, m), np.int32)
        cv = np.zeros((DR, m), np.float32); fv = np.zeros((DR, m), np.float32)
        self._L.oracle_ceos_get_buckets(self._h, cid, cv, fid, fv)
        return cid, cv, fid, fv

    def load_index(self, X, cid, cval, fid, fval):
        """Installs points and buckets built elsewhere (same layout as buckets())."""
        c = lambda a, t: np.ascontiguousarray(a, dtype=t)
        self._L.oracle_ceos_load_index(self._h, _f32(X), c(cid, np.int32), c(cval, np.float32),
                                     c(fid, np.int32), c(fval, np.float32))

    def set_sizes(self, csize, fsize):
        """Valid entries per close / far list after load_index (default: top_m each)."""
        self._L.oracle_ceos_set_sizes(self._h, np.ascontiguousarray(csize, np.int32),
                                    np.ascontiguousarray(fsize, np.int32))

    def bucket_columns(self, X, cols):
        """The build's close / far lists of the given columns over the rows X, projected in
        chunks (src/CEOs.cpp:343-432): (cid, cval, fid, fval), each [len(cols)][top_m]."""
        X = _f32(X)
        cols = np.ascontiguousarray(cols, np.int32)
        nc, m = len(cols), self.top_m
        cid = np.zeros((nc, m), np.int32); fid = np.zeros((nc, m), np.int32)
        cv = np.zeros((nc, m), np.float32); fv = np.zeros((nc, m), np.float32)
        self._L.oracle_ceos_bucket_columns(self._h, X, X.shape[0], cols, nc, cid, cv, fid, fv)
        return cid, cv, fid, fv

    def sign_bits(self):
        out = np.zeros(self.F * 3 * self.R, np.uint8)
        self._L.oracle_ceos_get_bits(self._h, out)
        return out

    def project(self, rows):
        rows = _f32(rows)
        out = np.zeros((rows.shape[0], self.D * self.R), np.float32)
        self._L.oracle_ceos_project(self._h, rows, rows.shape[0], out)
        return out

    def search_coCEOs_Est(self, Q, n_neighbors, verbose=False, *, ks=None, ss=None, cs=None,
                          stats=False):
        Q = _f32(Q)
        nq = Q.shape[0]
        k = np.full(nq, n_neighbors, np.int32) if ks is None else np.asarray(ks, np.int32)
        s = np.full(nq, self.n_probed_vectors, np.int32) if ss is None else np.asarray(ss, np.int32)
        c = np.full(nq, self.n_cand, np.int32) if cs is None else np.asarray(cs, np.int32)
        off = np.zeros(nq, np.int64)
        off[1:] = np.cumsum(k[:-1])
        tot = int(k.sum())
        ids = np.zeros(tot, np.int32)
        ips = np.zeros(tot, np.float32)
        distinct = np.zeros(nq, np.int32)
        margin = np.zeros(nq, np.float32)
        rc = self._L.oracle_ceos_search_coceos(self._h, Q, nq, k, s, self.n_probed_points, c, off, ids,
                                             ips, distinct, margin)
        if rc:
            raise ValueError("invalid query parameters")
        if ks is None:
            ids, ips = ids.reshape(nq, n_neighbors), ips.reshape(nq, n_neighbors)
        if stats:
            return ids, ips, distinct, margin
        return ids, ips

    centering = False  # src/CEOs.h:42
    center = None

    def build_CEOs_Hash(self, X, center=None):
        """CEOs-Hash keeps the coCEOs lists' ids (src/CEOs.cpp:709-902). With centering the
        stored points are X minus its column mean (or minus `center` when given); the
        reference's own expression there (src/CEOs.cpp:721-728, a row-wise mean) only
        type-checks when n equals d."""
        X = _f32(X)
        if self.centering:
            self.center = column_mean(X) if center is None else np.asarray(center, np.float32)
            X = np.ascontiguousarray(X - self.center[None, :], dtype=np.float32)
        self.build_coCEOs_Est(X)

    def search_CEOs_Hash(self, Q, n_neighbors, verbose=False, *, stats=False):
        """src/CEOs.cpp:904-1066: rerank the union of the probed lists' first ids."""
        Q = _f32(Q)
        nq = Q.shape[0]
        ids = np.zeros((nq, n_neighbors), np.int32)
        ips = np.zeros((nq, n_neighbors), np.float32)
        cnt = np.zeros(nq, np.int32)
        rc = self._L.oracle_ceos_search_hash(self._h, Q, nq, n_neighbors, self.n_probed_vectors,
                                           self.n_probed_points, ids, ips, cnt)
        if rc:
            raise ValueError("invalid query parameters")
        return (ids, ips, cnt) if stats else (ids, ips)

    def build(self, X):
        self._L.oracle_ceos_build_est(self._h, _f32(X))

    def search(self, Q, n_neighbors, verbose=False):
        Q = _f32(Q)
        nq = Q.shape[0]
        ids = np.zeros((nq, n_neighbors), np.int32)
        ips = np.zeros((nq, n_neighbors), np.float32)
        self._L.oracle_ceos_search_est(self._h, Q, nq, n_neighbors, self.n_probed_vectors, self.n_cand,
                                     ids, ips)
        return ids, ips


class OracleStream:
    """Restatement of class streamCEOs (src/streamCEOs.h:12-107)."""

    def __init__(self, n_features: int, timing: bool = False):
        self._L = lib(timing)
        if self._L is None:
            raise FileNotFoundError(TIMING_PATH)
        self.d = n_features
        self._h = self._L.oracle_stream_new(n_features)
        self.n_probed_vectors = 10
        self.n_probed_points = 10
        self.n_cand = 10

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_stream_free(self._h)
            self._h = None

    def setIndexParam(self, n_proj, n_repeats=1, top_points=1, n_threads=-1, random_seed=-1):
        self.D, self.R, self.top_m = n_proj, n_repeats, top_points
        self.n_probed_vectors = top_points  # src/streamCEOs.h:58
        self._L.oracle_stream_set_index_param(self._h, n_proj, n_repeats, top_points, n_threads,
                                            random_seed)

    def set_centering(self, on, center=None):
        """centering (src/streamCEOs.h:44); `center` replaces the computed mean at the next build."""
        c = None if center is None else np.ascontiguousarray(center, np.float32)
        self._given = c  # kept alive while the library may read it
        self._L.oracle_stream_set_centering(self._h, 1 if on else 0, None if c is None else c.ctypes.data)

    def center(self):
        out = np.zeros(self.d, np.float32)
        self._L.oracle_stream_get_center(self._h, out)
        return out

    def build(self, X):
        X = _f32(X)
        self._L.oracle_stream_build(self._h, X, X.shape[0])

    def update(self, new_X, n_delPoints=0):
        new_X = _f32(new_X)
        if self._L.oracle_stream_update(self._h, new_X, new_X.shape[0], n_delPoints):
            raise ValueError("invalid update")

    def size(self):
        return self._L.oracle_stream_size(self._h)

    def buckets(self):
        DR, m = self.D * self.R, self.top_m
        cid = np.zeros((DR, m), np.int32); fid = np.zeros((DR, m), np.int32)
        cv = np.zeros((DR, m), np.float32); fv = np.zeros((DR, m), np.float32)
        cs = np.zeros(DR, np.int32); fs = np.zeros(DR, np.int32)
        self._L.oracle_stream_get_buckets(self._h, cid, cv, fid, fv, cs, fs)
        return cid, cv, fid, fv, cs, fs

    def hash_search(self, Q, n_neighbors, verbose=False):
        """streamCEOs.hash_search (src/streamCEOs.cpp:621-877)."""
        Q = _f32(Q)
        nq = Q.shape[0]
        ids = np.zeros((nq, n_neighbors), np.int32)
        ips = np.zeros((nq, n_neighbors), np.float32)
        rc = self._L.oracle_stream_hash_search(self._h, Q, nq, n_neighbors, self.n_probed_vectors,
                                             self.n_probed_points, ids, ips)
        if rc:
            raise ValueError("invalid query parameters")
        return ids, ips

    def estimate_search(self, Q, n_neighbors, verbose=False):
        Q = _f32(Q)
        nq = Q.shape[0]
        ids = np.zeros((nq, n_neighbors), np.int32)
        ips = np.zeros((nq, n_neighbors), np.float32)
        rc = self._L.oracle_stream_search(self._h, Q, nq, n_neighbors, self.n_probed_vectors,
                                        self.n_probed_points, self.n_cand, ids, ips)
        if rc:
            raise ValueError("invalid query parameters")
        return ids, ips
