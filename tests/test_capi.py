"""CPU checks of the C ABI and the drop-in module: the library loads, exports every
symbol include/ceos_amd.h declares, and the host-side logic (defaults, parameter
rules, error mapping) mirrors the reference without touching a GPU."""
import ctypes as C

import numpy as np
import pytest

from ceos_amd import capi


def test_header_symbols_exported():
    names = capi.header_symbols()
    assert len(names) >= 35
    L = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_codes_without_device():
    L = capi.lib()
    h = C.c_void_p()
    assert L.ceos_create(100, 16, C.byref(h)) == 0
    try:
        assert L.ceos_set_index_param(h, 0, 1, 10, -1, 1) == 1          # n_proj must be positive
        assert L.ceos_set_index_param(h, 8192, 1, 10, -1, 1) == 1       # fhtDim above 4096
        assert b"4096" in L.ceos_last_error()
        assert L.ceos_set_index_param(h, 64, 2, 10, -1, 42) == 0
        q = np.zeros((2, 16), np.float32)
        ids = np.zeros((2, 5), np.int32)
        ips = np.zeros((2, 5), np.float32)
        rc = L.ceos_search_coceos(h, q.ctypes.data, 2, 16, 5, ids.ctypes.data, ips.ctypes.data)
        assert rc == 3 and b"not built" in L.ceos_last_error()           # search before build
        assert L.ceos_search_coceos(h, q.ctypes.data, 2, 15, 5, ids.ctypes.data, ips.ctypes.data) == 1
    finally:
        L.ceos_destroy(h)
    assert L.ceos_create(10, 0, C.byref(h)) == 1


def test_module_defaults_mirror_reference(CEOs):
    # src/CEOs.h:52-55 defaults; src/CEOs.h:69 n_probed_points := top_m
    g = CEOs.CEOs(1000, 32)
    assert (g.n_probed_vectors, g.n_probed_points, g.n_cand, g.centering) == (10, 10, 10, False)
    g.setIndexParam(256, 2, 50, 4, 42)
    assert g.n_probed_points == 50
    assert g.fhtDim == 256
    g.n_probed_vectors, g.n_cand, g.centering = 7, 99, True
    assert (g.n_probed_vectors, g.n_cand, g.centering) == (7, 99, True)
    # streamCEOs: setIndexParam sets n_probed_vectors := top_points (src/streamCEOs.h:58)
    s = CEOs.streamCEOs(32)
    assert (s.n_probed_vectors, s.n_probed_points, s.n_cand) == (10, 10, 10)
    s.setIndexParam(128, 1, 33, -1, 1)
    assert s.n_probed_vectors == 33 and s.n_probed_points == 10


def test_module_errors_raise_instead_of_exit(CEOs):
    g = CEOs.CEOs(100, 8)
    with pytest.raises(ValueError):
        g.setIndexParam(0)
    g.setIndexParam(64, 1, 10, -1, 1)
    with pytest.raises(RuntimeError):
        g.search_coCEOs_Est(np.zeros((3, 8), np.float32), 5)             # not built
    with pytest.raises(ValueError):
        g.build_coCEOs_Est(np.zeros(8, np.float32))                      # not 2-D
    with pytest.raises((ValueError, RuntimeError)):  # top_m > n_points (or no device here)
        g.build_CEOs_Hash(np.zeros((3, 8), np.float32))
    s = CEOs.streamCEOs(8)
    with pytest.raises(RuntimeError):                                     # not built
        s.hash_search(np.zeros((3, 8), np.float32), 5)


def test_read_bandwidth_argument_checks():
    """The calibration entry point rejects bad arguments before touching a device."""
    L = capi.lib()
    g = C.c_double()
    buf = np.zeros(1024, np.float32)
    p = buf.ctypes.data
    assert L.ceos_read_bandwidth(None, 1024, p, 256, 1, None, C.byref(g)) == 1      # null array
    assert L.ceos_read_bandwidth(p, 1022, p, 256, 1, None, C.byref(g)) == 1         # n not a multiple of 4
    assert L.ceos_read_bandwidth(p, 1024, p, 100, 1, None, C.byref(g)) == 1         # n_out not a multiple of 256
    assert L.ceos_read_bandwidth(p, 1024, p, 256, 0, None, C.byref(g)) == 1         # iters
    assert b"iters" in L.ceos_last_error()


def test_rerank_bounds_toggle_without_device(CEOs):
    """The rerank-bounds switch is host state until an index is built: on by default,
    reported off for n_features % 4 != 0 (no int8 copies then), null handles rejected."""
    L = capi.lib()
    L.ceos_set_rerank_bounds.argtypes = [C.c_void_p, C.c_int]
    L.ceos_get_rerank_bounds.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    h = C.c_void_p()
    v = C.c_int(-1)
    assert L.ceos_create(100, 16, C.byref(h)) == 0
    try:
        assert L.ceos_get_rerank_bounds(h, C.byref(v)) == 0 and v.value == 1
        assert L.ceos_set_rerank_bounds(h, 0) == 0
        assert L.ceos_get_rerank_bounds(h, C.byref(v)) == 0 and v.value == 0
        assert L.ceos_set_rerank_bounds(h, 1) == 0
        assert L.ceos_get_rerank_bounds(h, C.byref(v)) == 0 and v.value == 1
        assert L.ceos_get_rerank_bounds(h, None) == 1
    finally:
        L.ceos_destroy(h)
    assert L.ceos_set_rerank_bounds(None, 1) == 1
    g = CEOs.CEOs(100, 30)
    assert g.rerank_bounds is False          # 30 % 4 != 0
    s = CEOs.streamCEOs(32)
    assert s.rerank_bounds is True
    s.rerank_bounds = False
    assert s.rerank_bounds is False


def _index_file(path, n=20, d=8, D=64, R=1, m=10, F=64, L=6, ids=None, sizes=None, truncate=0):
    """Bytes of a coCEOs index file (the layout Core::save writes)."""
    rng = np.random.default_rng(5)
    nl = 2 * D * R
    ent = np.zeros((nl, m, 2), np.uint32)
    ent[:, :, 0] = rng.integers(0, max(n, 1), (nl, m)) if ids is None else ids
    ent[:, :, 1] = rng.standard_normal((nl, m)).astype(np.float32).view(np.uint32)
    sz = np.full(nl, m, np.int32) if sizes is None else np.asarray(sizes, np.int32)
    blob = b"CEOSIDX1" + np.array([0, d, D, R, m, F, L, 10, 10, 10, 0], np.int32).tobytes()
    blob += np.uint32(42).tobytes() + np.int64(n).tobytes()
    blob += rng.standard_normal((max(n, 0), d)).astype(np.float32).tobytes()
    blob += ent.tobytes() + sz.tobytes() + np.zeros((max(n, 0) + 31) // 32, np.uint32).tobytes()
    with open(path, "wb") as f:
        f.write(blob[: len(blob) - truncate])


@pytest.mark.parametrize("case", ["id_out_of_range", "size_above_top_m", "top_m_above_n", "wrong_fht_dim",
                                  "truncated", "negative_n"])
def test_load_rejects_damaged_files_before_device_work(tmp_path, case):
    """Every field of an index file is checked before anything reaches the device, so a
    damaged or foreign file fails with CEOS_EINVAL on a host without a GPU."""
    path = str(tmp_path / "bad.ceos")
    if case == "id_out_of_range":
        ids = np.zeros((128, 10), np.uint32)
        ids[77, 3] = 20  # n = 20
        _index_file(path, ids=ids)
    elif case == "size_above_top_m":
        sizes = np.full(128, 10)
        sizes[5] = 11
        _index_file(path, sizes=sizes)
    elif case == "top_m_above_n":
        _index_file(path, n=8, m=10)
    elif case == "wrong_fht_dim":
        _index_file(path, F=128, L=7)
    elif case == "truncated":
        _index_file(path, truncate=9)
    else:
        _index_file(path, n=-1)
    L = capi.lib()
    L.ceos_load.argtypes = [C.c_void_p, C.c_char_p]
    h = C.c_void_p()
    assert L.ceos_create(1, 8, C.byref(h)) == 0
    try:
        assert L.ceos_load(h, path.encode()) == 1, L.ceos_last_error()
        assert b"index file" in L.ceos_last_error()
    finally:
        L.ceos_destroy(h)
