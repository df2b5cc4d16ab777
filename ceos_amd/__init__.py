"""ceos_amd — MI355X-native coCEOs (HIP kernels for gfx950 behind a C ABI).

The drop-in Python surface is the pybind11 module ``ceos_amd.CEOs`` with the
reference's classes ``CEOs`` and ``streamCEOs``::

    from ceos_amd import CEOs
    index = CEOs.CEOs(n, d)
    index.setIndexParam(D, n_repeats, top_m, n_threads, seed)
    index.build_coCEOs_Est(X)
    ids, ips = index.search_coCEOs_Est(Q, k)

Importing fails loudly when the native library has not been built
(``python ceos_amd/build.py``); there is no fallback path.
"""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libceos_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build the HIP library with `python ceos_amd/build.py`")

from . import CEOs  # noqa: E402  (pybind11 extension, links libceos_amd.so)

__all__ = ["CEOs", "LIB_PATH"]
