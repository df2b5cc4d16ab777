"""Builds libceos_amd.so (HIP kernels + C ABI) and the pybind11 module `CEOs` in-tree.

    python ceos_amd/build.py

Every HIP source is compiled for gfx950 only, with packed FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libceos_amd.so")
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
PYMOD = os.path.join(HERE, "CEOs" + EXT_SUFFIX)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["ceos_build.hip", "ceos_query.hip", "ceos_search.hip", "ceos_static.hip", "ceos_serve.cpp",
               "ceos_runtime.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(jobs: int = 4) -> str:
    objs = []
    procs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(CSRC, "_obj", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB])
    return LIB


def build_pymod() -> str:
    import pybind11

    inc = [pybind11.get_include(), sysconfig.get_paths()["include"]]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden"]
    for i in inc:
        cmd += ["-I", i]
    cmd += [os.path.join(CSRC, "ceos_pybind.cpp"), "-L", HERE, "-lceos_amd", "-Wl,-rpath,$ORIGIN", "-o", PYMOD]
    _run(cmd)
    return PYMOD


def build() -> None:
    build_lib()
    build_pymod()


if __name__ == "__main__":
    build()
    sys.exit(0)
