"""In-tree build of the native libraries (no JIT cache, no pip install).

  _lib/libpakhip.so  HIP kernels + C ABI (include/pakhip.h), gfx950 only
  _lib/libpkgen.so   host generator of synthetic read sets
  tests/cpp/_bin/builder_adapter_test  host driver of include/pakhip_builders.hpp

Run `python -m pakman_plus_amd.build` or call build_all().
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Compile flags for device code.  Packed-FP32 instructions are switched off
# for all device code of this library.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]

HIP_SOURCES = ["pk_device.hip", "pk_server.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_gen() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libpkgen.so")
    tmp = f"{out}.{os.getpid()}.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "pk_gen.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libpakhip.so")
    tmp = f"{out}.{os.getpid()}.tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_adapter_test() -> str:
    """Host-only driver of the C++ builder adapter (include/pakhip_builders.hpp),
    linked against _lib/libpakhip.so; used by the GPU tests."""
    src = os.path.join(ROOT, "tests", "cpp", "builder_adapter_test.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "_bin")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "builder_adapter_test")
    tmp = f"{out}.{os.getpid()}.tmp"  # concurrent builders each rename their own
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", src, f"-L{LIBDIR}", "-lpakhip",
          "-Wl,-rpath,$ORIGIN/../../../pakman_plus_amd/_lib", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_prefix_order_test() -> str:
    """Host-only driver of pakhip::reference_prefix_order (CPU tests)."""
    src = os.path.join(ROOT, "tests", "cpp", "prefix_order_test.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "_bin")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "prefix_order_test")
    tmp = f"{out}.{os.getpid()}.tmp"
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_gen()
    build_hip()
    build_adapter_test()
    build_prefix_order_test()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gen": build_gen, "hip": build_hip, "adapter": build_adapter_test,
         "prefix": build_prefix_order_test}[w]()
