"""The C ABI: libpakhip.so loads without a GPU and exports every function
declared in include/pakhip.h; the Python mirror binds exactly those; the C++
adapter header compiles against it.  No compute calls (CPU only)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pakhip.h")
LIB = os.path.join(ROOT, "pakman_plus_amd", "_lib", "libpakhip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \t\*]*?\b(pk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_parses_to_known_entry_points():
    names = declared_functions()
    assert "pk_count" in names and "pk_builder_count" in names and "pk_macro_nodes" in names
    assert len(names) >= 14


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from pakman_plus_amd import build
        build.build_hip()
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared_functions()) <= exported


def test_python_mirror_binds_the_header():
    from pakman_plus_amd import api
    assert sorted(api.EXPORTS) == declared_functions()


def test_cpp_adapter_header_compiles():
    src = os.path.join(ROOT, "tests", "cpp", "builder_adapter_test.cpp")
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                           f"-I{os.path.join(ROOT, 'include')}", src])


def test_header_is_plain_c():
    # the ABI header must compile as C (no C++ types in the signatures)
    subprocess.check_call(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", HEADER])
