"""CPU-side checks of the product boundary: the HIP library builds, loads and
exports every symbol include/virgo.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "virgo_amd", "libvirgo_hip.so")
HDR = os.path.join(ROOT, "include", "virgo.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "vg_prove" in names and "vg_ctx_create" in names and "vg_fft" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libvirgo_hip.so missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_export_list_matches_header():
    import virgo_amd
    assert sorted(virgo_amd.EXPORTS) == declared()


def test_cpp_mirror_header_compiles():
    """include/virgo_prover.hpp (the zk_prover mirror) compiles standalone."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = "#include \"virgo_prover.hpp\"\nint main() { return 0; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(root, "include"),
                        "-x", "c++", "-"], input=src.encode(), capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
