"""CPU checks of host logic of the library that needs no GPU: the proof
server's range allocator against a brute-force model (a C++ program
compiled here from tests/cpp/range_alloc_test.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_alloc_matches_model(tmp_path):
    exe = str(tmp_path / "range_alloc_test")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "range_alloc_test.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
