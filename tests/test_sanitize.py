"""The library's host-side file readers (.hmsh, .hsln and gmsh .msh) and writers,
built with the AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code
only: these sources use no HIP) and driven over valid files and seeded
corruptions of them (tests/sanitize/parsers_fuzz.cpp: truncations, byte flips,
count fields set to extreme values, text edits): every call returns a status
code, with no out-of-bounds access, leak or undefined behaviour on the way."""
import os
import shutil
import subprocess

import pytest

from test_host import write_square_msh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cfd_optimisations_amd", "csrc")
SOURCES = [os.path.join(CSRC, f) for f in ("errors.cpp", "mesh_host.cpp", "msh_import.cpp", "output_host.cpp")]
DRIVER = os.path.join(ROOT, "tests", "sanitize", "parsers_fuzz.cpp")


def test_file_readers_under_sanitizers(tmp_path, wedge_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path / "parsers_fuzz")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "include"), "-I", CSRC] + SOURCES + [DRIVER, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    msh = str(tmp_path / "square.msh")
    write_square_msh(msh, 9, 7, quads_left=3)
    hsln = os.path.join(ROOT, "tests", "golden", "Base_1.hsln")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path), wedge_path, hsln, msh], capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "accepted" in r.stdout
