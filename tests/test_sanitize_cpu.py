"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle and the library's host-side generators, built with gcc/g++
-fsanitize=address,undefined (no recovery: any report aborts) from their
sources into a temporary directory, then run (tests/sanitize/check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-fopenmp"]


@pytest.mark.skipif(not shutil.which("gcc") or not shutil.which("g++"), reason="needs gcc and g++")
def test_host_code_under_sanitizers(tmp_path):
    obj_c = tmp_path / "oracle.o"
    obj_g = tmp_path / "gen.o"
    exe = tmp_path / "check"
    inc = [f"-I{ROOT}/include", f"-I{ROOT}/oracle"]
    steps = [
        ["gcc", *SAN, "-std=gnu11", *inc, "-c", f"{ROOT}/oracle/bingo_oracle.c", "-o", str(obj_c)],
        ["g++", *SAN, "-std=c++17", *inc, "-c", f"{ROOT}/bingo_amd/csrc/gen.cpp", "-o", str(obj_g)],
        ["g++", *SAN, "-std=c++17", *inc, f"{ROOT}/tests/sanitize/check.cpp", str(obj_c), str(obj_g), "-lm", "-o",
         str(exe)],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    env = dict(os.environ, OMP_NUM_THREADS="2", UBSAN_OPTIONS="print_stacktrace=1",
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized host checks ok" in r.stdout
