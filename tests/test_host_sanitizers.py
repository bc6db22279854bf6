"""The host (CPU) code under AddressSanitizer and UndefinedBehaviorSanitizer: the PHAT reader
and writer, the host Rips builder, edge collapse and alpha generator of libssq, and the oracle, driven over
well-formed and malformed inputs by tests/asan/asan_driver.cpp."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "specseqplusplus_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
    inc = ["-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    objs = []
    for src in ("ssq_io.cpp", "ssq_gen.cpp", "ssq_collapse.cpp", "ssq_alpha.cpp"):
        o = str(tmp_path / (src + ".o"))
        subprocess.check_call(["g++", "-std=c++17", *san, *inc, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    o = str(tmp_path / "oracle.o")
    subprocess.check_call(["gcc", "-std=c11", *san, "-c", os.path.join(ROOT, "oracle", "ssq_oracle.c"), "-o", o])
    objs.append(o)
    exe = str(tmp_path / "asan_driver")
    subprocess.check_call(["g++", "-std=c++17", *san, *inc, os.path.join(ROOT, "tests", "asan", "asan_driver.cpp"),
                           *objs, "-o", exe, "-lpthread"])
    work = tmp_path / "files"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(work)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr, r.stderr
    assert "0 failed checks" in r.stdout
