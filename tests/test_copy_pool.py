"""The host copy threads behind the pinned-buffer copies of host operands
(linbox_amd/csrc/copy_pool.h), tested on the CPU: tests/cpp/copy_pool_harness.cpp
built with ThreadSanitizer and with AddressSanitizer + UBSan, each run to
completion (every byte copied, no write outside the destination, no data race)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "copy_pool_harness.cpp")
INC = os.path.join(ROOT, "linbox_amd", "csrc")
OUT = os.path.join(HERE, "cpp", "_build")


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_copy_pool_under_sanitizer(san):
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "copy_pool_" + san.replace(",", "_"))
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=" + san, "-fno-omit-frame-pointer",
                    "-I", INC, "-o", exe, SRC], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failures" in r.stdout
