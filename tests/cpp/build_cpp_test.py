"""Compile tests/cpp/test_linbox_shim.cpp -- LinBox-shaped test code against
include/lbx/linbox_mirror.hpp and the GPU specialisations of
include/lbx/linbox_shim.hpp, linked with liblbx.so and the CPU oracle (its
checker) -- into tests/cpp/_build/test_linbox_shim (g++, no GPU needed to build)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "test_linbox_shim.cpp")
OUT = os.path.join(HERE, "_build", "test_linbox_shim")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-o", OUT, "-L", os.path.join(ROOT, "linbox_amd"), "-llbx", "-L", os.path.join(ROOT, "oracle", "_build"),
           "-llbx_oracle", "-Wl,-rpath,$ORIGIN/../../../linbox_amd", "-Wl,-rpath,$ORIGIN/../../../oracle/_build"]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
