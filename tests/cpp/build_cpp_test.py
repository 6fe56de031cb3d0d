"""Compile tests/cpp/test_polymat.cpp against include/lbx/polymat.hpp and liblbx.so
into tests/cpp/_build/test_polymat (g++, no GPU needed to build)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "test_polymat.cpp")
OUT = os.path.join(HERE, "_build", "test_polymat")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-o", OUT,
           "-L", os.path.join(ROOT, "linbox_amd"), "-llbx", "-Wl,-rpath,$ORIGIN/../../../linbox_amd"]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
