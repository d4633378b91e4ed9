"""Compile tests/cpp/kernel_tests.cpp against libpgvd.so and the oracle.

    python tests/cpp/build_tests.py

The binary is written next to the source (git-ignored) so that GPU runs use
the copy built here.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "kernel_tests.cpp")
BIN = os.path.join(HERE, "kernel_tests")


def build() -> str:
    lib_dir = os.path.join(ROOT, "pgvd_amd")
    orc_dir = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{orc_dir}", SRC, "-o", BIN + ".tmp",
           f"-L{lib_dir}", "-lpgvd", f"-L{orc_dir}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../pgvd_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.run(cmd, check=True)
    os.replace(BIN + ".tmp", BIN)
    return BIN


if __name__ == "__main__":
    print(build())
    sys.exit(0)
