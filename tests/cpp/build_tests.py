"""Compile the C++ test programs next to their sources (git-ignored), so that
GPU runs use the copies built here.

    kernel_tests      tests/cpp/kernel_tests.cpp against libpgvd.so and the
                      oracle (GPU)
    brt_eval_check    the BRT node search (pgvd_amd/csrc/brt_eval.hpp)
                      compiled for the host against the oracle (CPU)

    python tests/cpp/build_tests.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "kernel_tests.cpp")
BIN = os.path.join(HERE, "kernel_tests")
CHECK_SRC = os.path.join(HERE, "brt_eval_check.cpp")
CHECK_BIN = os.path.join(HERE, "brt_eval_check")
HIPCC = "/opt/rocm/bin/hipcc"


def build() -> str:
    lib_dir = os.path.join(ROOT, "pgvd_amd")
    orc_dir = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{orc_dir}", SRC, "-o", BIN + ".tmp",
           f"-L{lib_dir}", "-lpgvd", f"-L{orc_dir}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../pgvd_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.run(cmd, check=True)
    os.replace(BIN + ".tmp", BIN)
    build_check()
    return BIN


def build_check() -> str:
    """Host-only HIP compile of the BRT node search check."""
    orc_dir = os.path.join(ROOT, "oracle")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O2",
           f"-I{os.path.join(ROOT, 'pgvd_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{orc_dir}", CHECK_SRC, "-o", CHECK_BIN + ".tmp", f"-L{orc_dir}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(CHECK_BIN + ".tmp", CHECK_BIN)
    return CHECK_BIN


if __name__ == "__main__":
    print(build())
    sys.exit(0)
