"""Compile the C++ test programs next to their sources (git-ignored), so that
GPU runs use the copies built here.

    kernel_tests      tests/cpp/kernel_tests.cpp against libpgvd.so and the
                      oracle (GPU)
    brt_eval_check    the BRT node search (pgvd_amd/csrc/brt_eval.hpp)
                      compiled for the host against the oracle (CPU)
    oracle_sanitize   tests/cpp/oracle_sanitize.c with the oracle, under
                      AddressSanitizer and UndefinedBehaviorSanitizer (CPU)
    san/              (--sanitized) libpgvd.so's sources with their host code
                      under AddressSanitizer and UndefinedBehaviorSanitizer
                      (device code as usual), and kernel_tests against that
                      library: san/kernel_tests (GPU); san/brt_eval_check,
                      brt_eval_check under both (CPU)

    python tests/cpp/build_tests.py [--sanitized]
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


# Sanitizers for host code only: each -fsanitize flag follows -Xarch_host.
HOST_SAN = ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
            "-Xarch_host -fno-sanitize-recover=all -Xarch_host -fno-omit-frame-pointer").split()
SAN_DIR = os.path.join(HERE, "san")
ORACLE_SAN_SRC = os.path.join(HERE, "oracle_sanitize.c")
ORACLE_SAN_BIN = os.path.join(HERE, "oracle_sanitize")


def build_oracle_sanitize() -> str:
    """The oracle and its driver under ASan and UBSan (gcc, host only)."""
    orc_dir = os.path.join(ROOT, "oracle")
    cmd = ["gcc", "-std=gnu11", "-g", "-O1", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", f"-I{orc_dir}", os.path.join(orc_dir, "pgvd_oracle.c"),
           ORACLE_SAN_SRC, "-lm", "-o", ORACLE_SAN_BIN + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(ORACLE_SAN_BIN + ".tmp", ORACLE_SAN_BIN)
    return ORACLE_SAN_BIN


def build_check_sanitized() -> str:
    """brt_eval_check with the BRT search's host build under ASan and UBSan."""
    orc_dir = os.path.join(ROOT, "oracle")
    os.makedirs(SAN_DIR, exist_ok=True)
    out = os.path.join(SAN_DIR, "brt_eval_check")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", *HOST_SAN,
           f"-I{os.path.join(ROOT, 'pgvd_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{orc_dir}", CHECK_SRC, "-o", out + ".tmp", f"-L{orc_dir}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../../oracle"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(out + ".tmp", out)
    return out


def build_sanitized() -> str:
    """libpgvd.so with its host code under ASan and UBSan, and kernel_tests
    linked against it with clang's sanitizer runtimes (the library's)."""
    sys.path.insert(0, ROOT)
    from pgvd_amd import build as B
    import concurrent.futures as cf
    os.makedirs(SAN_DIR, exist_ok=True)

    def comp(src):
        obj = os.path.join(SAN_DIR, os.path.basename(src)[:-4] + ".o")
        p = subprocess.run([B.HIPCC, *B.FLAGS, *HOST_SAN, "-g", "-c", src, "-o", obj],
                           capture_output=True, text=True)
        if p.returncode:
            raise RuntimeError(p.stderr[-3000:])
        return obj

    with cf.ThreadPoolExecutor(8) as ex:
        objs = list(ex.map(comp, B.sources()))
    lib = os.path.join(SAN_DIR, "libpgvd.so")
    subprocess.run([B.HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", *HOST_SAN,
                    "-shared-libsan", "-o", lib + ".tmp", *objs], check=True)
    os.replace(lib + ".tmp", lib)
    orc_dir = os.path.join(ROOT, "oracle")
    exe = os.path.join(SAN_DIR, "kernel_tests")
    clang = "/opt/rocm/llvm/bin/clang++"
    rt = os.path.dirname(subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"],
                                        check=True, capture_output=True,
                                        text=True).stdout.strip())
    subprocess.run([clang, "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-shared-libsan",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{orc_dir}", SRC, "-o", exe + ".tmp",
                    f"-L{SAN_DIR}", "-lpgvd", f"-L{orc_dir}", "-loracle",
                    "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/../../../oracle",
                    f"-Wl,-rpath,{rt}"], check=True)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    if "--sanitized" in sys.argv:
        print(build_sanitized())
    else:
        print(build())
        print(build_oracle_sanitize())
    sys.exit(0)
