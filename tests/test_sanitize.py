"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, where it
runs without a GPU: the oracle's C restatement driven through every entry
point (tests/cpp/oracle_sanitize.c), and the BRT node search of the
product's kernels (pgvd_amd/csrc/brt_eval.hpp) compiled for the host.  A
sanitizer report aborts the program."""
from __future__ import annotations

import os
import subprocess

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_oracle_under_asan_and_ubsan():
    from tests.cpp.build_tests import build_oracle_sanitize
    binary = build_oracle_sanitize()
    p = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=ENV)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stdout + p.stderr[-4000:]


def test_brt_search_under_asan_and_ubsan():
    import oracle
    from tests.cpp.build_tests import build_check_sanitized
    oracle.compile_libs(ref=False)
    binary = build_check_sanitized()
    p = subprocess.run([binary], capture_output=True, text=True, timeout=600, env=ENV)
    assert p.returncode == 0 and "ALL MATCH" in p.stdout, p.stdout[-3000:] + p.stderr[-4000:]
