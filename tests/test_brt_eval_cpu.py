"""The BRT node search every BRT kernel runs (pgvd_amd/csrc/brt_eval.hpp),
compiled for the host and checked node by node against the oracle's
BuildBinaryRadixTree (C/BuildBRT.c:90-181) for 32-, 63- and 126-bit keys,
uniform and clustered, including the LDS-window form the kernels use."""
from __future__ import annotations

import os
import subprocess


def test_brt_eval_matches_oracle_on_cpu():
    from tests.cpp.build_tests import build_check
    import oracle
    oracle.compile_libs(ref=False)
    binary = build_check()
    p = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL MATCH" in p.stdout
