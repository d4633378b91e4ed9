"""The C++ host layer (include/pgvd_kernels.hpp): it compiles against the
installed headers, and on the GPU its stage tests -- the reference's
tests/KernelTests.cpp scenarios restated against the oracle -- pass."""
from __future__ import annotations

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "kernel_tests")


def test_cpp_layer_compiles():
    from tests.cpp.build_tests import build
    assert os.path.exists(build())


def test_cpp_layer_symbols_exported():
    import pgvd_amd
    out = subprocess.run(["nm", "-DC", "--defined-only", pgvd_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in ["pgvd::Kernels::BuildOctree_p", "pgvd::Kernels::RadixSortBigUnsigned",
                 "pgvd::Kernels::UniqueSorted", "pgvd::Kernels::BuildBinaryRadixTree_p",
                 "pgvd::Kernels::ComputeLocalSplits_p", "pgvd::Kernels::BinaryRadixToOctree_p",
                 "pgvd::Kernels::StreamScan_p", "pgvd::Kernels::PointsToMorton_p",
                 "pgvd::Kernels::BitPredicate", "pgvd::Kernels::DoubleCompact",
                 "pgvd::Karras::Quantize", "pgvd::Karras::BuildOctreeInParallel",
                 "pgvd::OctreeUtils::FindLeaf", "pgvd::OctreeUtils::FindLeaves",
                 "pgvd::OctreeUtils::CellWalk", "pgvd::OctreeUtils::CellWalks",
                 "pgvd::Geom::FitBoxes", "pgvd::Octree2::build"]:
        assert name in out, name


@pytest.mark.gpu
def test_cpp_kernel_scenarios():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/kernel_tests was not built (run __graft_entry__.build())")
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL PASSED" in p.stdout


SAN_BIN = os.path.join(HERE, "cpp", "san", "kernel_tests")


@pytest.mark.gpu
def test_cpp_kernel_scenarios_under_host_sanitizers():
    """The same scenarios against libpgvd.so built with its host code under
    AddressSanitizer and UndefinedBehaviorSanitizer (device code as usual;
    tests/cpp/build_tests.py --sanitized, run by __graft_entry__.build()).
    Leak checking is off: the HIP runtime keeps allocations to the end."""
    if not os.path.exists(SAN_BIN):
        pytest.fail("tests/cpp/san/kernel_tests was not built (run __graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([SAN_BIN], capture_output=True, text=True, timeout=280, env=env)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr[-4000:]
    assert "ALL PASSED" in p.stdout
